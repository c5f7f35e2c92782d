/*
 * cvp9_hip/abi_types.h — the reference decoder's data types at the
 * reconstruction-backend boundary, restated field for field so this library
 * is binary-compatible with the structs the reference decoder passes in.
 *
 * Layouts must stay identical to the reference (gcc, x86-64 SysV):
 *   cvp9_err_t, cvp9_frame_info_t, cvp9_dmabuf_frame_t, cvp9_export_buffer_t
 *       include/compute_vp9/decoder.h:29-37, 47-56, 126-134, 152-158
 *   vp9_frame_header_t      src/decoder/vp9_bitstream.h:11-87   (76 bytes)
 *   cvp9_mv_t               src/decoder/vp9_parsed_frame.h:12-17 (16 bytes)
 *   vp9_macroblock_info_t   src/decoder/vp9_parsed_frame.h:20-33 (32 bytes)
 *   vp9_parsed_frame_t      src/decoder/vp9_parsed_frame.h:36-63
 * The offsets are pinned by static assertions below and, in the test suite,
 * by a probe compiled against the reference headers.
 *
 * When this library is built inside the reference tree, define
 * CVP9_HIP_USE_REFERENCE_HEADERS so the reference's own headers supply these
 * types instead (hip_backend.h includes them directly).
 */
#ifndef CVP9_HIP_ABI_TYPES_H
#define CVP9_HIP_ABI_TYPES_H

#include <stdint.h>
#include <stddef.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes, values as in the reference (decoder.h:29-37). */
typedef enum {
    CVP9_OK               =  0,
    CVP9_ERR_NOMEM        = -1,
    CVP9_ERR_NO_BACKEND   = -2,
    CVP9_ERR_INVALID_DATA = -3,
    CVP9_ERR_UNSUPPORTED  = -4,
    CVP9_ERR_GPU          = -5,
    CVP9_ERR_AGAIN        = -6,
} cvp9_err_t;

/* Backend enum (decoder.h:21-26) plus the HIP backend as a new value; the
 * existing values keep their numbers. */
typedef enum {
    CVP9_BACKEND_AUTO    = 0,
    CVP9_BACKEND_VULKAN  = 1,
    CVP9_BACKEND_OPENCL  = 2,
    CVP9_BACKEND_CPU     = 3,
    CVP9_BACKEND_HIP     = 4,
} cvp9_backend_t;

/* Planar I420 frame description (decoder.h:47-56). */
typedef struct {
    uint32_t    width;
    uint32_t    height;
    uint32_t    stride_y;
    uint32_t    stride_uv;
    uint8_t    *plane_y;
    uint8_t    *plane_u;
    uint8_t    *plane_v;
    int64_t     pts;
} cvp9_frame_info_t;

/* Decoded frame exported as one DMA-BUF holding I420 (decoder.h:126-134). */
typedef struct {
    int         fd;
    uint64_t    size;
    uint32_t    width;
    uint32_t    height;
    uint32_t    offsets[3];
    uint32_t    pitches[3];
    int64_t     pts;
} cvp9_dmabuf_frame_t;

/* CPU-mapped, DMA-BUF exportable device buffer (decoder.h:152-158). */
typedef struct {
    int         fd;
    void       *mapped;
    uint64_t    size;
    void       *priv_buf;
    void       *priv_mem;
} cvp9_export_buffer_t;

/* Frame type (vp9_bitstream.h:11-16). */
typedef enum {
    VP9_FRAME_KEY        = 0,
    VP9_FRAME_NON_KEY    = 1,
    VP9_FRAME_INTRA_ONLY = 2,
    VP9_FRAME_S          = 3,
} vp9_frame_type_t;

typedef enum {
    VP9_CS_UNKNOWN   = 0,
    VP9_CS_BT601     = 1,
    VP9_CS_BT709     = 2,
    VP9_CS_SMPTE170  = 3,
    VP9_CS_SMPTE240  = 4,
    VP9_CS_BT2020    = 5,
    VP9_CS_SRGB      = 7,
} vp9_color_space_t;

typedef enum {
    VP9_FILTER_EIGHTTAP        = 0,
    VP9_FILTER_EIGHTTAP_SMOOTH = 1,
    VP9_FILTER_EIGHTTAP_SHARP  = 2,
    VP9_FILTER_BILINEAR        = 3,
    VP9_FILTER_SWITCHABLE      = 4,
} vp9_interp_filter_t;

/* Uncompressed frame header (vp9_bitstream.h:39-87). The reconstruction
 * reads only width, height, frame_type, filter_level and sharpness_level. */
typedef struct {
    vp9_frame_type_t    frame_type;
    uint32_t            width;
    uint32_t            height;
    uint32_t            render_width;
    uint32_t            render_height;
    uint8_t             profile;
    uint8_t             bit_depth;
    vp9_color_space_t   color_space;
    bool                color_range;
    uint8_t             subsampling_x;
    uint8_t             subsampling_y;
    bool                show_frame;
    bool                show_existing_frame;
    uint8_t             frame_to_show_map_idx;
    bool                intra_only;
    uint8_t             reset_frame_context;
    bool                error_resilient;
    bool                allow_high_precision_mv;
    vp9_interp_filter_t interp_filter;
    bool                refresh_frame_context;
    bool                frame_parallel;
    uint8_t             frame_context_idx;
    uint8_t             log2_tile_cols;
    uint8_t             log2_tile_rows;
    uint32_t            uncompressed_header_bytes;
    uint16_t            first_partition_size;
    int16_t             base_qindex;
    int8_t              y_dc_delta_q;
    int8_t              uv_dc_delta_q;
    int8_t              uv_ac_delta_q;
    uint8_t             filter_level;
    uint8_t             sharpness_level;
    uint8_t             refresh_frame_flags;
    uint8_t             ref_frame_idx[3];
    uint8_t             ref_frame_sign_bias[4];
    bool                segmentation_enabled;
} vp9_frame_header_t;

/* Per-4x4 motion vector cell (vp9_parsed_frame.h:12-17). */
typedef struct {
    int32_t x;
    int32_t y;
    int32_t ref;
    int32_t pad;
} cvp9_mv_t;

/* One prediction block as emitted by the tile parser
 * (vp9_parsed_frame.h:20-33). mv[r][0] is x, mv[r][1] is y, in 1/8 pel. */
typedef struct {
    uint16_t x;
    uint16_t y;
    uint8_t  width;
    uint8_t  height;
    uint8_t  tx_size;
    uint8_t  is_intra;
    uint8_t  y_mode;
    uint8_t  uv_mode;
    uint8_t  ref_frame[2];
    int16_t  mv[2][2];
    uint8_t  skip;
    uint32_t coeff_offset;
    uint16_t eob;
} vp9_macroblock_info_t;

/* Parsed frame handed to the backend (vp9_parsed_frame.h:36-63). The
 * backend reads hdr, blocks[0..num_blocks) and coeffs[0..num_coeffs). */
typedef struct {
    vp9_frame_header_t      hdr;
    vp9_macroblock_info_t  *blocks;
    uint32_t                num_blocks;
    uint32_t                blocks_capacity;
    int16_t                *coeffs;
    uint32_t                num_coeffs;
    uint32_t                coeffs_capacity;
    cvp9_mv_t              *mv_grid;
    uint32_t                mv_grid_width;
    uint32_t                mv_grid_height;
    uint8_t                *mi_width_grid;
    uint8_t                *mi_height_grid;
    uint32_t                mi_grid_width;
    uint32_t                mi_grid_height;
    uint32_t               *mi_block_grid;
} vp9_parsed_frame_t;

#ifdef __cplusplus
}
#define CVP9_ABI_ASSERT(c, m) static_assert(c, m)
#else
#define CVP9_ABI_ASSERT(c, m) _Static_assert(c, m)
#endif

CVP9_ABI_ASSERT(sizeof(vp9_macroblock_info_t) == 32, "block record is 32 bytes");
CVP9_ABI_ASSERT(offsetof(vp9_macroblock_info_t, mv) == 12, "mv at 12");
CVP9_ABI_ASSERT(offsetof(vp9_macroblock_info_t, skip) == 20, "skip at 20");
CVP9_ABI_ASSERT(offsetof(vp9_macroblock_info_t, coeff_offset) == 24, "coeff_offset at 24");
CVP9_ABI_ASSERT(offsetof(vp9_macroblock_info_t, eob) == 28, "eob at 28");
CVP9_ABI_ASSERT(sizeof(vp9_frame_header_t) == 76, "frame header is 76 bytes");
CVP9_ABI_ASSERT(offsetof(vp9_frame_header_t, filter_level) == 63, "filter_level at 63");
CVP9_ABI_ASSERT(offsetof(vp9_frame_header_t, sharpness_level) == 64, "sharpness_level at 64");
CVP9_ABI_ASSERT(sizeof(vp9_parsed_frame_t) == 160, "parsed frame is 160 bytes");
CVP9_ABI_ASSERT(offsetof(vp9_parsed_frame_t, blocks) == 80, "blocks at 80");
CVP9_ABI_ASSERT(offsetof(vp9_parsed_frame_t, coeffs) == 96, "coeffs at 96");
CVP9_ABI_ASSERT(sizeof(cvp9_mv_t) == 16, "mv cell is 16 bytes");
CVP9_ABI_ASSERT(sizeof(cvp9_frame_info_t) == 48, "frame info is 48 bytes");

#endif /* CVP9_HIP_ABI_TYPES_H */
