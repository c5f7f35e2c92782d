/*
 * cvp9_oracle.h — CPU restatement of the reference's reconstruction path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the parity oracle: tests/, the smoke
 * check and the benchmark's CPU-baseline leg may call it, as the checker.
 * The product library never links, loads or calls it.
 *
 * Every function restates the reference's own C (file:line cited at each
 * definition). It keeps the reference's arithmetic exactly: FP64 matrix
 * inverse DCT with a per-call cos table and strictly ascending summation,
 * bilinear 1/8-pel motion compensation, whole-block intra prediction in list
 * order, and the 4-tap luma deblock. Build it with -ffp-contract=off.
 *
 * Pinning (what makes this oracle trustworthy, tested in tests/):
 *   - golden per-plane hashes and full small frames produced by the
 *     reference's own C compiled from its sources (oracle/_ref);
 *   - closed forms: DC-only IDCT, orthonormality, constant-neighbour intra,
 *     zero/integer-MV motion compensation, flat-image loop filter, threshold
 *     table, NV12 interleave;
 *   - brute force: a naive double-loop IDCT for N=4.
 */
#ifndef CVP9_ORACLE_H
#define CVP9_ORACLE_H

#include "cvp9_hip/abi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Frame buffers (vp9_frame.c:8-63). */
int  orc_frame_alloc(cvp9_frame_info_t *f, uint32_t width, uint32_t height);
void orc_frame_free(cvp9_frame_info_t *f);
int  orc_frame_copy(cvp9_frame_info_t *dst, const cvp9_frame_info_t *src);

/* IDCT matrix exactly as the reference builds it (vp9_reconstruct.c:19-28):
 * M[i*N + j] for N in {4,8,16,32}. */
void orc_idct_matrix(int n, double *m_out);
/* 2D inverse DCT of one N x N raster block (vp9_reconstruct.c:13-60). */
void orc_idct_2d(const int16_t *in, int16_t *out, int n);

/* Single stages, exposed for stage-level tests. */
void orc_intra_predict(uint8_t *dst, int stride, int w, int h,
                       const uint8_t *above, const uint8_t *left,
                       uint8_t top_left, int mode);
void orc_motion_compensate(uint8_t *dst, int dst_stride,
                           const uint8_t *ref, int ref_stride,
                           int w, int h, int block_x, int block_y,
                           int mv_x, int mv_y, int ref_w, int ref_h);
void orc_loop_filter(cvp9_frame_info_t *f, uint8_t level, uint8_t sharpness);
int  orc_lf_threshold(uint8_t level, uint8_t sharpness);

/* Whole-frame reconstruction (vp9_reconstruct.c:194-370). With
 * apply_lf == 0 the frame is returned before the deblocking pass. */
int orc_reconstruct_frame_ex(const vp9_parsed_frame_t *pf,
                             const cvp9_frame_info_t *refs[8],
                             cvp9_frame_info_t *out, int apply_lf);
int orc_reconstruct_frame(const vp9_parsed_frame_t *pf,
                          const cvp9_frame_info_t *refs[8],
                          cvp9_frame_info_t *out);

/* Decoder-level state of the CPU fallback: the 8-slot reference pool
 * (vp9_decoder.c:209-231). */
typedef struct orc_decoder orc_decoder_t;
orc_decoder_t *orc_decoder_new(void);
void orc_decoder_free(orc_decoder_t *d);
/* Reconstructs pf against the pool, updates the pool, and copies the frame
 * into *out (caller frees with orc_frame_free). Returns 0 on success. */
int orc_decoder_frame(orc_decoder_t *d, const vp9_parsed_frame_t *pf,
                      int64_t pts, cvp9_frame_info_t *out);
/* Reconstructs a run of frames; writes each frame's I420 bytes
 * (W*H + 2*cw*ch) consecutively into out (NULL to discard). Returns frames
 * done. Used by the CPU-baseline timing. */
int orc_decoder_run(orc_decoder_t *d, const vp9_parsed_frame_t *const *pfs,
                    int n, uint8_t *out);

/* I420 -> NV12 (vaapi_entrypoint.c:131-153): Y with pitch W, then UV
 * interleaved with pitch 2*cw. */
void orc_nv12_pack(const cvp9_frame_info_t *f, uint8_t *dst);

/* Packed I420 copy of a frame (Y W*H, U cw*ch, V cw*ch), for comparisons. */
void orc_frame_pack_i420(const cvp9_frame_info_t *f, uint8_t *dst);

#ifdef __cplusplus
}
#endif
#endif
