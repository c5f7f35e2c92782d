/*
 * cvp9_hip/synth.h — deterministic generator of already-parsed VP9 frames.
 *
 * No counterpart in the reference: its generator (tests/vp9_synth.h) writes
 * bitstream packets whose tiles decode to all-DC 64x64 blocks. This one
 * writes vp9_parsed_frame_t directly, i.e. what the tile parser
 * (src/decoder/vp9_tile.c, vp9_kf_tile.c:967-1232) hands to the
 * reconstruction backend, so workloads exercise every reconstruction stage.
 *
 * Layout of the output, identical to the parser's:
 *   - blocks in Z order inside each 64x64 superblock, superblocks in raster
 *     order; blocks wholly outside the frame are omitted, partial ones kept;
 *   - coefficients are int16, raster order per transform, transforms in
 *     raster order per block, stored densely from coeff_offset for the
 *     floor(clamped_w/T) x floor(clamped_h/T) grid (nothing for skip blocks);
 *   - nonzero coefficients sit at the lowest positions of the zigzag scan of
 *     src/decoder/vp9_entropy.c:18-38.
 * All randomness comes from a counter-based splitmix64 stream and integer or
 * IEEE double arithmetic without library calls, so the bytes are identical
 * on every machine.
 */
#ifndef CVP9_HIP_SYNTH_H
#define CVP9_HIP_SYNTH_H

#include "cvp9_hip/abi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t width, height;
    uint64_t seed;
    uint32_t split_permille;     /* P(split) per 64/32/16 node, x1000 */
    uint32_t intra_permille;     /* P(intra) per block on non-key frames */
    uint32_t skip_permille;      /* P(skip) per block */
    uint32_t tx32_permille;      /* 0: TX uniform over allowed sizes;
                                    else P(32x32 where allowed), rest uniform
                                    over 4/8/16 */
    uint32_t coef_fixed;         /* >0: this many leading zigzag positions */
    uint32_t coef_permille;      /* else: this fraction of N*N positions */
    uint32_t lap_scale;          /* Laplacian scale of coefficient values */
    int32_t  mv_range;           /* MV components uniform in [-r, r], 1/8 pel */
    int32_t  filter_level;       /* 0..63, or -1 for U[0,63] per frame */
    uint32_t sharpness;          /* 0..7 */
    uint32_t key_interval;       /* frame k is KEY iff k % key_interval == 0;
                                    0 = only frame 0 */
    int32_t  qindex_lo, qindex_hi; /* base_qindex drawn per frame (no effect
                                      on reconstruction) */
} cvp9_synth_params_t;

/* Fills p with preset 0..4 (the five benchmark workloads; see the
 * generator source) at the given size and seed. Returns 0, or -1 for an
 * unknown preset. */
int cvp9_synth_preset(cvp9_synth_params_t *p, int preset,
                      uint32_t width, uint32_t height, uint64_t seed);

/* Generates frame `index` of the stream described by p. The result and its
 * arrays are owned by the caller: release with cvp9_synth_free. */
vp9_parsed_frame_t *cvp9_synth_frame(const cvp9_synth_params_t *p, uint32_t index);
void cvp9_synth_free(vp9_parsed_frame_t *pf);

/* Algorithmic bytes of one frame (reads of coefficients, block records and
 * reference footprints, I420 write; +NV12 when nv12 != 0). */
uint64_t cvp9_synth_algo_bytes(const vp9_parsed_frame_t *pf, int nv12);

#ifdef __cplusplus
}
#endif
#endif
