/*
 * cvp9_hip/testing.h — test and diagnostics hooks of the HIP backend.
 *
 * No counterpart in the reference. These let the test suite pin pieces of
 * the backend on their own: the IDCT tables the kernels use, the host-side
 * validation that decides between this backend and the CPU path, and
 * stage-level parity (deblocking switched off to compare with the CPU
 * reconstruction before its deblocking pass).
 */
#ifndef CVP9_HIP_TESTING_H
#define CVP9_HIP_TESTING_H

#include "cvp9_hip/abi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Library version string. */
const char *cvp9_hip_version(void);

/* Number of HIP devices visible (0 without a GPU; never initialises one). */
int cvp9_hip_device_count(void);

/* Host copy of the N x N inverse-DCT matrix uploaded to the kernels
 * (row-major, M[i][j]); N in {4, 8, 16, 32}. Returns 0, or -1 for other N. */
int cvp9_hip_idct_table(int n, double *out);

/* Runs the backend's validation of pf without a device: CVP9_OK when the
 * HIP path takes the frame, else the code hip_decode_frame would return. */
cvp9_err_t cvp9_hip_plan_check(const vp9_parsed_frame_t *pf);

/* The device image of pf as the planner lays it out (blocks, work tables,
 * intra order, coefficients), planned on at most `threads` host threads
 * (0: as many as the backend uses; frames below 2048 blocks are planned on
 * the calling thread). Writes it to dst when cap is large enough and
 * returns its size in bytes, or a negative cvp9_err_t. The image never
 * depends on the number of threads. Needs no device. */
int64_t cvp9_hip_plan_image(const vp9_parsed_frame_t *pf, int threads, uint8_t *dst, uint64_t cap);
/* The host part of pf's device image with compact coefficients, as the
 * drop-in boundary stages it: the tables as above, then per chunk of 64
 * pool entries a 64-bit nonzero mask (at info[0]) and the position of its
 * first value among the packed nonzero values (u32, at info[1]), then the
 * values (int16, at info[2]); info[3] the number of values, info[4] of
 * chunks, info[5] the bytes that cross to the device, info[6] the offset of
 * the dense pool the device writes from them and info[7] the whole image's
 * size. Writes dst (when cap is large enough) and returns the host part's
 * size, or a negative cvp9_err_t. Needs no device. */
int64_t cvp9_hip_plan_image_compact(const vp9_parsed_frame_t *pf, int threads, uint8_t *dst,
                                    uint64_t cap, uint64_t info[8]);
/* The order in which intra workers take pf's intra blocks: indices into
 * pf->blocks, by dependency level (no earlier intra block among the
 * above/left/above-left neighbours: level 0), then list index. Returns the
 * number of intra blocks written to order (at most cap), or a negative
 * cvp9_err_t. Needs no device. */
int cvp9_hip_intra_order(const vp9_parsed_frame_t *pf, uint32_t *order, uint32_t cap);

/* The chains of intra blocks a prepared frame lists for chain workers:
 * returns the number of chains K (or a negative cvp9_err_t) and with out !=
 * NULL writes K + 1 offsets, then the chained entries (positions in the
 * order above, chain by chain, each chain in dependency order), then every
 * other position in list order: K + 1 + (number of intra blocks) words, at
 * most cap. Needs no device. */
int cvp9_hip_intra_chains(const vp9_parsed_frame_t *pf, uint32_t *out, uint32_t cap);

/* Dependency levels of pf's intra blocks (the order above groups them by
 * level): returns the number of levels, and with starts != NULL writes the
 * order position where each level starts followed by the number of intra
 * blocks (depth + 1 entries, at most cap), or a negative cvp9_err_t. */
int cvp9_hip_intra_levels(const vp9_parsed_frame_t *pf, uint32_t *starts, uint32_t cap);

/* The code hip_decode_frame would return for each frame of a sequence
 * before any device work (CVP9_OK, CVP9_ERR_UNSUPPORTED or
 * CVP9_ERR_INVALID_DATA), applying the validator and the rule that after a
 * declined frame every frame is declined until the next KEY frame. Needs no
 * device. Returns n or a negative cvp9_err_t. */
int cvp9_hip_route_check(const vp9_parsed_frame_t *const *frames, int n, int *codes);
/* Copies the first n bytes of an export buffer (hip_export_buffer_alloc)
 * to host memory through HIP, after the backend's queued work; for checking
 * what a render-target decode wrote independently of the CPU mapping. */
cvp9_err_t cvp9_hip_export_read(void *ctx, const cvp9_export_buffer_t *buf, uint8_t *dst,
                                uint64_t n);

/* Reads n bytes at `offset` of a DMA-BUF of `size` bytes by importing it
 * as HIP external memory (as another GPU API would): for checking the
 * frames hip_get_frame_dmabuf hands out where the buffer cannot be mapped
 * into the CPU. The caller keeps fd. CVP9_ERR_UNSUPPORTED if the import
 * is refused. */
cvp9_err_t cvp9_hip_dmabuf_read(int fd, uint64_t size, uint64_t offset, uint8_t *dst, uint64_t n);

/* Copies `bytes` (multiple of 16) device to device `iters` times with a
 * plain streaming kernel on the current device; returns the achieved GB/s
 * (read + write bytes / time), or a negative value on error. Used to
 * calibrate HBM traffic counters and to report a measured copy peak. */
double cvp9_hip_copy_probe(uint64_t bytes, int iters);

/* Plain device memory for tests of entry points that take device buffers:
 * zero-filled memory on the current device (NULL on failure), its release,
 * and a blocking copy of n bytes from device memory to host memory. */
void *cvp9_hip_dev_alloc(uint64_t bytes);
void cvp9_hip_dev_free(void *p);
cvp9_err_t cvp9_hip_dev_read(const void *src, uint8_t *dst, uint64_t n);
/* Debug flags for one backend context (bit 0: skip the deblocking stage;
 * bit 1: every dependency wait of later frames gives up after one poll, so
 * a test can force the device-error path while every kernel still ends). */
#define CVP9_HIP_DEBUG_NO_DEBLOCK 1u
#define CVP9_HIP_DEBUG_SPIN_ONE   2u
cvp9_err_t hip_backend_set_debug(void *ctx, uint32_t flags);

/* Bound on every dependency wait of the engine's later batches, in polls
 * (0: the default, about 2 s). A tiny bound makes waits give up, so tests
 * can check that a timed-out wait is reported as CVP9_ERR_GPU while every
 * kernel still ends. The engine type is declared in cvp9_hip/engine.h. */
struct cvp9_engine;
cvp9_err_t cvp9_engine_set_spin_limit(struct cvp9_engine *e, uint32_t polls);

#ifdef __cplusplus
}
#endif
#endif
