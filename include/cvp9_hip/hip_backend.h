/*
 * cvp9_hip/hip_backend.h — the HIP reconstruction backend, C ABI.
 *
 * Drop-in for the reference's Vulkan backend: the same eight entry points
 * with the same argument meaning, ownership and error behaviour, called from
 * the same places in src/decoder/vp9_decoder.c (dispatch at :110-119,
 * :131-134, :203-207, :256-259, :286-289, :298-301, :310-313, :320-323).
 *
 * Output is bit-exact, per plane and per frame, to the reference's CPU
 * reconstruction (src/decoder/vp9_reconstruct.c) followed by its reference
 * pool update (vp9_decoder.c:224-231: KEY refreshes all 8 slots, every other
 * frame slot 0), which this backend keeps on the device.
 *
 * Pipelining: up to CVP9_HIP_INFLIGHT frames in flight on one HIP stream,
 * one event per slot (vulkan_backend.h:36-81 semantics).
 */
#ifndef CVP9_HIP_BACKEND_H
#define CVP9_HIP_BACKEND_H

#ifdef CVP9_HIP_USE_REFERENCE_HEADERS
#include "compute_vp9/decoder.h"
#include "decoder/vp9_parsed_frame.h"
#else
#include "cvp9_hip/abi_types.h"
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define CVP9_HIP_INFLIGHT 3

/* Replaces vulkan_backend_init (vulkan_backend.h:141). Creates a context on
 * the current HIP device: stream, device reference pool, staging ring.
 * CVP9_ERR_NO_BACKEND when no HIP device is usable, CVP9_ERR_NOMEM. */
cvp9_err_t hip_backend_init(void **ctx);

/* Replaces vulkan_backend_destroy (vulkan_backend.h:142). Waits for
 * in-flight work, frees everything. NULL is a no-op. */
void hip_backend_destroy(void *ctx);

/* Replaces vulkan_decode_frame (vulkan_backend.h:146-148). pf is borrowed
 * for the call only (copied into the slot's staging). Returns
 *   CVP9_OK               frame queued
 *   CVP9_ERR_AGAIN        CVP9_HIP_INFLIGHT frames in flight: get one first
 *   CVP9_ERR_UNSUPPORTED  layout this backend does not take (blocks off the
 *                         4x4 grid or overlapping, tx_size > 3, W or H = 1
 *                         mod 8, ...), or any non-KEY frame after such a
 *                         frame: nothing queued, the caller runs the CPU
 *                         path (vp9_decoder.c:209) until the next KEY frame
 *                         (see hip_backend_read_ref)
 * After a device error (CVP9_ERR_GPU from hip_get_frame) every later frame
 * reports CVP9_ERR_GPU until a KEY frame, which starts the pool afresh.
 *   CVP9_ERR_INVALID_DATA NULL arguments or coefficient offsets past num_coeffs
 *   CVP9_ERR_GPU / CVP9_ERR_NOMEM on HIP failures. */
cvp9_err_t hip_decode_frame(void *ctx, const vp9_parsed_frame_t *pf, int64_t pts);

/* Replaces vulkan_get_frame (vulkan_backend.h:150-153). Oldest frame first.
 * wait=0 returns CVP9_ERR_AGAIN while it is still on the GPU, unless the
 * ring is full (then it blocks); wait=1 blocks. CVP9_ERR_UNSUPPORTED when
 * nothing is queued. Plane pointers (I420, stride_y = W, stride_uv =
 * ceil(W/2)) stay valid until the slot is reused, CVP9_HIP_INFLIGHT-1
 * decodes later. CVP9_ERR_GPU if the device reported an error. */
cvp9_err_t hip_get_frame(void *ctx, cvp9_frame_info_t *info, int wait);

/* Replaces vulkan_get_frame_dmabuf (vulkan_backend.h:155). Like
 * hip_get_frame(wait=1) but returns a dup()'d DMA-BUF fd (caller closes)
 * of the slot's I420 output. CVP9_ERR_UNSUPPORTED without DMA-BUF export. */
cvp9_err_t hip_get_frame_dmabuf(void *ctx, cvp9_dmabuf_frame_t *out);

/* Replaces vulkan_export_buffer_alloc (vulkan_backend.h:157): a CPU-mapped,
 * device-accessible, DMA-BUF exportable buffer. CVP9_ERR_UNSUPPORTED when
 * DMA-BUF export is unavailable. */
cvp9_err_t hip_export_buffer_alloc(void *ctx, uint64_t size, cvp9_export_buffer_t *out);

/* Replaces vulkan_export_buffer_free (vulkan_backend.h:158). */
void hip_export_buffer_free(void *ctx, cvp9_export_buffer_t *buf);

/* Replaces vulkan_set_render_target (vulkan_backend.h:163): arms a one-shot
 * NV12 target (an export buffer, at least W*H + 2*ceil(W/2)*ceil(H/2)
 * bytes) for the next hip_decode_frame, which packs the frame into it on
 * the GPU. NULL disarms. */
cvp9_err_t hip_set_render_target(void *ctx, const cvp9_export_buffer_t *target);

/* No counterpart in the reference's Vulkan backend, which never declines a
 * frame. For the reference's dispatch (src/decoder/vp9_decoder.c:209-247)
 * when this backend declines a frame mid-stream: the CPU path that takes
 * over reconstructs from ctx->ref_frames, which frames decoded here never
 * updated. Reads reference slot `slot` (0..7) of the device pool as of the
 * last frame accepted: with frame->plane_y == NULL it only sets
 * frame->width/height (0 x 0 for an empty slot); otherwise frame must be
 * that size and its I420 planes (stride_y, stride_uv) receive the slot's
 * samples. Waits for queued work. After declining a frame the backend
 * declines every frame until the next KEY frame (which writes all 8 slots),
 * so the CPU path carries the stream until then. */
cvp9_err_t hip_backend_read_ref(void *ctx, int slot, cvp9_frame_info_t *frame);

#ifdef __cplusplus
}
#endif
#endif
