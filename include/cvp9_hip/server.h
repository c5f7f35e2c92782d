/*
 * cvp9_hip/server.h — continuous-batching request loop over the engine.
 *
 * No counterpart in the reference (one stream per decoder context,
 * src/decoder/vp9_decoder.c). A server admits streams against an HBM
 * budget for their device-resident state (reference pool, scratch, NV12
 * output), queues frames per stream as they arrive, and launches batches
 * made of the oldest queued frame of every stream that has one, while
 * earlier batches are still on the GPU. Every frame is packed as NV12 into
 * its stream's device buffer; completion is reported per frame with the
 * time it was observed. Output per stream is bit-exact to decoding the
 * stream alone.
 *
 * The server is single-threaded: one caller drives open/submit/pump/poll.
 */
#ifndef CVP9_HIP_SERVER_H
#define CVP9_HIP_SERVER_H

#include "cvp9_hip/engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cvp9_server cvp9_server_t;

typedef struct {
    uint64_t hbm_budget;    /* bytes of stream state the server may hold */
    uint32_t max_batch;     /* frames per batch (<= 4096) */
    uint32_t max_inflight;  /* batches on the GPU at once (1..4) */
} cvp9_server_config_t;

cvp9_err_t cvp9_server_create(cvp9_engine_t *e, const cvp9_server_config_t *cfg,
                              cvp9_server_t **out);
/* Waits for all work, releases every stream. */
void cvp9_server_destroy(cvp9_server_t *s);

/* HBM the server charges a stream of this size against the budget. */
uint64_t cvp9_server_stream_cost(uint32_t width, uint32_t height);

/* Admits a stream of width x height. CVP9_ERR_AGAIN when its cost would
 * exceed the budget (retry after another stream closes). */
cvp9_err_t cvp9_server_open(cvp9_server_t *s, uint32_t width, uint32_t height, int *stream);
/* Releases a stream once its queued frames have completed (earlier frames
 * still queued are dropped). Its device memory is recycled, not freed. */
void cvp9_server_close(cvp9_server_t *s, int stream);

/* Queues a prepared frame (must match the stream's size and stay alive
 * until it completes). Frames of one stream complete in submission order. */
cvp9_err_t cvp9_server_submit(cvp9_server_t *s, int stream, cvp9_prep_t *prep, uint64_t tag);

/* cvp9_server_submit for n frames at once (a driver that receives frames of
 * many streams together pays one call): frame i is preps[i] on streams[i]
 * with tags[i], in order. Stops at the first invalid entry and returns its
 * error; returns CVP9_OK when all n are queued. */
cvp9_err_t cvp9_server_submit_n(cvp9_server_t *s, int n, const int *streams,
                                cvp9_prep_t *const *preps, const uint64_t *tags);
/* Launches one batch (the oldest queued frame of each stream) unless
 * max_inflight batches are already on the GPU. Returns frames launched, or
 * a negative cvp9_err_t. */
int cvp9_server_pump(cvp9_server_t *s);

/* Reports up to `max` completed frames: their tags and the CLOCK_MONOTONIC
 * time (ns) at which completion was observed. Returns the count, or a
 * negative cvp9_err_t. */
int cvp9_server_poll(cvp9_server_t *s, uint64_t *tags, uint64_t *done_ns, int max);

/* Frames queued or in flight. */
uint32_t cvp9_server_pending(cvp9_server_t *s);

/* NV12 outputs rotate through CVP9_SERVER_NV12_RING device buffers per
 * stream: the output of the stream's frame number `frame` (0-based count
 * of its submitted frames) stays valid until frame + CVP9_SERVER_NV12_RING
 * is launched. Copies it (W*H luma, then ceil(W/2)*2 x ceil(H/2)
 * interleaved chroma) to host memory; synchronises. */
#define CVP9_SERVER_NV12_RING 4
cvp9_err_t cvp9_server_read_nv12(cvp9_server_t *s, int stream, uint64_t frame, uint8_t *dst,
                                 uint64_t cap);

/* Waits for everything; CVP9_ERR_GPU if a device error was flagged. */
cvp9_err_t cvp9_server_drain(cvp9_server_t *s);

#ifdef __cplusplus
}
#endif
#endif
