/*
 * cvp9_hip/engine.h — multi-stream reconstruction engine, C ABI.
 *
 * No counterpart in the reference, which decodes one stream per context
 * (src/decoder/vp9_decoder.c). This layer sits beside hip_backend.h for
 * callers that serve many streams on one GPU: every stream keeps its
 * reference pool resident in HBM, parsed frames are staged into HBM once
 * ("prepared"), and a batch of frames from different streams goes through
 * each reconstruction stage in a single launch. Per-stream output is the
 * same, bit for bit, as decoding that stream alone through hip_backend.h.
 */
#ifndef CVP9_HIP_ENGINE_H
#define CVP9_HIP_ENGINE_H

#include "cvp9_hip/abi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cvp9_engine cvp9_engine_t;
typedef struct cvp9_prep   cvp9_prep_t;
typedef struct cvp9_graph  cvp9_graph_t;

/* Engine on HIP device `device` with its own stream. */
cvp9_err_t cvp9_engine_create(int device, cvp9_engine_t **out);
void       cvp9_engine_destroy(cvp9_engine_t *e);

/* A decode stream: an empty 8-slot reference pool on the device. */
cvp9_err_t cvp9_engine_stream_open(cvp9_engine_t *e, int *stream_id);
void       cvp9_engine_stream_close(cvp9_engine_t *e, int stream_id);
/* Empties a stream's reference pool (as a new stream) and clears its error
 * state, keeping its device memory for reuse. */
cvp9_err_t cvp9_engine_stream_reset(cvp9_engine_t *e, int stream_id);
/* The HIP stream (hipStream_t) the most recent cvp9_engine_run enqueued
 * on, for callers recording an event after a batch. */
void      *cvp9_engine_hip_stream(cvp9_engine_t *e);
/* HIP queues the engine spreads batches over (1 or 2; default 2). A batch
 * runs on the queue its streams last ran on; a batch of streams that have
 * not run takes the queue the previous batch did not use; a batch whose
 * streams last ran on both waits for the other queue first. So batches of
 * disjoint stream sets (a caller alternating between two halves of its
 * streams) overlap on the GPU, and each stream's frames stay in order.
 * Graph capture, fused mode and per-stage timing use one queue. */
cvp9_err_t cvp9_engine_set_queues(cvp9_engine_t *e, int n);
/* Frame pipelining (default on). A frame's owner map and transforms do not
 * depend on the reference frames; with pipelining they run first, writing
 * every transform's residual as int16, on a queue paired with the batch's
 * queue, while the previous batch's motion compensation, intra and
 * deblocking still run; motion compensation then adds the inter residuals.
 * Each stream keeps two sets of owner map and residual buffers. Off: the
 * stages run in the order of the reference's frame loop on one queue.
 * Graph capture, fused mode and per-stage timing never overlap batches.
 * Output never depends on it. */
cvp9_err_t cvp9_engine_set_pipeline(cvp9_engine_t *e, int on);
/* A stream's device error flags, read on the host with no copy: 0 while
 * healthy, nonzero once a dependency wait of one of its frames gave up
 * (bits 0-7), a neighbour record was inconsistent (bits 8-15) or a grid
 * barrier of the fused kernel timed out (bits 16-23). A stream in error
 * takes no more frames (cvp9_engine_run returns CVP9_ERR_GPU) until
 * cvp9_engine_stream_reset. */
uint32_t   cvp9_engine_stream_error(cvp9_engine_t *e, int stream_id);
/* HBM held by a stream (frame pool, scratch). */
uint64_t   cvp9_engine_stream_bytes(cvp9_engine_t *e, int stream_id);

/* Validates pf (same rules and codes as hip_decode_frame) and copies its
 * device image into HBM. pf is not referenced afterwards. */
cvp9_err_t cvp9_engine_prepare(cvp9_engine_t *e, const vp9_parsed_frame_t *pf,
                               cvp9_prep_t **out);
void       cvp9_engine_prep_free(cvp9_engine_t *e, cvp9_prep_t *p);
/* HBM bytes of a prepared frame, and its algorithmic bytes (coefficients,
 * block records, reference footprints and the I420 write). */
uint64_t   cvp9_engine_prep_bytes(const cvp9_prep_t *p);

/* Enqueues one batch: frame preps[i] on stream streams[i] (each stream at
 * most once per batch), one launch per stage. Asynchronous; a prep must
 * stay alive until the batch completes. nv12[i] (or nv12 == NULL) is an
 * optional device pointer receiving the frame as NV12 (pitch W). */
cvp9_err_t cvp9_engine_run(cvp9_engine_t *e, int n, const int *streams,
                           cvp9_prep_t *const *preps, uint8_t *const *nv12);
/* Waits for all enqueued batches (both queues). CVP9_ERR_GPU if a device
 * error was flagged for any stream. */
cvp9_err_t cvp9_engine_sync(cvp9_engine_t *e);

/* Copies the stream's most recent frame (packed I420: Y W*H, U, V
 * ceil(W/2)*ceil(H/2)) to host memory; synchronises. */
cvp9_err_t cvp9_engine_read(cvp9_engine_t *e, int stream_id, uint8_t *dst, uint64_t cap,
                            uint32_t *width, uint32_t *height);

/* Launch shape of the stage kernels: waves sharing one transform group for
 * N = 4, 8, 16, 32 (1, 2 or 4), threads per intra block (64 or 256), the
 * number of persistent intra workers, the width (intra blocks over the
 * batch) from which a leading dependency level gets a launch of its own
 * instead of going to the persistent workers (0: persistent workers only;
 * at most 64 leading levels), and whether the ordered intra stage only
 * publishes each block's edge samples and leaves the blocks' samples to a
 * parallel pass after it (intra_defer 1) or writes every block itself (0),
 * and whether a batch of one frame that lists chains of intra blocks (frames
 * prepared by cvp9_engine_prepare do) runs its most critical chains each on
 * a worker of its own, which takes a block right after its chain
 * predecessor and gets that predecessor's edge samples in registers
 * (intra_chains 1; 64-thread workers, no deferred samples, no level
 * launches).
 * Output does not depend on it; cvp9_engine_autotune picks it for a
 * workload class. */
typedef struct {
    uint8_t  res_split[4];
    uint16_t intra_lanes;
    uint32_t intra_workers;
    uint32_t intra_level_min;
    uint32_t intra_defer;
    uint32_t intra_chains;
} cvp9_tune_t;
cvp9_err_t cvp9_engine_get_tune(cvp9_engine_t *e, cvp9_tune_t *t);
cvp9_err_t cvp9_engine_set_tune(cvp9_engine_t *e, const cvp9_tune_t *t);
int        cvp9_engine_compute_units(cvp9_engine_t *e);

/* Small autotuner: times candidate launch shapes on a representative batch
 * (frame preps[i] of tuning stream streams[i]; those streams' reference
 * pools advance, so use streams opened for tuning) `iters` times each,
 * keeps the fastest shape per stage, installs it and returns it in *best.
 * Synchronises. */
cvp9_err_t cvp9_engine_autotune(cvp9_engine_t *e, int n, const int *streams,
                                cvp9_prep_t *const *preps, int iters, cvp9_tune_t *best);

/* Fused mode: cvp9_engine_run launches one persistent kernel per batch
 * (every stage, separated by grid barriers, all workgroups co-resident)
 * instead of one kernel per stage. CVP9_ERR_UNSUPPORTED if the device
 * cannot hold the fused kernel's grid. Output is identical. */
cvp9_err_t cvp9_engine_set_fused(cvp9_engine_t *e, int on);

/* Graph capture: batches enqueued by cvp9_engine_run between begin and end
 * are recorded (descriptor copy + kernels), not executed, into a HIP graph
 * that cvp9_engine_graph_launch replays. The host-side reference pools
 * advance at record time. A replay reproduces the recorded frames exactly
 * when the recorded sequence starts with a KEY frame on every stream and
 * nothing else ran on those streams since. Streams must have run once at
 * the same sizes first (no allocation while recording: otherwise
 * CVP9_ERR_UNSUPPORTED). */
cvp9_err_t cvp9_engine_capture_begin(cvp9_engine_t *e);
cvp9_err_t cvp9_engine_capture_end(cvp9_engine_t *e, cvp9_graph_t **g);
cvp9_err_t cvp9_engine_graph_launch(cvp9_engine_t *e, cvp9_graph_t *g);
void       cvp9_engine_graph_free(cvp9_engine_t *e, cvp9_graph_t *g);

/* Per-stage timing with HIP events on the engine's stream. When on, every
 * stage launch is bracketed by events; cvp9_engine_stage_stats returns, for
 * each of the CVP9_ENGINE_STAGES stages, the summed milliseconds and the
 * number of launches since the last reset (it synchronises). Stage order:
 * fill, owner, mc, residual4, residual8, residual16, residual32,
 * neighbours (of intra blocks), intra, deblock, nv12, fused, expand (the
 * compact coefficients of a frame staged through the drop-in boundary,
 * written out as the dense pool before the transforms). */
#define CVP9_ENGINE_STAGES 13
cvp9_err_t cvp9_engine_set_timing(cvp9_engine_t *e, int on);
cvp9_err_t cvp9_engine_stage_stats(cvp9_engine_t *e, double *ms, uint32_t *launches,
                                   int reset);

/* A frame buffer in device memory in the engine's layout: Y rows of
 * roundup(W, 64) bytes (H rows), then U and V rows of roundup(ceil(W/2), 64)
 * bytes (ceil(H/2) rows each), then 64 bytes of slack; base 256-byte
 * aligned (hipMalloc's alignment). */
typedef struct {
    uint8_t *base;
    uint32_t width, height;
} cvp9_frame_buf_t;
/* Bytes of such a buffer for a width x height frame. */
uint64_t cvp9_frame_bytes(uint32_t width, uint32_t height);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): cvp9_engine_run for one frame, with the reference
 * frames and the output frame held by the caller, and every per-frame
 * scratch buffer (batch descriptor, owner map, intra edge and neighbour
 * records, residuals, error word) taken from a fixed device array of
 * CVP9_STATIC_SCRATCH_BYTES that the library holds on each GPU until the
 * process ends. It calls no hipMalloc, hipFree or other allocating routine.
 *
 *   prep     a prepared frame (cvp9_engine_prepare), W x H;
 *   refs     the 8 reference slots (refs[i].base == NULL or refs == NULL:
 *            empty slot, predicted as flat 127), each any size;
 *   out      the output frame, W x H; must not be one of the refs;
 *   nv12     optional device pointer receiving the frame as NV12 (pitch W);
 *   status   optional device-visible word receiving the frame's device error
 *            flags (0: none; see cvp9_engine_stream_error), in stream order;
 *   stream   the hipStream_t to enqueue on (NULL: the null stream).
 * Asynchronous. The reference pool rule (KEY: the frame goes to all 8 slots,
 * else slot 0) is the caller's. Output is bit-exact to cvp9_engine_run on
 * the same frame and references. Limits: at most 2^20 4x4 luma cells,
 * 49152 blocks, 26880 intra blocks and 12 Mi coefficients per frame (a
 * 3840x2160 frame of the benchmark workloads fits); a larger frame returns
 * CVP9_ERR_UNSUPPORTED without enqueuing anything. Every call on a GPU shares
 * that GPU's scratch array: calls that overlap on one GPU (other streams,
 * other threads) race, and nothing beside the call may keep waiting kernels
 * resident (the intra stage needs its workers co-resident). */
#define CVP9_STATIC_SCRATCH_BYTES (60ull * 1024 * 1024 + 8192)
cvp9_err_t cvp9_engine_run_static(const cvp9_prep_t *prep, const cvp9_frame_buf_t *refs,
                                  const cvp9_frame_buf_t *out, uint8_t *nv12, uint32_t *status,
                                  void *stream);

#ifdef __cplusplus
}
#endif
#endif
