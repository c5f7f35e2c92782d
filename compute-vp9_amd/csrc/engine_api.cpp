/*
 * engine_api.cpp — multi-stream engine C ABI (cvp9_hip/engine.h).
 */
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <new>
#include <vector>

#include "cvp9_hip/engine.h"
#include "cvp9_hip/testing.h"
#include "engine.hpp"

static_assert(CVP9_ENGINE_STAGES == STAGE_COUNT, "stage list of the ABI and the kernels differ");

using namespace cvp9;


namespace {

std::atomic<int> g_live_engines{0};

constexpr int kBatchSlots = 4;
/* Flags of the engine's ordering events (batch done, phase and queue order,
 * graph replay). They order work on the device and tell the host when a
 * batch's descriptor slot is free; the host reads no data through them, and
 * the kernels' own acquire and release fences make each kernel's stores
 * visible to the kernels after it on any queue. Without the system-scope
 * fence, a marker skips its cache write-back and invalidate, which cost
 * 2.5 us of the gap between consecutive frames on one 1080p stream. */
constexpr unsigned kSyncEvent = hipEventDisableTiming | hipEventDisableSystemFence;
/* 1: a batch whose stream set differs from its queue's previous batch still
 * overlaps the other queue (see run_batch); the library is built with 0, the
 * schedule model of the tests builds both */
#ifndef CVP9_OVERLAP_REGROUPED
#define CVP9_OVERLAP_REGROUPED 0
#endif
constexpr uint32_t kSpinLimit = 1u << 21;

struct BatchSlot {
    uint8_t *host = nullptr, *dev = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool used = false;
    uint64_t serial = 0; /* of the batch that last used the slot */
};

/* Stage order without frame pipelining, then the two phases with it: the
 * first depends only on the frame's own data and on reference slots the
 * previous frame does not write (fill, owner map, neighbour records,
 * transforms to int16 residuals, motion compensation of the early units),
 * the second on the previous frame (motion compensation of the late units,
 * with the residual add like the early ones, intra, deblocking). */
constexpr int kOrderPlain[] = {STAGE_EXPAND, STAGE_FILL, STAGE_OWNER, STAGE_MC, STAGE_RES4, STAGE_RES8,
                               STAGE_RES16, STAGE_RES32, STAGE_NBR, STAGE_INTRA, STAGE_LF,
                               STAGE_NV12};
constexpr int kOrderA[] = {STAGE_EXPAND, STAGE_FILL, STAGE_OWNER, STAGE_RES4, STAGE_RES8,
                           STAGE_RES16, STAGE_RES32, STAGE_NBR, STAGE_MC};
constexpr int kOrderB[] = {STAGE_MC, STAGE_INTRA, STAGE_LF, STAGE_NV12};

struct TimedLaunch {
    int stage;
    hipEvent_t a, b;
};

bool stage_has_work(int st, const StageCounts &n)
{
    switch (st) {
    case STAGE_FILL: return n.fill != 0;
    case STAGE_OWNER: return true;
    case STAGE_MC: return n.inter != 0;
    case STAGE_RES4: return n.grp[0] != 0;
    case STAGE_RES8: return n.grp[1] != 0;
    case STAGE_RES16: return n.grp[2] != 0;
    case STAGE_RES32: return n.grp[3] != 0;
    case STAGE_NBR:
    case STAGE_INTRA: return n.intra != 0;
    case STAGE_LF: return n.lf != 0;
    case STAGE_NV12: return n.nv12 != 0;
    case STAGE_EXPAND: return n.exp != 0;
    }
    return false;
}

cvp9_err_t to_err(hipError_t e)
{
    if (e == hipSuccess) return CVP9_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return CVP9_ERR_NOMEM;
    return CVP9_ERR_GPU;
}

/* Order-independent signature of a set of stream ids (nonzero). */
uint64_t stream_set_signature(const int *streams, int n)
{
    uint64_t h = 0x9e3779b97f4a7c15ull ^ (uint64_t)n;
    for (int i = 0; i < n; i++) {
        uint64_t x = (uint64_t)(uint32_t)streams[i] + 0x632be59bd9b4e019ull;
        x ^= x >> 31;
        x *= 0xbf58476d1ce4e5b9ull;
        x ^= x >> 29;
        h += x; /* a sum: the order of the ids does not matter */
    }
    return h ? h : 1;
}

uint64_t prep_algo_bytes(const vp9_parsed_frame_t *pf)
{
    /* coefficients of transforms taken + 32-byte block records + reference
     * footprints of inter blocks + I420 write */
    const int W = (int)pf->hdr.width, H = (int)pf->hdr.height;
    uint64_t b = 0;
    for (uint32_t i = 0; i < pf->num_blocks; i++) {
        const vp9_macroblock_info_t &k = pf->blocks[i];
        const int bw = k.x + k.width > W ? W - k.x : k.width;
        const int bh = k.y + k.height > H ? H - k.y : k.height;
        b += 32;
        if (bw <= 0 || bh <= 0) continue;
        if (!k.skip && k.tx_size <= 3) {
            const int T = 4 << k.tx_size;
            b += 2ull * (uint64_t)((bw / T) * (bh / T) * T * T);
        }
        if (!k.is_intra) {
            const int cbw = (bw + 1) / 2, cbh = (bh + 1) / 2;
            b += (uint64_t)((bw + 1) * (bh + 1) + 2 * (cbw + 1) * (cbh + 1));
        }
    }
    return b + (uint64_t)W * H + 2ull * ((W + 1) / 2) * ((H + 1) / 2);
}

} // namespace

struct cvp9_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<uint8_t *, uint8_t *>> desc; /* pinned source, device copy */
    uint32_t batches = 0;
};

struct cvp9_engine {
    int device = 0;
    /* Two HIP queues. A batch runs on the queue its streams last ran on, or,
     * for streams that have not run, on the queue the previous batch did not
     * use; when its streams last ran on both, the chosen queue first waits for
     * the other one's tail. Batches of disjoint stream sets (the usual pattern:
     * a caller alternating two halves of its streams) thus overlap, and every
     * stream's frames stay in order. Capture, fused mode and per-stage timing
     * use queue 0 only. */
    hipStream_t stream = nullptr;  /* queue 0 */
    hipStream_t stream2 = nullptr; /* queue 1 */
    /* Frame pipelining (default on): a batch's first phase runs on the
     * phase queue paired with its queue, so it overlaps the previous batch's
     * second phase; the second phase waits for it (event aev). A stream's
     * owner map and residual buffer alternate between two sets; a first
     * phase waits for the batch that last read its set. Capture, fused mode,
     * per-stage timing and one queue run both phases in order on one queue. */
    bool pipeline = true;
    uint32_t spin_limit = kSpinLimit; /* bound on every dependency wait */
    hipStream_t astream[2] = {nullptr, nullptr};
    hipEvent_t aev[2] = {nullptr, nullptr};
    int last_mode = -1;            /* overlap of the previous batch (0/1), -1 none */
    uint64_t serial = 0;           /* batches run so far */
    hipEvent_t oev[3] = {nullptr, nullptr, nullptr}; /* queue 1 / phase queues into queue 0 */
    hipEvent_t gev = nullptr;      /* after the latest graph replay on queue 0 */
    bool graph_pending[2] = {false, false}; /* phase queue q has not waited for it */
    int nqueues = 2;
    int last_q = 1;
    /* per queue: the stream set of its latest batch (a signature) and that
     * batch's serial, for the overlap rule in run_batch */
    uint64_t qsig[2] = {0, 0};
    uint64_t qserial[2] = {0, 0};
    std::vector<int8_t> sq;        /* per stream id: queue of its last batch, -1 none */
    hipEvent_t xev[2] = {nullptr, nullptr};
    std::vector<StreamState *> streams;
    BatchSlot slots[kBatchSlots];
    int next = 0;
    Tune tune = cvp9_default_tune(256);
    int cus = 256;
    bool timing = false;
    std::vector<TimedLaunch> timed;
    std::vector<hipEvent_t> free_events;
    double stage_ms[STAGE_COUNT] = {0};
    uint32_t stage_n[STAGE_COUNT] = {0};
    bool fused = false;
    uint32_t fused_wgs = 0;
    cvp9_graph *capture = nullptr; /* recording batches into a graph */
    bool counted = false;          /* counted in g_live_engines */

    hipStream_t queue(int q) const { return q ? stream2 : stream; }

    hipError_t sync_all()
    {
        hipError_t r = hipSuccess;
        for (hipStream_t s : {stream, stream2, astream[0], astream[1]}) {
            hipError_t x = s ? hipStreamSynchronize(s) : hipSuccess;
            if (r == hipSuccess) r = x;
        }
        return r;
    }

    /* queue 0 waits (on the device) for everything enqueued on the others */
    hipError_t order_all_into_0()
    {
        const hipStream_t src[3] = {stream2, astream[0], astream[1]};
        for (int i = 0; i < 3; i++) {
            hipError_t r = hipEventRecord(oev[i], src[i]);
            if (r == hipSuccess) r = hipStreamWaitEvent(stream, oev[i], 0);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    }

    /* the latest batch on queue q may still be running */
    bool queue_busy(int q) const
    {
        hipEvent_t ev = batch_event(qserial[q]);
        return ev && hipEventQuery(ev) == hipErrorNotReady;
    }

    /* done event of the batch with this serial while it may still run */
    hipEvent_t batch_event(uint64_t s) const
    {
        for (const BatchSlot &b : slots)
            if (b.used && b.serial == s) return b.done;
        return nullptr;
    }

    /* queue dst waits (on the device) for everything enqueued on src so far */
    hipError_t order(int src, int dst)
    {
        hipError_t e = hipEventRecord(xev[src], queue(src));
        return e != hipSuccess ? e : hipStreamWaitEvent(queue(dst), xev[src], 0);
    }

    hipEvent_t take_event()
    {
        if (!free_events.empty()) {
            hipEvent_t e = free_events.back();
            free_events.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }

    void resolve_timing()
    {
        for (TimedLaunch &t : timed) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
                stage_ms[t.stage] += ms;
                stage_n[t.stage]++;
            }
            free_events.push_back(t.a);
            free_events.push_back(t.b);
        }
        timed.clear();
    }
};

int cvp9_live_engines()
{
    return g_live_engines.load(std::memory_order_relaxed);
}

extern "C" {

cvp9_err_t cvp9_engine_create(int device, cvp9_engine_t **out)
{
    if (!out) return CVP9_ERR_INVALID_DATA;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return CVP9_ERR_NO_BACKEND;
    if (init_device(device) != hipSuccess) return CVP9_ERR_NO_BACKEND;
    cvp9_engine_t *e = new (std::nothrow) cvp9_engine();
    if (!e) return CVP9_ERR_NOMEM;
    e->device = device;
    hipSetDevice(device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    {
        e->cus = prop.multiProcessorCount;
        e->tune = cvp9_default_tune(prop.multiProcessorCount);
    }
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return CVP9_ERR_GPU;
    }
    if (hipStreamCreateWithFlags(&e->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->xev[0], kSyncEvent) != hipSuccess ||
        hipEventCreateWithFlags(&e->xev[1], kSyncEvent) != hipSuccess ||
        hipStreamCreateWithFlags(&e->astream[0], hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&e->astream[1], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->aev[0], kSyncEvent) != hipSuccess ||
        hipEventCreateWithFlags(&e->aev[1], kSyncEvent) != hipSuccess ||
        hipEventCreateWithFlags(&e->oev[0], kSyncEvent) != hipSuccess ||
        hipEventCreateWithFlags(&e->oev[1], kSyncEvent) != hipSuccess ||
        hipEventCreateWithFlags(&e->oev[2], kSyncEvent) != hipSuccess ||
        hipEventCreateWithFlags(&e->gev, kSyncEvent) != hipSuccess) {
        cvp9_engine_destroy(e);
        return CVP9_ERR_GPU;
    }
    e->fused_wgs = cvp9_fused_max_wgs(device);
    for (BatchSlot &s : e->slots) {
        if (hipEventCreateWithFlags(&s.done, kSyncEvent) != hipSuccess) {
            cvp9_engine_destroy(e);
            return CVP9_ERR_GPU;
        }
    }
    g_live_engines.fetch_add(1, std::memory_order_relaxed);
    e->counted = true;
    *out = e;
    return CVP9_OK;
}

void cvp9_engine_destroy(cvp9_engine_t *e)
{
    if (!e) return;
    if (e->counted) g_live_engines.fetch_sub(1, std::memory_order_relaxed);
    hipSetDevice(e->device);
    if (e->stream) e->sync_all();
    for (StreamState *s : e->streams) delete s;
    for (BatchSlot &s : e->slots) {
        if (s.done) hipEventDestroy(s.done);
        if (s.host) hipHostFree(s.host);
        if (s.dev) hipFree(s.dev);
    }
    e->resolve_timing();
    for (hipEvent_t ev : e->free_events) hipEventDestroy(ev);
    for (hipEvent_t ev : e->xev)
        if (ev) hipEventDestroy(ev);
    for (hipEvent_t ev : e->aev)
        if (ev) hipEventDestroy(ev);
    for (hipEvent_t ev : e->oev)
        if (ev) hipEventDestroy(ev);
    if (e->gev) hipEventDestroy(e->gev);
    for (hipStream_t s : e->astream)
        if (s) hipStreamDestroy(s);
    if (e->stream2) hipStreamDestroy(e->stream2);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
}

cvp9_err_t cvp9_engine_stream_open(cvp9_engine_t *e, int *id)
{
    if (!e || !id) return CVP9_ERR_INVALID_DATA;
    StreamState *s = new (std::nothrow) StreamState(e->device);
    if (!s) return CVP9_ERR_NOMEM;
    for (size_t i = 0; i < e->streams.size(); i++)
        if (!e->streams[i]) {
            e->streams[i] = s;
            e->sq[i] = -1;
            *id = (int)i;
            return CVP9_OK;
        }
    e->streams.push_back(s);
    e->sq.push_back(-1);
    *id = (int)e->streams.size() - 1;
    return CVP9_OK;
}

void cvp9_engine_stream_close(cvp9_engine_t *e, int id)
{
    if (!e || id < 0 || id >= (int)e->streams.size() || !e->streams[id]) return;
    hipSetDevice(e->device);
    e->sync_all(); /* in-flight batches may still read it */
    delete e->streams[id];
    e->streams[id] = nullptr;
}

cvp9_err_t cvp9_engine_stream_reset(cvp9_engine_t *e, int id)
{
    if (!e || id < 0 || id >= (int)e->streams.size() || !e->streams[id]) return CVP9_ERR_INVALID_DATA;
    hipSetDevice(e->device);
    hipError_t he = e->sync_all();
    if (he != hipSuccess) return to_err(he);
    e->sq[id] = 0;
    return to_err(e->streams[id]->reset());
}

void *cvp9_engine_hip_stream(cvp9_engine_t *e)
{
    return e ? (void *)e->queue(e->last_q) : nullptr;
}

cvp9_err_t cvp9_engine_set_queues(cvp9_engine_t *e, int n)
{
    if (!e || n < 1 || n > 2) return CVP9_ERR_INVALID_DATA;
    e->nqueues = n;
    return CVP9_OK;
}

cvp9_err_t cvp9_engine_set_spin_limit(cvp9_engine_t *e, uint32_t polls)
{
    if (!e) return CVP9_ERR_INVALID_DATA;
    e->spin_limit = polls ? polls : kSpinLimit;
    return CVP9_OK;
}

cvp9_err_t cvp9_engine_set_pipeline(cvp9_engine_t *e, int on)
{
    if (!e) return CVP9_ERR_INVALID_DATA;
    e->pipeline = on != 0;
    return CVP9_OK;
}

uint32_t cvp9_engine_stream_error(cvp9_engine_t *e, int id)
{
    if (!e || id < 0 || id >= (int)e->streams.size() || !e->streams[id]) return 0;
    return e->streams[id]->error();
}

uint64_t cvp9_engine_stream_bytes(cvp9_engine_t *e, int id)
{
    if (!e || id < 0 || id >= (int)e->streams.size() || !e->streams[id]) return 0;
    return e->streams[id]->resident_bytes();
}

cvp9_err_t cvp9_engine_prepare(cvp9_engine_t *e, const vp9_parsed_frame_t *pf, cvp9_prep_t **out)
{
    if (!e || !pf || !out) return CVP9_ERR_INVALID_DATA;
    *out = nullptr;
    cvp9_prep_t *p = new (std::nothrow) cvp9_prep();
    if (!p) return CVP9_ERR_NOMEM;
    p->L.chains = true; /* chain workers may take it (resident inputs: planned once) */
    cvp9_err_t err = plan_measure(pf, p->L);
    if (err != CVP9_OK) {
        delete p;
        return err;
    }
    hipSetDevice(e->device);
    std::vector<uint8_t> host(p->L.bytes);
    plan_write(pf, p->L, host.data());
    hipError_t he = hipMalloc(&p->dev, p->L.bytes);
    if (he == hipSuccess) he = hipMemcpy(p->dev, host.data(), p->L.bytes, hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        if (p->dev) hipFree(p->dev);
        delete p;
        return to_err(he);
    }
    p->algo = prep_algo_bytes(pf);
    *out = p;
    return CVP9_OK;
}

void cvp9_engine_prep_free(cvp9_engine_t *e, cvp9_prep_t *p)
{
    if (!p) return;
    if (e) {
        hipSetDevice(e->device);
        e->sync_all();
    }
    hipFree(p->dev);
    delete p;
}

uint64_t cvp9_engine_prep_bytes(const cvp9_prep_t *p)
{
    return p ? p->algo : 0;
}

} // extern "C"

namespace {
/* a prepared image still in host memory, copied to its device place on the
 * batch's first queue before anything reads it */
struct Upload {
    const uint8_t *host;
    uint8_t *dev;
    size_t bytes;
};

cvp9_err_t run_batch(cvp9_engine_t *e, int n, const int *streams, cvp9_prep_t *const *preps,
                     uint8_t *const *nv12, const Upload *up);
} // namespace

cvp9_err_t cvp9_engine_run(cvp9_engine_t *e, int n, const int *streams, cvp9_prep_t *const *preps,
                           uint8_t *const *nv12)
{
    return run_batch(e, n, streams, preps, nv12, nullptr);
}

namespace cvp9 {
DFrame engine_stream_ref(cvp9_engine_t *e, int sid, int slot)
{
    DFrame f = {};
    if (!e || sid < 0 || sid >= (int)e->streams.size() || !e->streams[sid] || slot < 0 || slot > 7)
        return f;
    const FrameBuf *b = e->streams[sid]->ref(slot);
    return b ? b->dframe() : f;
}

cvp9_err_t engine_run_staged(cvp9_engine_t *e, int sid, const PrepLayout &L, const uint8_t *host,
                             uint8_t *dev, uint8_t *nv12, StagedResult *res)
{
    if (!e || !host || !dev || !res) return CVP9_ERR_INVALID_DATA;
    cvp9_prep p;
    p.L = L;
    p.dev = dev;
    cvp9_prep_t *pp = &p;
    const Upload up{host, dev, (size_t)L.upload};
    const cvp9_err_t err = run_batch(e, 1, &sid, &pp, nv12 ? &nv12 : nullptr, &up);
    p.dev = nullptr; /* not the engine's to free */
    if (err != CVP9_OK) return err;
    StreamState *s = e->streams[sid];
    res->out = s->ref(0)->dframe(); /* every frame becomes slot 0 */
    res->queue = e->queue(e->last_q);
    res->err = s->err_word();
    return CVP9_OK;
}
} // namespace cvp9

namespace {
/* The second phase of a one-job batch follows its first phase on another
 * queue. A queue wait on an event that has already finished when the wait
 * is queued costs that queue ~0.3 us; one still pending then costs it
 * 2.4-4.4 us at the moment the queue reaches it, which for one stream is on
 * the chain between frames. So while the queue still has the previous
 * batch's work ahead, the host polls the first phase's end (a fraction of a
 * microsecond per query) and queues the second phase with no wait once it
 * has finished. It gives up (and queues the wait) when the previous batch
 * finishes first, when the queue has nothing ahead (a host-bound caller
 * never spins), or after kHostWaitUs. */
constexpr double kHostWaitUs = 2000.0;

bool phase_a_done(hipEvent_t aev, hipEvent_t prev)
{
    if (!prev || hipEventQuery(prev) != hipErrorNotReady) return false;
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        const hipError_t r = hipEventQuery(aev);
        if (r == hipSuccess) return true;
        if (r != hipErrorNotReady || hipEventQuery(prev) != hipErrorNotReady) return false;
        if (std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() >
            kHostWaitUs)
            return false;
    }
}

cvp9_err_t run_batch(cvp9_engine_t *e, int n, const int *streams, cvp9_prep_t *const *preps,
                     uint8_t *const *nv12, const Upload *up)
{
    if (!e || n <= 0 || n > CVP9_MAX_BATCH || !streams || !preps) return CVP9_ERR_INVALID_DATA;
    hipSetDevice(e->device);
    std::vector<DJob> jobs((size_t)n);
    /* one frame per stream per batch: the stream's scratch is per batch */
    for (int i = 0; i < n; i++) {
        const int id = streams[i];
        if (id < 0 || id >= (int)e->streams.size() || !e->streams[id] || !preps[i])
            return CVP9_ERR_INVALID_DATA;
        for (int k = 0; k < i; k++)
            if (streams[k] == id) return CVP9_ERR_INVALID_DATA;
    }
    /* A stream whose earlier frame reported a device error (a dependency
     * wait that gave up) takes no more frames until it is reset: its
     * reference pool is not trustworthy, and a wait that keeps timing out
     * would otherwise repeat batch after batch. The word is host memory
     * the kernels write, so this costs no copy. */
    for (int i = 0; i < n; i++)
        if (e->streams[streams[i]]->error()) return CVP9_ERR_GPU;
    hipError_t he;
    /* the queue (see the engine's description) */
    const bool single = e->capture || e->fused || e->timing || e->nqueues < 2;
    int q = -1;
    for (int i = 0; i < n && q < 0; i++) q = e->sq[streams[i]];
    if (single) q = 0;
    else if (q < 0) q = 1 - e->last_q;
    bool other = false;
    for (int i = 0; i < n; i++) other |= e->sq[streams[i]] >= 0 && e->sq[streams[i]] != q;
    /* Batches overlap on the two queues only in the pattern the tests cover:
     * each queue's consecutive batches hold the same streams (a caller
     * alternating two halves of its streams). A batch whose stream set
     * differs from its queue's previous batch while the other queue is
     * still busy (three or more stream groups alternating over the two
     * queues) runs after the other queue's work instead: the overlapped
     * form of that schedule stalled once on MI355X with no cause found. */
    const uint64_t sig = stream_set_signature(streams, n);
    if (!CVP9_OVERLAP_REGROUPED && !single && !other && e->qsig[q] && e->qsig[q] != sig &&
        e->queue_busy(1 - q))
        other = true;
    if (other && (he = e->order(1 - q, q)) != hipSuccess) return to_err(he);
    const hipStream_t qs = e->queue(q);
    const bool resid_all = e->pipeline && !e->fused;
    const bool overlap = resid_all && !single;
    const hipStream_t qa = overlap ? e->astream[q] : qs; /* first phase */
    if (e->last_mode >= 0 && e->last_mode != (int)overlap && !e->capture &&
        (he = e->sync_all()) != hipSuccess) /* sets and queues change meaning */
        return to_err(he);
    e->last_mode = overlap;
    for (int i = 0; i < n; i++) e->streams[streams[i]]->set_two_sets(overlap);
    if (e->capture) { /* nothing may synchronise or memset while recording */
        for (int i = 0; i < n; i++)
            if (!e->streams[streams[i]]->scratch_fits(preps[i]->L)) return CVP9_ERR_UNSUPPORTED;
    } else if (overlap) {
        /* buffers that in-flight work reads are reallocated only when idle */
        bool fits = true;
        for (int i = 0; i < n; i++) fits &= e->streams[streams[i]]->scratch_fits(preps[i]->L);
        if (!fits && (he = e->sync_all()) != hipSuccess) return to_err(he);
    }
    const uint64_t serial = ++e->serial;
    std::vector<hipEvent_t> readers; /* batches still reading the sets written now */
    for (int i = 0; i < n; i++) {
        StreamState *s = e->streams[streams[i]];
        const PrepLayout &L = preps[i]->L;
        if (!e->capture && (he = s->ensure_scratch(L, qa)) != hipSuccess) return to_err(he);
        if (overlap) {
            hipEvent_t ev = e->batch_event(s->reader(s->parity()));
            if (ev && std::find(readers.begin(), readers.end(), ev) == readers.end())
                readers.push_back(ev);
        }
        FrameBuf *out = nullptr;
        if ((he = s->acquire(L.W, L.H, &out, !e->capture, overlap)) != hipSuccess) return to_err(he);
        const uint32_t early = overlap ? s->mc_independent(L) : L.n_mc;
        jobs[i] = s->make_job(L, preps[i]->dev, out, nv12 ? nv12[i] : nullptr, serial);
        jobs[i].n_mc_a = early;
        s->update_refs(L.key != 0, out);
    }

    const size_t need = BatchBuilder::bytes_for((uint32_t)n);
    uint8_t *dhost = nullptr, *ddev = nullptr;
    BatchSlot *bs = nullptr;
    if (e->capture) {
        /* a recorded batch owns its descriptor for the graph's lifetime */
        if ((he = hipHostMalloc((void **)&dhost, need, hipHostMallocDefault)) != hipSuccess) return to_err(he);
        if ((he = hipMalloc((void **)&ddev, need)) != hipSuccess) {
            hipHostFree(dhost);
            return to_err(he);
        }
        e->capture->desc.emplace_back(dhost, ddev);
        e->capture->batches++;
    } else {
        bs = &e->slots[e->next];
        e->next = (e->next + 1) % kBatchSlots;
        if (bs->used && hipEventSynchronize(bs->done) != hipSuccess) return CVP9_ERR_GPU;
        if (bs->cap < need) {
            if (bs->host) hipHostFree(bs->host);
            if (bs->dev) hipFree(bs->dev);
            bs->host = bs->dev = nullptr;
            bs->cap = 0;
            if ((he = hipHostMalloc((void **)&bs->host, need, hipHostMallocDefault)) != hipSuccess ||
                (he = hipMalloc((void **)&bs->dev, need)) != hipSuccess)
                return to_err(he);
            bs->cap = need;
        }
        dhost = bs->host;
        ddev = bs->dev;
    }
    BatchBuilder bb;
    std::vector<const PrepLayout *> lays((size_t)n);
    for (int i = 0; i < n; i++) lays[i] = &preps[i]->L;
    Tune tune = e->tune;
    if (e->fused) tune.intra_level_min = 0; /* the fused kernel runs persistent workers only */
    bb.build(jobs.data(), lays.data(), (uint32_t)n, dhost, ddev, tune, e->spin_limit, e->fused);
    bb.set_resid_all(resid_all);
    /* (a reader that has finished needs no wait: a queue wait costs the
     * host several microseconds, a query a fraction of one, and when the
     * host is what limits the rate the device is usually done) */
    for (hipEvent_t ev : readers)
        if (hipEventQuery(ev) != hipSuccess && (he = hipStreamWaitEvent(qa, ev, 0)) != hipSuccess)
            return to_err(he);
    if (overlap && e->graph_pending[q]) {
        if ((he = hipStreamWaitEvent(qa, e->gev, 0)) != hipSuccess) return to_err(he);
        e->graph_pending[q] = false;
    }
    if (up && (he = hipMemcpyAsync(up->dev, up->host, up->bytes, hipMemcpyHostToDevice, qa)) !=
                  hipSuccess)
        return to_err(he);
    if ((he = hipMemcpyAsync(ddev, dhost, bb.used(), hipMemcpyHostToDevice, qa)) != hipSuccess)
        return to_err(he);
    const bool timing = e->timing && !e->capture;
    if (e->fused) {
        TimedLaunch t{STAGE_FUSED, nullptr, nullptr};
        if (timing) {
            t.a = e->take_event();
            t.b = e->take_event();
            if (!t.a || !t.b) return CVP9_ERR_GPU;
            hipEventRecord(t.a, qs);
        }
        /* compact coefficients are expanded by a launch of their own first */
        if (bb.counts().exp &&
            (he = cvp9_launch_stage(STAGE_EXPAND, bb.batch(), bb.counts(), qs)) != hipSuccess)
            return to_err(he);
        /* every workgroup of a fused launch must be resident at once: with
         * several engines in the process, each takes its share */
        const int live = cvp9_live_engines();
        const uint32_t wgs = live > 1 ? std::max(e->fused_wgs / (uint32_t)live, 8u) : e->fused_wgs;
        if ((he = cvp9_launch_fused(bb.batch(), wgs, qs)) != hipSuccess) return to_err(he);
        if (timing) {
            hipEventRecord(t.b, qs);
            e->timed.push_back(t);
        }
    }
    /* stages in order; late: the second phase (its motion compensation takes
     * the late units). done: an event for the phase's end; its last launch
     * records it by its own completion when that launch is a single kernel
     * (the marker packet of a separate record sits on the chain between
     * frames), else it stays set for the caller to record. */
    auto run_stages = [&](const int *order, int count, hipStream_t qx, bool late,
                          hipEvent_t *done) -> hipError_t {
        auto has_work = [&](int st) {
            return late && st == STAGE_MC ? bb.counts().inter2 != 0 : stage_has_work(st, bb.counts());
        };
        int last = -1;
        for (int i = 0; i < count; i++)
            if (has_work(order[i])) last = i;
        for (int i = 0; i < count; i++) {
            const int st = order[i];
            const bool timed = timing && has_work(st);
            hipEvent_t *const d = i == last && !timing ? done : nullptr;
            TimedLaunch t{st, nullptr, nullptr};
            if (timed) {
                t.a = e->take_event();
                t.b = e->take_event();
                if (!t.a || !t.b) return hipErrorOutOfMemory;
                hipEventRecord(t.a, qx);
            }
            hipError_t r;
            int span = 1;
            if (late && st == STAGE_MC) {
                r = cvp9_launch_mc_late(bb.batch(), bb.counts(), qx, d);
            } else {
                /* outside per-stage timing, independent stages share launches */
                span = timing ? 1 : cvp9_stage_span(st, bb.batch(), bb.counts());
                r = cvp9_launch_stages(st, span, bb.batch(), bb.counts(), qx, d);
            }
            if (r != hipSuccess) return r;
            i += span - 1;
            if (timed) {
                hipEventRecord(t.b, qx);
                e->timed.push_back(t);
            }
        }
        return hipSuccess;
    };
    hipEvent_t done = bs ? bs->done : nullptr; /* cleared once a kernel records it */
    const hipEvent_t prev_on_q = e->batch_event(e->qserial[q]); /* the queue's previous batch */
    if (!e->fused) {
        if (!resid_all) {
            he = run_stages(kOrderPlain, (int)(sizeof(kOrderPlain) / sizeof(int)), qs, false, &done);
        } else {
            /* (phase A's end is recorded by a marker: it is off the chain
             * between frames, and a kernel-recorded event costs the host
             * more per launch) */
            he = run_stages(kOrderA, (int)(sizeof(kOrderA) / sizeof(int)), qa, false, nullptr);
            if (he == hipSuccess && overlap) {
                he = hipEventRecord(e->aev[q], qa);
                if (he == hipSuccess && !(n == 1 && phase_a_done(e->aev[q], prev_on_q)))
                    he = hipStreamWaitEvent(qs, e->aev[q], 0);
            }
            if (he == hipSuccess)
                he = run_stages(kOrderB, (int)(sizeof(kOrderB) / sizeof(int)), qs, true, &done);
        }
        if (he != hipSuccess) return to_err(he);
    }
    if (bs) {
        if (done && (he = hipEventRecord(done, qs)) != hipSuccess) return to_err(he);
        bs->used = true;
        bs->serial = serial;
    }
    for (int i = 0; i < n; i++) e->sq[streams[i]] = (int8_t)q;
    e->last_q = q;
    e->qsig[q] = sig;
    e->qserial[q] = serial;
    return CVP9_OK;
}
} // namespace

extern "C" {

cvp9_err_t cvp9_engine_set_fused(cvp9_engine_t *e, int on)
{
    if (!e) return CVP9_ERR_INVALID_DATA;
    if (on && e->fused_wgs == 0) return CVP9_ERR_UNSUPPORTED;
    e->fused = on != 0;
    return CVP9_OK;
}

static void free_graph(cvp9_graph *g)
{
    if (!g) return;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    for (auto &d : g->desc) {
        hipHostFree(d.first);
        hipFree(d.second);
    }
    delete g;
}

cvp9_err_t cvp9_engine_capture_begin(cvp9_engine_t *e)
{
    if (!e || e->capture) return CVP9_ERR_INVALID_DATA;
    hipSetDevice(e->device);
    if (e->sync_all() != hipSuccess) return CVP9_ERR_GPU;
    for (int8_t &q : e->sq) q = q >= 0 ? 0 : -1; /* queue 1 is idle: all work is ordered on 0 */
    e->capture = new (std::nothrow) cvp9_graph();
    if (!e->capture) return CVP9_ERR_NOMEM;
    if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
        free_graph(e->capture);
        e->capture = nullptr;
        return CVP9_ERR_GPU;
    }
    return CVP9_OK;
}

cvp9_err_t cvp9_engine_capture_end(cvp9_engine_t *e, cvp9_graph_t **out)
{
    if (!e || !e->capture || !out) return CVP9_ERR_INVALID_DATA;
    cvp9_graph *g = e->capture;
    e->capture = nullptr;
    *out = nullptr;
    if (hipStreamEndCapture(e->stream, &g->graph) != hipSuccess || !g->graph ||
        hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
        free_graph(g);
        return CVP9_ERR_GPU;
    }
    *out = g;
    return CVP9_OK;
}

cvp9_err_t cvp9_engine_graph_launch(cvp9_engine_t *e, cvp9_graph_t *g)
{
    if (!e || !g || !g->exec) return CVP9_ERR_INVALID_DATA;
    hipSetDevice(e->device);
    /* the recorded batches run on queue 0, after everything queued so far;
     * later first phases (which write the sets the replay reads) wait for it */
    if (e->order_all_into_0() != hipSuccess || hipGraphLaunch(g->exec, e->stream) != hipSuccess ||
        hipEventRecord(e->gev, e->stream) != hipSuccess)
        return CVP9_ERR_GPU;
    e->graph_pending[0] = e->graph_pending[1] = true;
    for (int8_t &q : e->sq) q = q >= 0 ? 0 : -1;
    e->last_q = 0;
    return CVP9_OK;
}

void cvp9_engine_graph_free(cvp9_engine_t *e, cvp9_graph_t *g)
{
    if (!g) return;
    if (e) {
        hipSetDevice(e->device);
        e->sync_all();
    }
    free_graph(g);
}

cvp9_err_t cvp9_engine_sync(cvp9_engine_t *e)
{
    if (!e) return CVP9_ERR_INVALID_DATA;
    hipSetDevice(e->device);
    if (e->sync_all() != hipSuccess) return CVP9_ERR_GPU;
    e->resolve_timing();
    for (StreamState *s : e->streams)
        if (s && s->error()) return CVP9_ERR_GPU;
    return CVP9_OK;
}

cvp9_err_t cvp9_engine_read(cvp9_engine_t *e, int id, uint8_t *dst, uint64_t cap, uint32_t *width,
                            uint32_t *height)
{
    if (!e || id < 0 || id >= (int)e->streams.size() || !e->streams[id]) return CVP9_ERR_INVALID_DATA;
    FrameBuf *f = e->streams[id]->ref(0);
    if (!f) return CVP9_ERR_UNSUPPORTED;
    const uint32_t w = f->w, h = f->h, cw = (w + 1) / 2, ch = (h + 1) / 2;
    if (width) *width = w;
    if (height) *height = h;
    const uint64_t need = (uint64_t)w * h + 2ull * cw * ch;
    if (!dst || cap < need) return CVP9_ERR_INVALID_DATA;
    hipSetDevice(e->device);
    if (e->sync_all() != hipSuccess) return CVP9_ERR_GPU;
    const DFrame d = f->dframe();
    if (hipMemcpy2D(dst, w, d.y, d.sy, w, h, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy2D(dst + (size_t)w * h, cw, d.u, d.suv, cw, ch, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy2D(dst + (size_t)w * h + (size_t)cw * ch, cw, d.v, d.suv, cw, ch,
                    hipMemcpyDeviceToHost) != hipSuccess)
        return CVP9_ERR_GPU;
    return e->streams[id]->error() ? CVP9_ERR_GPU : CVP9_OK;
}

cvp9_err_t cvp9_engine_get_tune(cvp9_engine_t *e, cvp9_tune_t *t)
{
    if (!e || !t) return CVP9_ERR_INVALID_DATA;
    for (int i = 0; i < 4; i++) t->res_split[i] = e->tune.res_split[i];
    t->intra_lanes = e->tune.intra_lanes;
    t->intra_workers = e->tune.intra_workers;
    t->intra_level_min = e->tune.intra_level_min;
    t->intra_defer = e->tune.intra_defer;
    t->intra_chains = e->tune.intra_chains;
    return CVP9_OK;
}

cvp9_err_t cvp9_engine_set_tune(cvp9_engine_t *e, const cvp9_tune_t *t)
{
    if (!e || !t) return CVP9_ERR_INVALID_DATA;
    for (int i = 0; i < 4; i++)
        if (t->res_split[i] != 1 && t->res_split[i] != 2 && t->res_split[i] != 4)
            return CVP9_ERR_INVALID_DATA;
    if (t->intra_lanes != 64 && t->intra_lanes != 256) return CVP9_ERR_INVALID_DATA;
    if (t->intra_workers < 1 || t->intra_workers > 65536) return CVP9_ERR_INVALID_DATA;
    if (t->intra_defer > 1 || t->intra_chains > 1) return CVP9_ERR_INVALID_DATA;
    for (int i = 0; i < 4; i++) e->tune.res_split[i] = t->res_split[i];
    e->tune.intra_lanes = t->intra_lanes;
    e->tune.intra_workers = t->intra_workers;
    e->tune.intra_level_min = t->intra_level_min;
    e->tune.intra_defer = t->intra_defer;
    e->tune.intra_chains = t->intra_chains;
    return CVP9_OK;
}

int cvp9_engine_compute_units(cvp9_engine_t *e)
{
    return e ? e->cus : 0;
}

cvp9_err_t cvp9_engine_autotune(cvp9_engine_t *e, int n, const int *streams,
                                cvp9_prep_t *const *preps, int iters, cvp9_tune_t *best)
{
    if (!e || n <= 0 || !streams || !preps || iters < 1) return CVP9_ERR_INVALID_DATA;
    const uint32_t cu = (uint32_t)(e->cus > 0 ? e->cus : 256);
    /* each candidate run measures every stage at once; the transform split
     * is uniform over sizes inside a run and chosen per size afterwards */
    struct Cand { uint8_t split; uint16_t lanes; uint32_t workers; uint32_t level_min; uint32_t defer;
                  uint32_t chains = 0; };
    /* (the intra launch holds the worker count to what stays resident) */
    const Cand cands[] = {
        {1, 256, 4 * cu, 0, 0},      {2, 256, 4 * cu, 0, 0},      {4, 256, 3 * cu, 0, 0},
        {1, 64, 8 * cu, 0, 0},       {2, 64, 16 * cu, 0, 0},      {4, 64, 12 * cu, 0, 0},
        {1, 256, 2 * cu, 4 * cu, 0}, {2, 256, 4 * cu, 4 * cu, 0}, {4, 256, 4 * cu, 16 * cu, 0},
        {1, 64, 16 * cu, 4 * cu, 0}, {1, 256, 4 * cu, 0, 1},      {2, 64, 16 * cu, 0, 1},
        {4, 64, 8 * cu, 0, 1},       {1, 64, 8 * cu, 0, 0, 1},   {2, 64, 16 * cu, 0, 0, 1},
    };
    const bool was_timing = e->timing;
    const Tune saved = e->tune;
    double best_res[4] = {1e30, 1e30, 1e30, 1e30}, best_intra = 1e30;
    Tune out = saved;
    cvp9_err_t err = CVP9_OK;
    cvp9_engine_stage_stats(e, nullptr, nullptr, 1);
    for (const Cand &c : cands) {
        for (int k = 0; k < 4; k++) e->tune.res_split[k] = c.split;
        e->tune.intra_lanes = c.lanes;
        e->tune.intra_workers = c.workers;
        e->tune.intra_level_min = c.level_min;
        e->tune.intra_defer = c.defer;
        e->tune.intra_chains = c.chains;
        e->timing = false;
        if ((err = cvp9_engine_run(e, n, streams, preps, nullptr)) != CVP9_OK) break; /* warm */
        e->timing = true;
        for (int i = 0; i < iters && err == CVP9_OK; i++)
            err = cvp9_engine_run(e, n, streams, preps, nullptr);
        if (err != CVP9_OK) break;
        double ms[STAGE_COUNT];
        uint32_t cnt[STAGE_COUNT];
        if ((err = cvp9_engine_stage_stats(e, ms, cnt, 1)) != CVP9_OK) break;
        for (int k = 0; k < 4; k++) {
            const int st = STAGE_RES4 + k;
            const double t = cnt[st] ? ms[st] / cnt[st] : 0.0;
            if (cnt[st] && t < best_res[k]) {
                best_res[k] = t;
                out.res_split[k] = c.split;
            }
        }
        if (cnt[STAGE_INTRA]) {
            const double t = ms[STAGE_INTRA] / cnt[STAGE_INTRA];
            if (t < best_intra) {
                best_intra = t;
                out.intra_lanes = c.lanes;
                out.intra_workers = c.workers;
                out.intra_level_min = c.level_min;
                out.intra_defer = c.defer;
                out.intra_chains = c.chains;
            }
        }
    }
    e->timing = was_timing;
    e->tune = err == CVP9_OK ? out : saved;
    if (err == CVP9_OK && best) {
        for (int k = 0; k < 4; k++) best->res_split[k] = out.res_split[k];
        best->intra_lanes = out.intra_lanes;
        best->intra_workers = out.intra_workers;
        best->intra_level_min = out.intra_level_min;
        best->intra_defer = out.intra_defer;
        best->intra_chains = out.intra_chains;
    }
    return err;
}

cvp9_err_t cvp9_engine_set_timing(cvp9_engine_t *e, int on)
{
    if (!e) return CVP9_ERR_INVALID_DATA;
    e->timing = on != 0;
    return CVP9_OK;
}

cvp9_err_t cvp9_engine_stage_stats(cvp9_engine_t *e, double *ms, uint32_t *launches, int reset)
{
    if (!e) return CVP9_ERR_INVALID_DATA;
    hipSetDevice(e->device);
    if (e->sync_all() != hipSuccess) return CVP9_ERR_GPU;
    e->resolve_timing();
    for (int i = 0; i < STAGE_COUNT; i++) {
        if (ms) ms[i] = e->stage_ms[i];
        if (launches) launches[i] = e->stage_n[i];
        if (reset) {
            e->stage_ms[i] = 0;
            e->stage_n[i] = 0;
        }
    }
    return CVP9_OK;
}

} // extern "C"
