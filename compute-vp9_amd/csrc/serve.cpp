/*
 * serve.cpp — continuous-batching request loop (cvp9_hip/server.h).
 *
 * Streams are admitted against an HBM budget, frames queue per stream, and
 * pump() launches one batch holding the oldest queued frame of every stream
 * that has one (round-robin when the batch limit is lower than the number
 * of streams), while up to max_inflight earlier batches are still on the
 * GPU. Batches execute in order on the engine's HIP stream, so frames of
 * one stream in consecutive batches keep their order. Completion is seen
 * through one event per batch.
 */
#include <hip/hip_runtime.h>
#include <string.h>
#include <time.h>

#include <deque>
#include <new>
#include <vector>

#include "cvp9_hip/server.h"
#include "engine.hpp"

namespace {

uint64_t now_ns()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

uint64_t nv12_bytes(uint32_t w, uint32_t h)
{
    return (uint64_t)w * h + 2ull * ((w + 1) / 2) * ((h + 1) / 2);
}

struct Pending {
    cvp9_prep_t *prep;
    uint64_t tag;
    uint64_t frame;
};

struct SStream {
    int engine_id = -1;
    bool active = false;
    uint32_t w = 0, h = 0;
    uint64_t cost = 0;
    uint64_t submitted = 0;
    uint8_t *nv12 = nullptr;     /* CVP9_SERVER_NV12_RING buffers back to back */
    uint64_t nv12_cap = 0;       /* bytes per ring entry */
    std::deque<Pending> queue;
};

struct Batch {
    hipEvent_t ev;
    std::vector<uint64_t> tags;
};

} // namespace

struct cvp9_server {
    cvp9_engine_t *e = nullptr;
    cvp9_server_config_t cfg{};
    uint64_t reserved = 0;
    std::vector<SStream> streams;
    std::vector<int> free_handles;
    std::deque<Batch> inflight;
    std::vector<hipEvent_t> events;
    std::deque<std::pair<uint64_t, uint64_t>> done; /* tag, ns */
    uint32_t queued = 0, running = 0;
    size_t rr = 0;

    /* batches may finish out of launch order (the engine runs batches of
     * disjoint streams on two queues); frames of one stream still finish in
     * order, since its batches run in order */
    cvp9_err_t retire()
    {
        for (size_t i = 0; i < inflight.size();) {
            hipError_t q = hipEventQuery(inflight[i].ev);
            if (q == hipErrorNotReady) {
                i++;
                continue;
            }
            if (q != hipSuccess) return CVP9_ERR_GPU;
            const uint64_t t = now_ns();
            for (uint64_t tag : inflight[i].tags) done.emplace_back(tag, t);
            running -= (uint32_t)inflight[i].tags.size();
            events.push_back(inflight[i].ev);
            inflight.erase(inflight.begin() + (long)i);
        }
        return CVP9_OK;
    }
};

extern "C" {

uint64_t cvp9_server_stream_cost(uint32_t w, uint32_t h)
{
    /* frames: KEY frame, previous, current, and the frame the previous one
     * read (not reused while that frame may still run: frame pipelining);
     * two sets (frame pipelining) of owner map, neighbour records (312 B
     * per intra block, 8x8 blocks) and residual scratch; edge records
     * (512 B per block, 8x8 blocks); NV12 ring */
    const uint64_t px = (uint64_t)w * h;
    return 4ull * cvp9::frame_bytes(w, h, nullptr, nullptr) + 2 * (px / 4 + 5 * px + 2 * px) +
           8 * px + CVP9_SERVER_NV12_RING * nv12_bytes(w, h);
}

cvp9_err_t cvp9_server_create(cvp9_engine_t *e, const cvp9_server_config_t *cfg, cvp9_server_t **out)
{
    if (!e || !cfg || !out) return CVP9_ERR_INVALID_DATA;
    if (cfg->max_batch < 1 || cfg->max_batch > CVP9_MAX_BATCH || cfg->max_inflight < 1 ||
        cfg->max_inflight > 4)
        return CVP9_ERR_INVALID_DATA;
    cvp9_server_t *s = new (std::nothrow) cvp9_server();
    if (!s) return CVP9_ERR_NOMEM;
    s->e = e;
    s->cfg = *cfg;
    *out = s;
    return CVP9_OK;
}

void cvp9_server_destroy(cvp9_server_t *s)
{
    if (!s) return;
    cvp9_engine_sync(s->e);
    for (Batch &b : s->inflight) hipEventDestroy(b.ev);
    for (hipEvent_t ev : s->events) hipEventDestroy(ev);
    for (SStream &st : s->streams) {
        if (st.engine_id >= 0) cvp9_engine_stream_close(s->e, st.engine_id);
        hipFree(st.nv12);
    }
    delete s;
}

cvp9_err_t cvp9_server_open(cvp9_server_t *s, uint32_t w, uint32_t h, int *stream)
{
    if (!s || !stream || w < 2 || h < 2) return CVP9_ERR_INVALID_DATA;
    const uint64_t cost = cvp9_server_stream_cost(w, h);
    if (s->reserved + cost > s->cfg.hbm_budget) return CVP9_ERR_AGAIN;
    int hnd = -1;
    /* recycle a closed stream whose NV12 ring is large enough */
    for (size_t i = 0; i < s->free_handles.size(); i++) {
        SStream &c = s->streams[s->free_handles[i]];
        if (c.nv12_cap >= nv12_bytes(w, h)) {
            hnd = s->free_handles[i];
            s->free_handles.erase(s->free_handles.begin() + i);
            break;
        }
    }
    if (hnd < 0) {
        SStream st;
        cvp9_err_t err = cvp9_engine_stream_open(s->e, &st.engine_id);
        if (err != CVP9_OK) return err;
        st.nv12_cap = (nv12_bytes(w, h) + 255) & ~255ull;
        if (hipMalloc((void **)&st.nv12, CVP9_SERVER_NV12_RING * st.nv12_cap) != hipSuccess) {
            cvp9_engine_stream_close(s->e, st.engine_id);
            return CVP9_ERR_NOMEM;
        }
        s->streams.push_back(std::move(st));
        hnd = (int)s->streams.size() - 1;
    } else {
        cvp9_err_t err = cvp9_engine_stream_reset(s->e, s->streams[hnd].engine_id);
        if (err != CVP9_OK) return err;
    }
    SStream &st = s->streams[hnd];
    st.active = true;
    st.w = w;
    st.h = h;
    st.cost = cost;
    st.submitted = 0;
    st.queue.clear();
    s->reserved += cost;
    *stream = hnd;
    return CVP9_OK;
}

void cvp9_server_close(cvp9_server_t *s, int stream)
{
    if (!s || stream < 0 || stream >= (int)s->streams.size() || !s->streams[stream].active) return;
    SStream &st = s->streams[stream];
    s->queued -= (uint32_t)st.queue.size();
    st.queue.clear();
    st.active = false;
    s->reserved -= st.cost;
    s->free_handles.push_back(stream);
}

cvp9_err_t cvp9_server_submit(cvp9_server_t *s, int stream, cvp9_prep_t *prep, uint64_t tag)
{
    if (!s || !prep || stream < 0 || stream >= (int)s->streams.size() || !s->streams[stream].active)
        return CVP9_ERR_INVALID_DATA;
    SStream &st = s->streams[stream];
    st.queue.push_back(Pending{prep, tag, st.submitted++});
    s->queued++;
    return CVP9_OK;
}

cvp9_err_t cvp9_server_submit_n(cvp9_server_t *s, int n, const int *streams,
                                cvp9_prep_t *const *preps, const uint64_t *tags)
{
    if (!s || n < 0 || (n && (!streams || !preps || !tags))) return CVP9_ERR_INVALID_DATA;
    for (int i = 0; i < n; i++) {
        const cvp9_err_t err = cvp9_server_submit(s, streams[i], preps[i], tags[i]);
        if (err != CVP9_OK) return err;
    }
    return CVP9_OK;
}

int cvp9_server_pump(cvp9_server_t *s)
{
    if (!s) return CVP9_ERR_INVALID_DATA;
    cvp9_err_t err = s->retire();
    if (err != CVP9_OK) return err;
    if (s->inflight.size() >= s->cfg.max_inflight || s->queued == 0) return 0;
    std::vector<int> ids;
    std::vector<cvp9_prep_t *> preps;
    std::vector<uint8_t *> outs;
    Batch b;
    const size_t n = s->streams.size();
    size_t last = s->rr;
    for (size_t k = 0; k < n && ids.size() < s->cfg.max_batch; k++) {
        const size_t i = (s->rr + k) % n;
        SStream &st = s->streams[i];
        if (!st.active || st.queue.empty()) continue;
        const Pending p = st.queue.front();
        st.queue.pop_front();
        ids.push_back(st.engine_id);
        preps.push_back(p.prep);
        outs.push_back(st.nv12 + (p.frame % CVP9_SERVER_NV12_RING) * st.nv12_cap);
        b.tags.push_back(p.tag);
        last = i;
    }
    s->rr = (last + 1) % (n ? n : 1);
    if (ids.empty()) return 0;
    err = cvp9_engine_run(s->e, (int)ids.size(), ids.data(), preps.data(), outs.data());
    if (err != CVP9_OK) return err;
    if (s->events.empty()) {
        hipEvent_t ev;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return CVP9_ERR_GPU;
        s->events.push_back(ev);
    }
    b.ev = s->events.back();
    s->events.pop_back();
    if (hipEventRecord(b.ev, (hipStream_t)cvp9_engine_hip_stream(s->e)) != hipSuccess) return CVP9_ERR_GPU;
    s->queued -= (uint32_t)ids.size();
    s->running += (uint32_t)ids.size();
    s->inflight.push_back(std::move(b));
    return (int)ids.size();
}

int cvp9_server_poll(cvp9_server_t *s, uint64_t *tags, uint64_t *done_ns, int max)
{
    if (!s || max < 0 || (max > 0 && (!tags || !done_ns))) return CVP9_ERR_INVALID_DATA;
    cvp9_err_t err = s->retire();
    if (err != CVP9_OK) return err;
    int k = 0;
    while (k < max && !s->done.empty()) {
        tags[k] = s->done.front().first;
        done_ns[k] = s->done.front().second;
        s->done.pop_front();
        k++;
    }
    return k;
}

uint32_t cvp9_server_pending(cvp9_server_t *s)
{
    return s ? s->queued + s->running : 0;
}

cvp9_err_t cvp9_server_read_nv12(cvp9_server_t *s, int stream, uint64_t frame, uint8_t *dst, uint64_t cap)
{
    if (!s || !dst || stream < 0 || stream >= (int)s->streams.size()) return CVP9_ERR_INVALID_DATA;
    const SStream &st = s->streams[stream];
    const uint64_t need = nv12_bytes(st.w, st.h);
    if (cap < need || !st.nv12) return CVP9_ERR_INVALID_DATA;
    cvp9_err_t err = cvp9_engine_sync(s->e);
    if (err != CVP9_OK) return err;
    const uint8_t *src = st.nv12 + (frame % CVP9_SERVER_NV12_RING) * st.nv12_cap;
    return hipMemcpy(dst, src, need, hipMemcpyDeviceToHost) == hipSuccess ? CVP9_OK : CVP9_ERR_GPU;
}

cvp9_err_t cvp9_server_drain(cvp9_server_t *s)
{
    if (!s) return CVP9_ERR_INVALID_DATA;
    while (s->queued) {
        int r = cvp9_server_pump(s);
        if (r < 0) return (cvp9_err_t)r;
        if (r == 0) {
            cvp9_err_t err = cvp9_engine_sync(s->e);
            if (err != CVP9_OK) return err;
        }
    }
    cvp9_err_t err = cvp9_engine_sync(s->e);
    cvp9_err_t r2 = s->retire();
    return err != CVP9_OK ? err : r2;
}

} // extern "C"
