/*
 * fake_hip.cpp — a stand-in for the HIP runtime and for the kernel launch
 * functions, under which the engine's own host code (engine.cpp,
 * engine_api.cpp, compiled unchanged for the host) runs with no GPU and
 * writes the dependency graph of everything it would enqueue.
 *
 * Test infrastructure only: it runs no kernel. Device memory is host memory,
 * every copy and fill happens at once (so the launch functions below can read
 * the batch descriptors the engine uploaded), and each call that would put
 * work on a queue appends one line to the trace named by CVP9_MODEL_TRACE:
 *
 *   K id stream name share acc...   kernel; share: the fraction of the
 *                                    device's resident workgroups it needs
 *                                    all at once (0: it never waits on other
 *                                    workgroups); acc: r:KEY / w:KEY
 *   C id stream acc...               copy or fill
 *   R id stream event                event record
 *   W id stream rec                  wait for record node rec (-1: none)
 *   H id dep...                      the host learnt that nodes dep finished
 *   F key                            host frees allocation key
 *   M text                           a note from the driver
 *
 * A KEY is the base address of the allocation an access falls in. Queue
 * semantics follow HIP's: a queue runs its nodes in order, a wait holds its
 * queue until the record it names has finished, the queues the engine
 * creates are non-blocking (no implicit order with the null queue), and
 * work enqueued after a host synchronisation starts after it. hipFree is
 * taken to wait for the whole device (HIP's documented behaviour for memory
 * a kernel may still use) and is written as an H over every queue's tail
 * followed by the F.
 *
 * The kernels' memory accesses are restated per stage from what each kernel
 * reads and writes (job descriptor fields), with the motion-compensation
 * units' reference slots read from the uploaded unit tables, so that the
 * early and late launches touch exactly the slots their units name.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <set>
#include <string>
#include <vector>

#include "dev_types.h"
#include "kernels.h"

namespace {

FILE *g_out = nullptr;
int g_next = 0;
int g_stream_ids = 1; /* 0: the null queue */
int g_event_ids = 1;
std::map<int, int> g_tail;       /* queue -> its last node */
std::map<int, int> g_rec;        /* event -> node of its latest record */
std::map<uintptr_t, size_t> g_alloc; /* base -> size, device and pinned host */
unsigned g_query_state = 0;      /* query policy (CVP9_MODEL_QUERY) */
int g_query_mode = -1;           /* 0 busy, 1 done, 2 pseudo-random */

FILE *out()
{
    if (!g_out) {
        const char *p = getenv("CVP9_MODEL_TRACE");
        g_out = p ? fopen(p, "w") : stdout;
        if (!g_out) g_out = stdout;
    }
    return g_out;
}

int sid(hipStream_t s) { return (int)(intptr_t)s; }
int eid(hipEvent_t e) { return (int)(intptr_t)e; }

uintptr_t key_of(const void *p)
{
    const uintptr_t a = (uintptr_t)p;
    auto it = g_alloc.upper_bound(a);
    if (it == g_alloc.begin()) return 0;
    --it;
    return a < it->first + (it->second ? it->second : 1) ? it->first : 0;
}

struct Acc {
    std::map<uintptr_t, char> m; /* key -> 'r' or 'w' */
    void r(const void *p)
    {
        const uintptr_t k = key_of(p);
        if (k && !m.count(k)) m[k] = 'r';
    }
    void w(const void *p)
    {
        const uintptr_t k = key_of(p);
        if (k) m[k] = 'w';
    }
    std::string str() const
    {
        std::string s;
        char b[40];
        for (auto &kv : m) {
            snprintf(b, sizeof b, " %c:%lx", kv.second, (unsigned long)kv.first);
            s += b;
        }
        return s;
    }
};

int node(char kind, int stream, const std::string &body)
{
    const int id = g_next++;
    fprintf(out(), "%c %d %d%s\n", kind, id, stream, body.c_str());
    fflush(out());
    g_tail[stream] = id;
    return id;
}

void host_wait(const std::vector<int> &deps)
{
    std::string s;
    char b[16];
    for (int d : deps)
        if (d >= 0) {
            snprintf(b, sizeof b, " %d", d);
            s += b;
        }
    fprintf(out(), "H %d%s\n", g_next++, s.c_str());
    fflush(out());
}

std::vector<int> all_tails()
{
    std::vector<int> v;
    for (auto &kv : g_tail) v.push_back(kv.second);
    return v;
}

void *alloc(size_t n)
{
    void *p = nullptr;
    if (posix_memalign(&p, 256, n ? n : 1) != 0) return nullptr;
    memset(p, 0, n ? n : 1);
    g_alloc[(uintptr_t)p] = n;
    return p;
}

void release(void *p)
{
    if (!p) return;
    host_wait(all_tails());
    fprintf(out(), "F %lx\n", (unsigned long)(uintptr_t)p);
    fflush(out());
    g_alloc.erase((uintptr_t)p);
    free(p);
}

bool query_done()
{
    if (g_query_mode < 0) {
        const char *q = getenv("CVP9_MODEL_QUERY");
        g_query_mode = 0;
        if (q && !strcmp(q, "done")) g_query_mode = 1;
        if (q && !strncmp(q, "random", 6)) {
            g_query_mode = 2;
            g_query_state = (unsigned)atoi(q + 6) * 2654435761u + 1u;
        }
    }
    if (g_query_mode < 2) return g_query_mode == 1;
    g_query_state = g_query_state * 1103515245u + 12345u;
    return (g_query_state >> 16) & 1;
}

/* ---- what each stage reads and writes, per job ---- */

/* reference slots named by the motion-compensation units [a, b) of a job */
unsigned ref_slots(const DJob &J, uint32_t a, uint32_t b)
{
    unsigned m = 0;
    for (uint32_t u = a; u < b; u++) m |= 1u << ((J.mc_tab[u].y >> 16) & 7);
    return m;
}

void mc_access(const DJob &J, bool late, const DBatch &B, Acc &a)
{
    const uint32_t lo = late ? J.n_mc_a : 0, hi = late ? J.n_mc : J.n_mc_a;
    if (lo >= hi) return;
    const unsigned m = ref_slots(J, lo, hi);
    for (int i = 0; i < 8; i++)
        if ((m >> i) & 1 && J.ref[i].y) a.r(J.ref[i].y);
    if (B.resid_all) a.r(J.resid);
    a.w(J.out.y);
}

void stage_access(int st, const DBatch &B, bool late, Acc &a)
{
    a.r(B.jobs);
    for (uint32_t j = 0; j < B.njobs; j++) {
        const DJob &J = B.jobs[j];
        a.r(J.blocks);
        switch (st) {
        case STAGE_EXPAND:
            if (J.n_chunks) a.w(J.coeffs);
            break;
        case STAGE_FILL:
            if (J.fill) {
                a.w(J.out.y);
                a.w(J.owner);
            }
            break;
        case STAGE_OWNER: /* also gives each job its frame epoch */
            a.w(J.owner);
            a.w(J.ectr);
            a.w(B.jobs);
            break;
        case STAGE_RES4:
        case STAGE_RES8:
        case STAGE_RES16:
        case STAGE_RES32:
            if (!J.cls_ntx[st - STAGE_RES4]) break;
            if (B.resid_all) {
                a.w(J.resid);
            } else {
                a.r(J.owner);
                a.w(J.resid);
                a.w(J.out.y);
            }
            break;
        case STAGE_NBR:
            if (J.n_intra) {
                a.r(J.owner);
                a.w(J.nbr);
            }
            break;
        case STAGE_MC:
            mc_access(J, late, B, a);
            break;
        case STAGE_INTRA:
            if (J.n_intra) {
                a.r(J.nbr);
                a.r(J.resid);
                a.r(J.owner);
                a.w(J.edge);
                a.w(J.out.y);
                a.w(J.err);
                a.w(B.jobs); /* ticket counters */
            }
            break;
        case STAGE_LF:
            if (J.lf_thr > 0) a.w(J.out.y);
            break;
        case STAGE_NV12:
            if (J.nv12) {
                a.r(J.out.y);
                a.w(J.nv12);
            }
            break;
        }
    }
}

const char *kStageName[] = {"fill", "owner", "mc", "res4", "res8", "res16", "res32",
                            "nbr",  "intra", "lf", "nv12", "fused", "expand"};

void kernel(const char *name, int stream, double share, const Acc &a)
{
    char b[64];
    snprintf(b, sizeof b, " %s %.6f", name, share);
    node('K', stream, std::string(b) + a.str());
}

void record_done(hipEvent_t *done, hipStream_t s)
{
    if (done && *done) {
        g_rec[eid(*done)] = node('R', sid(s), " " + std::to_string(eid(*done)));
        *done = nullptr;
    }
}

/* The persistent intra workers need every workgroup of their launch
 * resident at once; the launch is sized to half of what the device holds
 * for one engine, divided among the live engines of the process. */
double persistent_share()
{
    const int live = cvp9_live_engines();
    return 1.0 / (2.0 * (live > 1 ? live : 1));
}

} // namespace

/* ---- kernel launch functions of kernels.h ---- */

Tune cvp9_default_tune(int cus)
{
    Tune t;
    memset(&t, 0, sizeof t);
    t.res_split[0] = 1;
    t.res_split[1] = 1;
    t.res_split[2] = 2;
    t.res_split[3] = 4;
    t.intra_lanes = 256;
    t.intra_workers = 4u * (uint32_t)(cus > 0 ? cus : 256);
    return t;
}

hipError_t cvp9_set_idct_tables(const double *, const double *, const double *, const double *)
{
    return hipSuccess;
}

hipError_t cvp9_launch_copy_probe(const void *, void *, size_t, hipStream_t) { return hipErrorNotSupported; }

uint32_t cvp9_fused_max_wgs(int) { return 1024; }

int cvp9_stage_span(int, const DBatch &, const StageCounts &) { return 1; }

hipError_t cvp9_launch_stage(int st, const DBatch &b, const StageCounts &n, hipStream_t s,
                             hipEvent_t *done)
{
    const uint32_t work[] = {n.fill, n.blocks ? n.blocks : 1, n.inter, n.grp[0], n.grp[1], n.grp[2],
                             n.grp[3], n.intra, n.intra, n.lf, n.nv12, 0, n.exp};
    if (st < 0 || st >= STAGE_COUNT || st == STAGE_FUSED) return hipErrorInvalidValue;
    if (!work[st]) return hipSuccess;
    if (st == STAGE_INTRA) {
        /* leading levels: one launch each, no waits; then the persistent tail */
        for (uint32_t l = 0; l < n.levels; l++) {
            Acc a;
            stage_access(st, b, false, a);
            kernel("intra_level", sid(s), 0.0, a);
        }
        if (n.intra_tail) {
            Acc a;
            stage_access(st, b, false, a);
            kernel("intra", sid(s), persistent_share(), a);
        }
        return hipSuccess;
    }
    Acc a;
    stage_access(st, b, false, a);
    kernel(kStageName[st], sid(s), 0.0, a);
    record_done(done, s);
    return hipSuccess;
}

hipError_t cvp9_launch_stages(int st, int span, const DBatch &b, const StageCounts &n, hipStream_t s,
                              hipEvent_t *done)
{
    return span == 1 ? cvp9_launch_stage(st, b, n, s, done) : hipErrorInvalidValue;
}

hipError_t cvp9_launch_mc_late(const DBatch &b, const StageCounts &n, hipStream_t s, hipEvent_t *done)
{
    if (!n.inter2) return hipSuccess;
    Acc a;
    stage_access(STAGE_MC, b, true, a);
    kernel("mc_late", sid(s), 0.0, a);
    record_done(done, s);
    return hipSuccess;
}

/* the fused kernel: every stage, its workgroups all resident at once */
hipError_t cvp9_launch_fused(const DBatch &b, uint32_t, hipStream_t s)
{
    Acc a;
    for (int st = 0; st < STAGE_COUNT; st++)
        if (st != STAGE_FUSED) stage_access(st, b, st == STAGE_MC, a);
    stage_access(STAGE_MC, b, false, a);
    const int live = cvp9_live_engines();
    kernel("fused", sid(s), 1.0 / (live > 1 ? live : 1), a);
    return hipSuccess;
}

hipError_t cvp9_launch_copy_word(const uint32_t *, uint32_t *, hipStream_t) { return hipErrorNotSupported; }

/* ---- the HIP runtime ---- */

extern "C" {

hipError_t hipGetDeviceCount(int *n)
{
    *n = 1;
    return hipSuccess;
}

hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }

hipError_t hipGetDevice(int *d)
{
    *d = 0;
    return hipSuccess;
}

hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_tR0600 *p, int)
{
    memset(p, 0, sizeof *p);
    p->multiProcessorCount = 256;
    p->totalGlobalMem = 288ull << 30;
    strcpy(p->gcnArchName, "gfx950");
    return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned int)
{
    *s = (hipStream_t)(intptr_t)g_stream_ids++;
    return hipSuccess;
}

hipError_t hipExtStreamCreateWithCUMask(hipStream_t *s, uint32_t, const uint32_t *)
{
    return hipStreamCreateWithFlags(s, 0);
}

hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }

hipError_t hipStreamSynchronize(hipStream_t s)
{
    auto it = g_tail.find(sid(s));
    host_wait({it == g_tail.end() ? -1 : it->second});
    return hipSuccess;
}

hipError_t hipDeviceSynchronize(void)
{
    host_wait(all_tails());
    return hipSuccess;
}

hipError_t hipEventCreate(hipEvent_t *e)
{
    *e = (hipEvent_t)(intptr_t)g_event_ids++;
    return hipSuccess;
}

hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }

hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }

hipError_t hipEventRecord(hipEvent_t e, hipStream_t s)
{
    g_rec[eid(e)] = node('R', sid(s), " " + std::to_string(eid(e)));
    return hipSuccess;
}

hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned)
{
    auto it = g_rec.find(eid(e));
    node('W', sid(s), " " + std::to_string(it == g_rec.end() ? -1 : it->second));
    return hipSuccess;
}

hipError_t hipEventSynchronize(hipEvent_t e)
{
    auto it = g_rec.find(eid(e));
    host_wait({it == g_rec.end() ? -1 : it->second});
    return hipSuccess;
}

/* busy, done or a pseudo-random answer by CVP9_MODEL_QUERY; an answer of
 * done is knowledge the host then acts on, so it is an H like a sync */
hipError_t hipEventQuery(hipEvent_t e)
{
    auto it = g_rec.find(eid(e));
    if (it == g_rec.end()) return hipSuccess;
    if (!query_done()) return hipErrorNotReady;
    host_wait({it->second});
    return hipSuccess;
}

hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t)
{
    *ms = 0.f;
    return hipSuccess;
}

hipError_t hipMalloc(void **p, size_t n)
{
    *p = alloc(n);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}

hipError_t hipHostMalloc(void **p, size_t n, unsigned int) { return hipMalloc(p, n); }

hipError_t hipHostGetDevicePointer(void **d, void *h, unsigned int)
{
    *d = h;
    return hipSuccess;
}

hipError_t hipFree(void *p)
{
    release(p);
    return hipSuccess;
}

hipError_t hipHostFree(void *p) { return hipFree(p); }

hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind, hipStream_t s)
{
    memcpy(dst, src, n);
    Acc a;
    a.r(src);
    a.w(dst);
    node('C', sid(s), a.str());
    return hipSuccess;
}

hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind k)
{
    hipMemcpyAsync(dst, src, n, k, nullptr);
    return hipStreamSynchronize(nullptr);
}

hipError_t hipMemcpy2D(void *dst, size_t dp, const void *src, size_t sp, size_t w, size_t h,
                       hipMemcpyKind)
{
    for (size_t r = 0; r < h; r++) memcpy((char *)dst + r * dp, (const char *)src + r * sp, w);
    Acc a;
    a.r(src);
    a.w(dst);
    node('C', 0, a.str());
    return hipStreamSynchronize(nullptr);
}

hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t s)
{
    memset(p, v, n);
    Acc a;
    a.w(p);
    node('C', sid(s), a.str());
    return hipSuccess;
}

hipError_t hipMemset(void *p, int v, size_t n)
{
    hipMemsetAsync(p, v, n, nullptr);
    return hipStreamSynchronize(nullptr);
}

/* graphs are not modelled: capture reports an error */
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *) { return hipErrorNotSupported; }
hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, hipGraphNode_t *, char *, size_t)
{
    return hipErrorNotSupported;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }

} // extern "C"

/* a note in the trace, for the driver */
void cvp9_model_note(const char *text)
{
    fprintf(out(), "M %s\n", text);
    fflush(out());
}
