/*
 * backend.cpp — hip_backend_* C ABI (cvp9_hip/hip_backend.h).
 *
 * Mirrors the Vulkan backend's contract (src/backend/vulkan/vulkan_backend.c:
 * ring :1037-1062, :1391-1426; get :1428-1455; DMA-BUF :1457-1486; export
 * buffers :1488-1544) on a one-stream engine (engine_api.cpp: frame
 * pipelining, shared launches, the stream's reference pool):
 *   decode: validate + stage (pinned) -> engine batch (H2D on its queue,
 *           stage kernels, reference-pool update) -> pack to the slot's VRAM
 *           I420 buffer -> D2H into the slot's pinned output -> event, on
 *           the queue the batch finishes on
 *   get:    poll / wait the oldest slot's event, hand out pinned pointers.
 */
#include <hip/hip_runtime.h>
#include <hsa/hsa_ext_amd.h>
#include <sys/mman.h>
#include <unistd.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <vector>

#include "cvp9_hip/hip_backend.h"
#include "cvp9_hip/testing.h"
#include "engine.hpp"

using namespace cvp9;

namespace {


struct Slot {
    hipEvent_t done = nullptr;
    bool pending = false;
    int64_t pts = 0;
    uint32_t w = 0, h = 0;
    uint8_t *stage_host = nullptr;  size_t stage_host_cap = 0;
    uint8_t *stage_dev = nullptr;   size_t stage_dev_cap = 0;
    uint8_t *out_host = nullptr;    size_t out_host_cap = 0;
    uint8_t *out_dev = nullptr;     size_t out_dev_cap = 0;   /* packed I420, VRAM */
    int out_fd = -1;                /* its DMA-BUF, or -1 */
    uint64_t out_off = 0;           /* out_dev's offset inside that DMA-BUF */
    uint32_t *err_host = nullptr;   /* the stream's error word after this frame */
    uint32_t *err_dev = nullptr;    /* its device view */
};

struct Backend {
    int device = 0;
    cvp9_engine_t *eng = nullptr;
    int sid = -1;                       /* the engine stream this context decodes */
    Slot slots[CVP9_HIP_INFLIGHT];
    uint32_t head = 0, count = 0;
    uint8_t *pending_target = nullptr;   /* device address of the armed NV12 target */
    uint64_t pending_target_size = 0;
    struct ExportPriv *pending_priv = nullptr;
    uint32_t debug = 0;
    bool dmabuf = false;
    /* A frame was declined (CVP9_ERR_UNSUPPORTED): the caller reconstructed
     * it on the CPU, so the device reference pool lacks it. Every frame is
     * declined until the next KEY frame, which writes all 8 slots. */
    bool declined = false;
    /* Coefficients staged compact (nonzero masks and values, expanded on
     * the device) instead of verbatim: about a third of the PCIe bytes at
     * 25% nonzero coefficients, for more host time per decode call (the
     * values are packed on the host). Chosen per context at init from the
     * environment: CVP9_HIP_STAGING=compact. */
    bool compact = false;
};

/* An export buffer: dedicated VRAM with its DMA-BUF, and the CPU view the
 * caller gets as `mapped`: the DMA-BUF mapped into this process where the
 * driver allows it, else pinned host memory that every decode into the
 * buffer refreshes on the stream before the frame is reported done. */
struct ExportPriv {
    uint8_t *dev;
    size_t size;
    void *map;
    bool shadow;   /* map is a hipHostMalloc copy, not a mapping of dev */
};

inline size_t i420_bytes(uint32_t w, uint32_t h)
{
    return (size_t)w * h + 2 * (size_t)((w + 1) / 2) * ((h + 1) / 2);
}

/* DMA-BUF fd of a VRAM allocation and ptr's offset inside it, or -1 */
int export_fd(void *ptr, size_t size, uint64_t *off)
{
    int fd = -1;
    *off = 0;
    if (hsa_amd_portable_export_dmabuf(ptr, size, &fd, off) != HSA_STATUS_SUCCESS) return -1;
    return fd;
}

/* Device memory for a buffer handed out as a DMA-BUF: whole 2 MiB units, so
 * the allocator gives it an allocation of its own (smaller requests can
 * share one, and then export at an offset). */
hipError_t alloc_exportable(uint8_t **p, size_t *cap, size_t need)
{
    const size_t c = (need + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
    hipError_t e = hipMalloc((void **)p, c);
    *cap = e == hipSuccess ? c : 0;
    return e;
}

hipError_t grow_pinned(uint8_t **p, size_t *cap, size_t need)
{
    if (*cap >= need) return hipSuccess;
    if (*p) hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t c = need + need / 4;
    hipError_t e = hipHostMalloc((void **)p, c, hipHostMallocDefault);
    if (e == hipSuccess) *cap = c;
    return e;
}

hipError_t grow_device(uint8_t **p, size_t *cap, size_t need)
{
    if (*cap >= need) return hipSuccess;
    if (*p) hipFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t c = need + need / 4;
    hipError_t e = hipMalloc((void **)p, c);
    if (e == hipSuccess) *cap = c;
    return e;
}

cvp9_err_t to_err(hipError_t e)
{
    if (e == hipSuccess) return CVP9_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return CVP9_ERR_NOMEM;
    return CVP9_ERR_GPU;
}

void free_slot(Slot &s)
{
    if (s.done) hipEventDestroy(s.done);
    if (s.stage_host) hipHostFree(s.stage_host);
    if (s.stage_dev) hipFree(s.stage_dev);
    if (s.out_host) hipHostFree(s.out_host);
    if (s.out_fd >= 0) hsa_amd_portable_close_dmabuf(s.out_fd);
    if (s.out_dev) hipFree(s.out_dev);
    if (s.err_host) hipHostFree(s.err_host);
    s = Slot();
}

/* The backend's decision for one frame before any device work: CVP9_OK
 * (take it), CVP9_ERR_UNSUPPORTED (the caller's CPU path reconstructs it)
 * or CVP9_ERR_INVALID_DATA. A frame the validator declines sets `declined`;
 * from then on every frame is declined until a KEY frame (which writes all
 * 8 reference slots, so the device pool is whole again). */
cvp9_err_t route(bool &declined, const vp9_parsed_frame_t *pf, PrepLayout &L)
{
    cvp9_err_t err = plan_measure(pf, L);
    if (err == CVP9_ERR_UNSUPPORTED) declined = true;
    if (err != CVP9_OK) return err;
    if (declined && pf->hdr.frame_type != VP9_FRAME_KEY) return CVP9_ERR_UNSUPPORTED;
    declined = false;
    return CVP9_OK;
}

/* oldest in-flight frame, released from the ring (vulkan_backend.c:1391-1426) */
cvp9_err_t acquire_oldest(Backend *b, int wait, Slot **out)
{
    if (!b || b->count == 0) return CVP9_ERR_UNSUPPORTED;
    Slot &s = b->slots[b->head];
    if (!wait && b->count < CVP9_HIP_INFLIGHT) {
        hipError_t q = hipEventQuery(s.done);
        if (q == hipErrorNotReady) return CVP9_ERR_AGAIN;
        if (q != hipSuccess) return CVP9_ERR_GPU;
    } else {
        if (hipEventSynchronize(s.done) != hipSuccess) return CVP9_ERR_GPU;
    }
    s.pending = false;
    b->head = (b->head + 1) % CVP9_HIP_INFLIGHT;
    b->count--;
    *out = &s;
    return CVP9_OK;
}

} // namespace

extern "C" {

cvp9_err_t hip_backend_init(void **ctx)
{
    if (!ctx) return CVP9_ERR_INVALID_DATA;
    *ctx = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return CVP9_ERR_NO_BACKEND;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return CVP9_ERR_NO_BACKEND;
    if (init_device(dev) != hipSuccess) return CVP9_ERR_NO_BACKEND;
    Backend *b = new (std::nothrow) Backend();
    if (!b) return CVP9_ERR_NOMEM;
    b->device = dev;
    const char *staging = getenv("CVP9_HIP_STAGING");
    b->compact = staging && strcmp(staging, "compact") == 0;
    cvp9_err_t ce = cvp9_engine_create(dev, &b->eng);
    if (ce == CVP9_OK) ce = cvp9_engine_stream_open(b->eng, &b->sid);
    if (ce != CVP9_OK) {
        hip_backend_destroy(b);
        return ce;
    }
    for (Slot &s : b->slots) {
        if (hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void **)&s.err_host, 64, hipHostMallocCoherent) != hipSuccess ||
            hipHostGetDevicePointer((void **)&s.err_dev, s.err_host, 0) != hipSuccess) {
            hip_backend_destroy(b);
            return CVP9_ERR_GPU;
        }
        *s.err_host = 0;
    }
    /* DMA-BUF export works on VRAM allocations when the driver supports it */
    {
        void *probe = nullptr;
        if (hipMalloc(&probe, 4096) == hipSuccess) {
            uint64_t off = 0;
            int fd = export_fd(probe, 4096, &off);
            if (fd >= 0) {
                b->dmabuf = true;
                hsa_amd_portable_close_dmabuf(fd);
            }
            hipFree(probe);
        }
    }
    *ctx = b;
    return CVP9_OK;
}

void hip_backend_destroy(void *ctx)
{
    Backend *b = (Backend *)ctx;
    if (!b) return;
    hipSetDevice(b->device);
    if (b->eng) cvp9_engine_sync(b->eng);
    for (Slot &s : b->slots) free_slot(s);
    cvp9_engine_destroy(b->eng);
    delete b;
}

cvp9_err_t hip_decode_frame(void *ctx, const vp9_parsed_frame_t *pf, int64_t pts)
{
    Backend *b = (Backend *)ctx;
    if (!b || !pf) return CVP9_ERR_INVALID_DATA;
    if (b->count == CVP9_HIP_INFLIGHT) return CVP9_ERR_AGAIN;
    const bool key = pf->hdr.frame_type == VP9_FRAME_KEY;
    PrepLayout L;
    L.compact = b->compact;
    cvp9_err_t err = route(b->declined, pf, L);
    if (err != CVP9_OK) return err;
    /* an armed NV12 target must hold the whole frame */
    if (b->pending_target && b->pending_target_size < i420_bytes(L.W, L.H)) return CVP9_ERR_INVALID_DATA;
    hipSetDevice(b->device);
    /* After a device error (reported for that frame and every later one)
     * a KEY frame starts the stream afresh: frames still in flight finish
     * first, then the pool is emptied and the error cleared. */
    if (key && cvp9_engine_stream_error(b->eng, b->sid)) {
        if ((err = cvp9_engine_stream_reset(b->eng, b->sid)) != CVP9_OK) return err;
    }

    Slot &s = b->slots[(b->head + b->count) % CVP9_HIP_INFLIGHT];
    const size_t need = L.bytes;
    hipError_t e;
    if ((e = grow_pinned(&s.stage_host, &s.stage_host_cap, L.host_bytes)) != hipSuccess) return to_err(e);
    if ((e = grow_device(&s.stage_dev, &s.stage_dev_cap, need)) != hipSuccess) return to_err(e);
    const size_t ob = i420_bytes(L.W, L.H);
    if ((e = grow_pinned(&s.out_host, &s.out_host_cap, ob)) != hipSuccess) return to_err(e);
    if (s.out_dev_cap < ob) {
        if (s.out_fd >= 0) hsa_amd_portable_close_dmabuf(s.out_fd);
        s.out_fd = -1;
        if (s.out_dev) hipFree(s.out_dev);
        s.out_dev = nullptr;
        if ((e = alloc_exportable(&s.out_dev, &s.out_dev_cap, ob)) != hipSuccess) return to_err(e);
        if (b->dmabuf) s.out_fd = export_fd(s.out_dev, s.out_dev_cap, &s.out_off);
    }

    plan_write(pf, L, s.stage_host);
    if (b->debug & CVP9_HIP_DEBUG_NO_DEBLOCK) L.lf_thr = 0;
    StagedResult r;
    if ((err = engine_run_staged(b->eng, b->sid, L, s.stage_host, s.stage_dev, b->pending_target,
                                 &r)) != CVP9_OK)
        return err;
    const hipStream_t q = r.queue;

    /* packed I420 into the slot's VRAM buffer, then to the pinned output */
    const DFrame f = r.out;
    const uint32_t cw = (L.W + 1) / 2, ch = (L.H + 1) / 2;
    uint8_t *dy = s.out_dev, *du = dy + (size_t)L.W * L.H, *dv = du + (size_t)cw * ch;
    if ((e = hipMemcpy2DAsync(dy, L.W, f.y, f.sy, L.W, L.H, hipMemcpyDeviceToDevice, q)) ||
        (e = hipMemcpy2DAsync(du, cw, f.u, f.suv, cw, ch, hipMemcpyDeviceToDevice, q)) ||
        (e = hipMemcpy2DAsync(dv, cw, f.v, f.suv, cw, ch, hipMemcpyDeviceToDevice, q)) ||
        (e = hipMemcpyAsync(s.out_host, s.out_dev, ob, hipMemcpyDeviceToHost, q)) ||
        (b->pending_priv && b->pending_priv->shadow &&
         (e = hipMemcpyAsync(b->pending_priv->map, b->pending_priv->dev, ob, hipMemcpyDeviceToHost,
                             q))) ||
        (e = cvp9_launch_copy_word(r.err, s.err_dev, q)) ||
        (e = hipEventRecord(s.done, q)))
        return to_err(e);

    s.pending = true;
    s.pts = pts;
    s.w = L.W;
    s.h = L.H;
    b->count++;
    b->pending_target = nullptr; /* one-shot (vulkan_backend.c:1059-1062) */
    b->pending_target_size = 0;
    b->pending_priv = nullptr;
    return CVP9_OK;
}

cvp9_err_t hip_get_frame(void *ctx, cvp9_frame_info_t *info, int wait)
{
    Backend *b = (Backend *)ctx;
    if (!info) return CVP9_ERR_INVALID_DATA;
    if (b) hipSetDevice(b->device);
    Slot *s = nullptr;
    cvp9_err_t err = acquire_oldest(b, wait, &s);
    if (err != CVP9_OK) return err;
    if (__atomic_load_n(s->err_host, __ATOMIC_ACQUIRE)) return CVP9_ERR_GPU;
    const uint32_t w = s->w, h = s->h, cw = (w + 1) / 2, ch = (h + 1) / 2;
    info->width = w;
    info->height = h;
    info->stride_y = w;
    info->stride_uv = cw;
    info->plane_y = s->out_host;
    info->plane_u = s->out_host + (size_t)w * h;
    info->plane_v = info->plane_u + (size_t)cw * ch;
    info->pts = s->pts;
    return CVP9_OK;
}

cvp9_err_t hip_get_frame_dmabuf(void *ctx, cvp9_dmabuf_frame_t *out)
{
    Backend *b = (Backend *)ctx;
    if (!b || !b->dmabuf) return CVP9_ERR_UNSUPPORTED;
    if (!out) return CVP9_ERR_INVALID_DATA;
    hipSetDevice(b->device);
    Slot *s = nullptr;
    cvp9_err_t err = acquire_oldest(b, 1, &s);
    if (err != CVP9_OK) return err;
    if (__atomic_load_n(s->err_host, __ATOMIC_ACQUIRE)) return CVP9_ERR_GPU;
    if (s->out_fd < 0) return CVP9_ERR_UNSUPPORTED;
    const uint32_t w = s->w, h = s->h, cw = (w + 1) / 2, ch = (h + 1) / 2;
    out->fd = dup(s->out_fd);
    if (out->fd < 0) return CVP9_ERR_GPU;
    const uint32_t o = (uint32_t)s->out_off;
    out->size = s->out_off + s->out_dev_cap;
    out->width = w;
    out->height = h;
    out->offsets[0] = o;
    out->offsets[1] = o + w * h;
    out->offsets[2] = o + w * h + cw * ch;
    out->pitches[0] = w;
    out->pitches[1] = cw;
    out->pitches[2] = cw;
    out->pts = s->pts;
    return CVP9_OK;
}

cvp9_err_t hip_export_buffer_alloc(void *ctx, uint64_t size, cvp9_export_buffer_t *out)
{
    Backend *b = (Backend *)ctx;
    if (!b || !out) return CVP9_ERR_INVALID_DATA;
    if (!b->dmabuf || size == 0) return CVP9_ERR_UNSUPPORTED;
    hipSetDevice(b->device);
    uint8_t *dev = nullptr;
    size_t sz = 0;
    if (alloc_exportable(&dev, &sz, size) != hipSuccess) return CVP9_ERR_GPU;
    uint64_t off = 0;
    int fd = export_fd(dev, sz, &off);
    if (fd < 0 || off != 0) { /* the buffer must start the DMA-BUF: no offset field */
        if (fd >= 0) hsa_amd_portable_close_dmabuf(fd);
        hipFree(dev);
        return CVP9_ERR_UNSUPPORTED;
    }
    /* CPU view: the DMA-BUF mapped (BAR), else a pinned shadow */
    bool shadow = false;
    void *map = mmap(nullptr, sz, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
    if (map == MAP_FAILED) {
        map = nullptr;
        if (hipHostMalloc(&map, sz, hipHostMallocDefault) != hipSuccess) {
            hsa_amd_portable_close_dmabuf(fd);
            hipFree(dev);
            return CVP9_ERR_NOMEM;
        }
        memset(map, 0, sz);
        shadow = true;
    }
    ExportPriv *p = new (std::nothrow) ExportPriv{dev, sz, map, shadow};
    if (!p) {
        if (shadow) hipHostFree(map); else munmap(map, sz);
        hsa_amd_portable_close_dmabuf(fd);
        hipFree(dev);
        return CVP9_ERR_NOMEM;
    }
    out->fd = fd;
    out->mapped = map;
    out->size = size;
    out->priv_buf = p;
    out->priv_mem = dev;
    return CVP9_OK;
}

void hip_export_buffer_free(void *ctx, cvp9_export_buffer_t *buf)
{
    Backend *b = (Backend *)ctx;
    if (!b || !buf || !buf->priv_buf) return;
    hipSetDevice(b->device);
    cvp9_engine_sync(b->eng);
    ExportPriv *p = (ExportPriv *)buf->priv_buf;
    if (b->pending_priv == p) {
        b->pending_target = nullptr;
        b->pending_target_size = 0;
        b->pending_priv = nullptr;
    }
    if (p->shadow) hipHostFree(p->map); else munmap(p->map, p->size);
    if (buf->fd >= 0) hsa_amd_portable_close_dmabuf(buf->fd);
    hipFree(p->dev);
    delete p;
    memset(buf, 0, sizeof(*buf));
    buf->fd = -1;
}

cvp9_err_t hip_set_render_target(void *ctx, const cvp9_export_buffer_t *target)
{
    Backend *b = (Backend *)ctx;
    if (!b) return CVP9_ERR_INVALID_DATA;
    if (!target) {
        b->pending_target = nullptr;
        b->pending_target_size = 0;
        b->pending_priv = nullptr;
        return CVP9_OK;
    }
    if (!b->dmabuf || !target->priv_buf) return CVP9_ERR_UNSUPPORTED;
    const ExportPriv *p = (const ExportPriv *)target->priv_buf;
    b->pending_target = p->dev;
    b->pending_target_size = target->size;
    b->pending_priv = (ExportPriv *)target->priv_buf;
    return CVP9_OK;
}

cvp9_err_t hip_backend_read_ref(void *ctx, int slot, cvp9_frame_info_t *frame)
{
    Backend *b = (Backend *)ctx;
    if (!b || !frame || slot < 0 || slot > 7) return CVP9_ERR_INVALID_DATA;
    hipSetDevice(b->device);
    cvp9_err_t err = cvp9_engine_sync(b->eng);
    if (err != CVP9_OK) return err;
    const DFrame f = engine_stream_ref(b->eng, b->sid, slot);
    if (!f.y) {
        frame->width = frame->height = 0;
        return CVP9_OK;
    }
    if (!frame->plane_y) {
        frame->width = f.w;
        frame->height = f.h;
        return CVP9_OK;
    }
    const uint32_t cw = (f.w + 1) / 2, ch = (f.h + 1) / 2;
    if (frame->width != f.w || frame->height != f.h || !frame->plane_u || !frame->plane_v ||
        frame->stride_y < f.w || frame->stride_uv < cw)
        return CVP9_ERR_INVALID_DATA;
    if (hipMemcpy2D(frame->plane_y, frame->stride_y, f.y, f.sy, f.w, f.h, hipMemcpyDeviceToHost) ||
        hipMemcpy2D(frame->plane_u, frame->stride_uv, f.u, f.suv, cw, ch, hipMemcpyDeviceToHost) ||
        hipMemcpy2D(frame->plane_v, frame->stride_uv, f.v, f.suv, cw, ch, hipMemcpyDeviceToHost))
        return CVP9_ERR_GPU;
    return CVP9_OK;
}

/* ---- testing hooks (cvp9_hip/testing.h) --------------------------------- */

const char *cvp9_hip_version(void) { return "cvp9_hip 0.1.0 (gfx950)"; }

int cvp9_hip_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int cvp9_hip_idct_table(int n, double *out)
{
    if (!out || (n != 4 && n != 8 && n != 16 && n != 32)) return -1;
    idct_matrix(n, out);
    return 0;
}

cvp9_err_t cvp9_hip_plan_check(const vp9_parsed_frame_t *pf)
{
    PrepLayout L;
    return plan_measure(pf, L);
}

int64_t cvp9_hip_plan_image(const vp9_parsed_frame_t *pf, int threads, uint8_t *dst, uint64_t cap)
{
    PrepLayout L;
    cvp9_err_t err = plan_measure(pf, L, threads);
    if (err != CVP9_OK) return err;
    if (dst && cap >= L.bytes) {
        memset(dst, 0, L.bytes); /* the alignment gaps between the tables */
        plan_write(pf, L, dst);
    }
    return (int64_t)L.bytes;
}

int64_t cvp9_hip_plan_image_compact(const vp9_parsed_frame_t *pf, int threads, uint8_t *dst,
                                    uint64_t cap, uint64_t info[8])
{
    PrepLayout L;
    L.compact = true;
    cvp9_err_t err = plan_measure(pf, L, threads);
    if (err != CVP9_OK) return err;
    if (dst && cap >= L.host_bytes) {
        memset(dst, 0, L.host_bytes);
        plan_write(pf, L, dst);
    }
    if (info) {
        info[0] = L.off_cmask;
        info[1] = L.off_cstart;
        info[2] = L.off_cvals;
        info[3] = L.nnz;
        info[4] = L.n_chunks;
        info[5] = L.upload;
        info[6] = L.off_coeffs;
        info[7] = L.bytes;
    }
    return (int64_t)L.host_bytes;
}

int cvp9_hip_intra_chains(const vp9_parsed_frame_t *pf, uint32_t *out, uint32_t cap)
{
    PrepLayout L;
    L.chains = true;
    cvp9_err_t err = plan_measure(pf, L);
    if (err != CVP9_OK) return err;
    std::vector<uint8_t> img(L.bytes);
    plan_write(pf, L, img.data());
    const uint32_t n = L.n_chains + 1 + L.n_intra;
    if (out) {
        if (cap < n) return CVP9_ERR_INVALID_DATA;
        memcpy(out, img.data() + L.off_chain, (size_t)n * 4);
    }
    return (int)L.n_chains;
}

int cvp9_hip_intra_order(const vp9_parsed_frame_t *pf, uint32_t *order, uint32_t cap)
{
    PrepLayout L;
    cvp9_err_t err = plan_measure(pf, L);
    if (err != CVP9_OK) return err;
    if (L.n_intra > cap || (L.n_intra && !order)) return CVP9_ERR_INVALID_DATA;
    std::vector<uint8_t> img(L.bytes);
    plan_write(pf, L, img.data());
    /* device list index -> index in pf->blocks (blocks clamped away are not listed) */
    std::vector<uint32_t> src;
    src.reserve(L.n_blocks);
    for (uint32_t i = 0; i < pf->num_blocks; i++) {
        const vp9_macroblock_info_t &b = pf->blocks[i];
        if (b.x < pf->hdr.width && b.y < pf->hdr.height) src.push_back(i);
    }
    const uint32_t *intra = (const uint32_t *)(img.data() + L.off_intra);
    for (uint32_t i = 0; i < L.n_intra; i++) order[i] = src[intra[i]];
    return (int)L.n_intra;
}

int cvp9_hip_intra_levels(const vp9_parsed_frame_t *pf, uint32_t *starts, uint32_t cap)
{
    PrepLayout L;
    cvp9_err_t err = plan_measure(pf, L);
    if (err != CVP9_OK) return err;
    std::vector<uint8_t> img(L.bytes);
    plan_write(pf, L, img.data());
    if (L.levels.empty()) return 0;
    const uint32_t depth = (uint32_t)L.levels.size() - 1;
    if (starts) {
        if (cap < depth + 1) return CVP9_ERR_INVALID_DATA;
        memcpy(starts, L.levels.data(), (size_t)(depth + 1) * 4);
    }
    return (int)depth;
}

cvp9_err_t cvp9_hip_export_read(void *ctx, const cvp9_export_buffer_t *buf, uint8_t *dst,
                                uint64_t n)
{
    Backend *b = (Backend *)ctx;
    if (!b || !buf || !buf->priv_buf || !dst) return CVP9_ERR_INVALID_DATA;
    const ExportPriv *p = (const ExportPriv *)buf->priv_buf;
    if (n > p->size) return CVP9_ERR_INVALID_DATA;
    hipSetDevice(b->device);
    if (cvp9_engine_sync(b->eng) != CVP9_OK) return CVP9_ERR_GPU;
    return hipMemcpy(dst, p->dev, n, hipMemcpyDeviceToHost) == hipSuccess ? CVP9_OK : CVP9_ERR_GPU;
}

cvp9_err_t cvp9_hip_dmabuf_read(int fd, uint64_t size, uint64_t offset, uint8_t *dst, uint64_t n)
{
    if (fd < 0 || !dst || offset + n > size) return CVP9_ERR_INVALID_DATA;
    /* the import takes ownership of the descriptor it is given */
    const int own = dup(fd);
    if (own < 0) return CVP9_ERR_GPU;
    hipExternalMemoryHandleDesc d;
    memset(&d, 0, sizeof(d));
    d.type = hipExternalMemoryHandleTypeOpaqueFd;
    d.handle.fd = own;
    d.size = size;
    hipExternalMemory_t ext = nullptr;
    if (hipImportExternalMemory(&ext, &d) != hipSuccess) {
        close(own);
        return CVP9_ERR_UNSUPPORTED;
    }
    hipExternalMemoryBufferDesc bd;
    memset(&bd, 0, sizeof(bd));
    bd.offset = 0;
    bd.size = size;
    void *ptr = nullptr;
    cvp9_err_t err = CVP9_OK;
    if (hipExternalMemoryGetMappedBuffer(&ptr, ext, &bd) != hipSuccess)
        err = CVP9_ERR_UNSUPPORTED;
    else if (hipMemcpy(dst, (uint8_t *)ptr + offset, n, hipMemcpyDeviceToHost) != hipSuccess)
        err = CVP9_ERR_GPU;
    if (ptr) hipFree(ptr);
    hipDestroyExternalMemory(ext);
    return err;
}

double cvp9_hip_copy_probe(uint64_t bytes, int iters)
{
    if (bytes < 16 || (bytes & 15) || iters < 1) return -1.0;
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double gbps = -1.0;
    if (hipMalloc(&a, bytes) == hipSuccess && hipMalloc(&b, bytes) == hipSuccess &&
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
        hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
        hipMemsetAsync(a, 1, bytes, s) == hipSuccess &&
        cvp9_launch_copy_probe(a, b, bytes, s) == hipSuccess &&
        hipEventRecord(e0, s) == hipSuccess) {
        bool ok = true;
        for (int i = 0; i < iters && ok; i++) ok = cvp9_launch_copy_probe(a, b, bytes, s) == hipSuccess;
        float ms = 0.f;
        if (ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
            hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0.f)
            gbps = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (s) hipStreamDestroy(s);
    hipFree(a);
    hipFree(b);
    return gbps;
}

cvp9_err_t hip_backend_set_debug(void *ctx, uint32_t flags)
{
    Backend *b = (Backend *)ctx;
    if (!b) return CVP9_ERR_INVALID_DATA;
    b->debug = flags;
    /* dependency waits give up after one poll (0: the default bound) */
    return cvp9_engine_set_spin_limit(b->eng, (flags & CVP9_HIP_DEBUG_SPIN_ONE) ? 1u : 0u);
}

void *cvp9_hip_dev_alloc(uint64_t bytes)
{
    void *p = nullptr;
    if (!bytes || hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    if (hipMemset(p, 0, bytes) != hipSuccess) {
        hipFree(p);
        return nullptr;
    }
    return p;
}

void cvp9_hip_dev_free(void *p)
{
    if (p) hipFree(p);
}

cvp9_err_t cvp9_hip_dev_read(const void *src, uint8_t *dst, uint64_t n)
{
    if (!src || !dst) return CVP9_ERR_INVALID_DATA;
    if (hipDeviceSynchronize() != hipSuccess) return CVP9_ERR_GPU;
    return hipMemcpy(dst, src, n, hipMemcpyDeviceToHost) == hipSuccess ? CVP9_OK : CVP9_ERR_GPU;
}

int cvp9_hip_route_check(const vp9_parsed_frame_t *const *frames, int n, int *codes)
{
    if (n < 0 || (n && (!frames || !codes))) return CVP9_ERR_INVALID_DATA;
    bool declined = false;
    for (int i = 0; i < n; i++) {
        PrepLayout L;
        codes[i] = frames[i] ? route(declined, frames[i], L) : CVP9_ERR_INVALID_DATA;
    }
    return n;
}

} // extern "C"
