/*
 * noalloc.hip — allocation-free single-frame reconstruction
 * (cvp9_engine_run_static, cvp9_hip/engine.h).
 *
 * Every per-frame scratch buffer of the stage kernels lives in one fixed
 * device array, `scratch`, partitioned into regions of fixed offsets:
 *
 *   [0, 4 KiB)         batch descriptor of one job (DJob + prefix rows)
 *   [4 KiB, 8 KiB)     error words (byte flags), epoch counter
 *   owner map          4 MiB   (2^20 4x4 luma cells)
 *   edge records       24 MiB  (49152 blocks x 64 words)
 *   neighbour records  8 MiB   (26880 intra blocks x 39 words)
 *   residuals          24 MiB  (12 Mi int16)
 *
 * The regions never move, so a region only ever holds its own kind of data:
 * in particular the edge region holds nothing but edge words, whose epochs
 * all predate the epoch counter's next value (the counter lives in the
 * array too and only grows), so a word left by an earlier call never looks
 * valid. The descriptor is written by a one-thread-block kernel from its
 * kernel arguments, which also clears the error words: no memcpy, no memset.
 *
 * In host code `scratch` is only a stand-in for the symbol: its device
 * address comes from hipGetSymbolAddress, once per device, and only that
 * address is used.
 */
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>

#include "cvp9_hip/engine.h"
#include "engine.hpp"

using namespace cvp9;

namespace {

constexpr size_t kOffDesc = 0;
constexpr size_t kOffCtrl = 4096;            /* error words at +0 and +4, epoch counter at +64 */
constexpr size_t kOffOwner = 8192;
constexpr size_t kOwnerBytes = 4ull << 20;
constexpr size_t kOffEdge = kOffOwner + kOwnerBytes;
constexpr size_t kEdgeBytes = 24ull << 20;
constexpr size_t kOffNbr = kOffEdge + kEdgeBytes;
constexpr size_t kNbrBytes = 8ull << 20;
constexpr size_t kOffResid = kOffNbr + kNbrBytes;
constexpr size_t kResidBytes = 24ull << 20;
constexpr size_t kScratchBytes = kOffResid + kResidBytes;
static_assert(kScratchBytes == CVP9_STATIC_SCRATCH_BYTES, "scratch size of the header");

constexpr uint32_t kSpinLimit = 1u << 21;

/* The descriptor bytes, passed by value to the setup kernel. */
constexpr uint32_t kImageWords = 512;
struct DescImage {
    uint32_t w[kImageWords];
};

} // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

/* Writes the batch descriptor and clears the error words (one block). */
__global__ __launch_bounds__(256) void k_static_setup(uint8_t *base, DescImage img, uint32_t nwords)
{
    uint32_t *d = (uint32_t *)(base + kOffDesc);
    for (uint32_t i = threadIdx.x; i < nwords; i += blockDim.x) d[i] = img.w[i];
    if (threadIdx.x == 0) {
        *(uint32_t *)(base + kOffCtrl) = 0u;
        *(uint32_t *)(base + kOffCtrl + 4) = 0u;
    }
}

/* Zeroes [lo, hi) of the array (16-byte multiples): the control words and
 * the edge region, once per device before the first frame, so that no word
 * there can match an epoch the counter will hand out. */
__global__ __launch_bounds__(256) void k_static_clear(uint8_t *base, size_t lo, size_t hi)
{
    uint4 *p = (uint4 *)(base + lo);
    const size_t n = (hi - lo) / 16;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        p[i] = make_uint4(0, 0, 0, 0);
}

namespace {

struct DevState {
    uint8_t *ptr = nullptr; /* device address of `scratch` on this device */
    Tune tune;
    bool ready = false;
    bool cleared = false;   /* k_static_clear has been enqueued */
};

DevState g_dev[64];
std::mutex g_mu;

/* Once per device, with that device current: the IDCT tables, the device
 * address of the scratch array, the default launch shape. */
cvp9_err_t setup(int dev, DevState **out)
{
    if (dev < 0 || dev >= 64) return CVP9_ERR_NO_BACKEND;
    std::lock_guard<std::mutex> g(g_mu);
    DevState &d = g_dev[dev];
    if (!d.ready) {
        if (init_device(dev) != hipSuccess) return CVP9_ERR_NO_BACKEND;
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CVP9_ERR_GPU;
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 256;
        d.ptr = (uint8_t *)p;
        d.tune = cvp9_default_tune(cus);
        d.ready = true;
    }
    *out = &d;
    return CVP9_OK;
}

DFrame view(const cvp9_frame_buf_t &b)
{
    DFrame f;
    memset(&f, 0, sizeof(f));
    if (!b.base) return f;
    frame_bytes(b.width, b.height, &f.sy, &f.suv);
    f.y = b.base;
    f.u = b.base + (size_t)f.sy * b.height;
    f.v = f.u + (size_t)f.suv * ((b.height + 1) / 2);
    f.w = b.width;
    f.h = b.height;
    return f;
}

bool usable(const cvp9_frame_buf_t &b)
{
    return b.width >= 2 && b.height >= 2 && b.width <= 16384 && b.height <= 16384 &&
           ((uintptr_t)b.base & 255) == 0;
}

} // namespace

extern "C" {

uint64_t cvp9_frame_bytes(uint32_t width, uint32_t height)
{
    return frame_bytes(width, height, nullptr, nullptr);
}

cvp9_err_t cvp9_engine_run_static(const cvp9_prep_t *prep, const cvp9_frame_buf_t *refs,
                                  const cvp9_frame_buf_t *out, uint8_t *nv12, uint32_t *status,
                                  void *stream)
{
    if (!prep || !prep->dev || !out || !out->base) return CVP9_ERR_INVALID_DATA;
    const PrepLayout &L = prep->L;
    if (out->width != L.W || out->height != L.H || !usable(*out)) return CVP9_ERR_INVALID_DATA;
    for (int i = 0; refs && i < 8; i++) {
        if (!refs[i].base) continue;
        if (!usable(refs[i]) || refs[i].base == out->base) return CVP9_ERR_INVALID_DATA;
    }
    /* the frame must fit the fixed regions */
    const uint64_t cells = (((uint64_t)((L.W + 3) / 4) * ((L.H + 3) / 4)) + 3) & ~3ull;
    if (cells * 4 > kOwnerBytes ||
        (uint64_t)(L.n_blocks ? L.n_blocks : 1) * CVP9_EDGE_WORDS * 8 > kEdgeBytes ||
        (uint64_t)(L.n_intra ? L.n_intra : 1) * CVP9_NB_WORDS * 8 > kNbrBytes ||
        (uint64_t)L.num_coeffs * 2 > kResidBytes)
        return CVP9_ERR_UNSUPPORTED;

    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return CVP9_ERR_NO_BACKEND;
    DevState *d = nullptr;
    cvp9_err_t err = setup(dev, &d);
    if (err != CVP9_OK) return err;
    uint8_t *const base = d->ptr;

    /* the job, as StreamState::make_job lays it out, with its scratch in
     * the fixed regions */
    DJob j;
    memset(&j, 0, sizeof(j));
    j.out = view(*out);
    for (int i = 0; refs && i < 8; i++) j.ref[i] = view(refs[i]);
    j.blocks = (const DBlock *)(prep->dev + L.off_blocks);
    j.coeffs = (const int16_t *)(prep->dev + L.off_coeffs);
    j.mc_tab = (const uint4 *)(prep->dev + L.off_mc_tab);
    j.px = (const uint32_t *)(prep->dev + L.off_px);
    j.n_px_small = L.n_px_small;
    j.intra = (const uint32_t *)(prep->dev + L.off_intra);
    for (int k = 0; k < 4; k++) {
        j.tx_tab[k] = (const uint32_t *)(prep->dev + L.off_tx_tab[k]);
        j.cls_ntx[k] = L.cls_ntx[k];
    }
    j.owner = (uint32_t *)(base + kOffOwner);
    j.edge = (uint64_t *)(base + kOffEdge);
    j.nbr = (uint64_t *)(base + kOffNbr);
    j.resid = (int16_t *)(base + kOffResid);
    j.nv12 = nv12;
    j.err = (uint32_t *)(base + kOffCtrl);
    j.err_host = (uint32_t *)(base + kOffCtrl + 4); /* unread here: the status copy reads j.err */
    j.ectr = (uint32_t *)(base + kOffCtrl + 64);
    j.n_blocks = L.n_blocks;
    j.n_inter = L.n_inter;
    j.n_mc = j.n_mc_a = L.n_mc;
    j.n_intra = L.n_intra;
    j.ow = (L.W + 3) / 4;
    j.oh = (L.H + 3) / 4;
    j.lf_thr = L.lf_thr;
    j.lf_wx = ((L.W + 3) / 8 + 1 + 63) / 64; /* waves of 64 deblocking units per band */
    j.lf_wy = (L.H + 1) / 8 + 1;             /* bands of 8 rows at 8k - 2 */
    j.fill = L.fill;

    /* the descriptor image, built on the stack: one job, persistent intra
     * workers only (no level tables) */
    alignas(16) uint8_t host[4096];
    static_assert(sizeof(host) >= 4 * kImageWords, "image buffer");
    Tune tune = d->tune;
    tune.intra_level_min = 0;
    const PrepLayout *lays[1] = {&L};
    BatchBuilder bb;
    bb.build(&j, lays, 1, host, base + kOffDesc, tune, kSpinLimit);
    bb.set_resid_all(false);
    if (bb.used() > 4 * kImageWords) return CVP9_ERR_UNSUPPORTED;
    DescImage img;
    memcpy(img.w, host, bb.used());

    const hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    {
        std::lock_guard<std::mutex> g(g_mu);
        if (!d->cleared) {
            hipLaunchKernelGGL(k_static_clear, dim3(1024), dim3(256), 0, s, base, kOffCtrl,
                               kOffEdge + kEdgeBytes);
            if ((e = hipGetLastError()) != hipSuccess) return CVP9_ERR_GPU;
            d->cleared = true;
        }
    }
    hipLaunchKernelGGL(k_static_setup, dim3(1), dim3(256), 0, s, base, img,
                       (uint32_t)((bb.used() + 3) / 4));
    e = hipGetLastError();
    if (e == hipSuccess) e = BatchBuilder::launch(bb.batch(), bb.counts(), s);
    if (e == hipSuccess && status) e = cvp9_launch_copy_word(j.err, status, s);
    return e == hipSuccess ? CVP9_OK : CVP9_ERR_GPU;
}

} // extern "C"
