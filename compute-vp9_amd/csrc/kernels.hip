/*
 * kernels.hip — CDNA4 (gfx950) reconstruction kernels.
 *
 * Every kernel runs over a batch of jobs (frames of independent streams);
 * a single-stream decode is a batch of one. Work items are found from
 * per-stage exclusive prefix sums over the jobs (dev_types.h).
 *
 * Arithmetic follows the reference CPU reconstruction exactly
 * (src/decoder/vp9_reconstruct.c): FP64 matrix IDCT with separately rounded
 * multiplies and adds in ascending k, bilinear 1/8-pel MC with edge clamp,
 * whole-block intra in list order, 4-tap luma deblock. Compile with
 * -ffp-contract=off; the IDCT also uses explicit __dmul_rn/__dadd_rn.
 *
 * Stage order per batch (one stream, in order):
 *   k_fill (only jobs not fully covered) -> k_owner -> k_mc -> k_residual<N>
 *   (inter: added to the prediction; intra: int16 residual kept aside)
 *   -> k_intra -> k_lf -> k_nv12
 */
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include "dev_types.h"
#include "kernels.h"

/* Global-memory views of the descriptor's pointers. The descriptors hold
 * generic pointers, through which the compiler emits FLAT accesses; those
 * count against the LDS/scalar wait counter too, so every LDS or scalar
 * wait would also wait for all memory loads in flight. Accesses through
 * these views are plain global loads and stores. */
#define GMEM __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ GMEM T *gm(T *p)
{
    return (GMEM T *)p;
}

/* a 16- or 8-byte record through a global view, as 32-bit words */
__device__ __forceinline__ uint4 gload4(const void *p)
{
    const GMEM uint32_t *q = (const GMEM uint32_t *)gm(p);
    return make_uint4(q[0], q[1], q[2], q[3]);
}
__device__ __forceinline__ uint2 gload2(const void *p)
{
    const GMEM uint32_t *q = (const GMEM uint32_t *)gm(p);
    return make_uint2(q[0], q[1]);
}

__constant__ double c_M4[16];
__constant__ double c_M8[64];
__constant__ double c_M16[256];
__constant__ double c_M32[1024];

template <int N> __device__ __forceinline__ double mat(int idx);
template <> __device__ __forceinline__ double mat<4>(int idx) { return c_M4[idx]; }
template <> __device__ __forceinline__ double mat<8>(int idx) { return c_M8[idx]; }
template <> __device__ __forceinline__ double mat<16>(int idx) { return c_M16[idx]; }
template <> __device__ __forceinline__ double mat<32>(int idx) { return c_M32[idx]; }

__device__ __forceinline__ int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

/* largest j with pre[j] <= w  (pre non-decreasing, pre[0] = 0 <= w < pre[n]) */
__device__ __forceinline__ uint32_t find_job(const uint32_t *pre, uint32_t n, uint32_t w)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (gm(pre)[mid] <= w) lo = mid; else hi = mid;
    }
    return lo;
}

/* The same lookup by a whole wave for a wave-uniform w (every lane active,
 * n <= 4096): the lanes compare 64 prefix entries at once, so one or two
 * rounds of loads replace the binary search's dependent chain. */
__device__ __forceinline__ uint32_t find_job_wave(const uint32_t *pre, uint32_t n, uint32_t w)
{
    /* a batch of one job (a single stream) needs no lookup: its kernels
     * start on the job's descriptor one memory round trip earlier */
    if (n == 1) return 0;
    /* loads with no branch around them (lanes past n read entry n - 1 and
     * ignore it), so nothing waits before the ballot */
    const uint32_t lane = __lane_id();
    if (n <= 64) {
        const uint32_t v = gm(pre)[min(lane, n - 1)];
        return (uint32_t)__builtin_popcountll(__ballot(lane < n && v <= w)) - 1;
    }
    const uint32_t step = (n + 63) >> 6;
    const uint32_t i = lane * step;
    const uint32_t v = gm(pre)[min(i, n - 1)];
    const uint32_t base = ((uint32_t)__builtin_popcountll(__ballot(i < n && v <= w)) - 1) * step;
    const uint32_t k = base + lane;
    const uint32_t v2 = gm(pre)[min(k, n - 1)];
    return base + (uint32_t)__builtin_popcountll(__ballot(lane < step && k < n && v2 <= w)) - 1;
}

/* Workgroup b of n runs on XCD b % 8 (the dispatcher deals workgroups out
 * round-robin over the 8 XCDs, each with its own L2). Returns the logical
 * index that gives each XCD one contiguous run of the work list, so work
 * items next to each other in the list (and in the frame) share an L2. */
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n)
{
    const uint32_t per = n >> 3, rem = n & 7, x = b & 7, i = b >> 3;
    return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

__device__ __forceinline__ uint64_t wave_ballot(bool p)
{
    return __ballot(p);
}

/* Sum of x over the 64 lanes of a whole wave (wave-uniform result): an
 * inclusive scan inside each row of 16 lanes with DPP row shifts (lanes
 * whose source lies outside the row add 0), then the four row totals. */
__device__ __forceinline__ uint32_t wave_sum(uint32_t x)
{
    int v = (int)x;
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true); /* row_shr:1 */
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true); /* row_shr:2 */
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true); /* row_shr:4 */
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true); /* row_shr:8 */
    /* lane 15 of each row holds the row's sum: row_bcast:15 adds it into
     * the next row, row_bcast:31 adds lane 31 into rows 2 and 3, so lane 63
     * ends with the total */
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false); /* row_bcast:15 */
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false); /* row_bcast:31 */
    return (uint32_t)__builtin_amdgcn_readlane(v, 63);
}

/* A block record, as three 8-byte global loads (records are 24 bytes on an
 * 8-byte grid). */
__device__ __forceinline__ DBlock load_block(const DBlock *p)
{
    const GMEM uint64_t *w = (const GMEM uint64_t *)gm(p);
    const uint64_t kw[3] = {w[0], w[1], w[2]};
    DBlock k;
    memcpy(&k, kw, sizeof(k));
    return k;
}

/* Error reporting. Each cause sets its own byte of the stream's error
 * word: in device memory, which waits poll to give up early once another
 * wait of the stream has (a host-memory word there would put a bus read in
 * every long wait), and in a host-visible copy with a plain system-scope
 * byte store (no read-modify-write crosses the bus), which the host reads
 * without a copy. Both go through the global address space: through a
 * generic pointer they would be flat instructions, which wait on every
 * outstanding access of the wave. */
__device__ __forceinline__ void flag_error(uint32_t *err, uint32_t *err_host, int byte)
{
    __hip_atomic_fetch_or(gm(err), 1u << (8 * byte), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(gm((uint8_t *)err_host) + byte, (uint8_t)1, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t read_error(const uint32_t *err)
{
    return __hip_atomic_load(gm(err), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* The batch's first kernel gives every job its frame epoch: the stream's
 * counter advanced by one (0 is skipped: zeroed edge words never look
 * valid). A stream appears at most once per batch and its batches run in
 * order, so one thread per job updates the counter. Later kernels of the
 * batch read the epoch from the job descriptor. */
__device__ __forceinline__ uint32_t next_epoch(uint32_t *ctr)
{
    uint32_t e = *gm(ctr) + 1;
    if (e == 0) e = 1;
    *gm(ctr) = e;
    return e;
}
__device__ __forceinline__ void assign_epochs(const DBatch &B)
{
    for (uint32_t j = threadIdx.x; j < B.njobs; j += blockDim.x) {
        DJob &J = ((DJob *)B.jobs)[j];
        gm(&J.epoch)[0] = next_epoch(J.ectr);
    }
}

/* ------------------------------------------------------------------------ */
/* 128 fill of the output frame and NONE fill of the owner map, for jobs    */
/* whose blocks do not cover the frame (vp9_frame.c:27-30 initial value).   */

__device__ __forceinline__ void fill_item(const DBatch &B, uint32_t w)
{
    const uint32_t j = find_job_wave(B.pre_fill, B.njobs, w);
    const DJob &J = B.jobs[j];
    const uint32_t chunk = w - B.pre_fill[j];
    const uint64_t fbytes = (uint64_t)J.out.sy * J.out.h +
                            2ull * J.out.suv * ((J.out.h + 1) / 2);
    const uint64_t obytes = ((uint64_t)J.ow * J.oh * 4 + 15) & ~15ull;
    const uint64_t fchunks = (fbytes + 65535) >> 16;
    uint8_t *base;
    uint64_t len;
    uint4 v;
    if (chunk < fchunks) {
        base = J.out.y + ((uint64_t)chunk << 16);
        len = fbytes - ((uint64_t)chunk << 16);
        v = make_uint4(0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u);
    } else {
        const uint64_t c = chunk - fchunks;
        base = (uint8_t *)J.owner + (c << 16);
        len = obytes - (c << 16);
        v = make_uint4(CVP9_OWNER_NONE, CVP9_OWNER_NONE, CVP9_OWNER_NONE, CVP9_OWNER_NONE);
    }
    if (len > 65536) len = 65536;
    GMEM uint32_t *p = (GMEM uint32_t *)gm(base);
    for (uint32_t i = threadIdx.x; i < (uint32_t)(len >> 4); i += 256) {
        p[4 * i] = v.x;
        p[4 * i + 1] = v.y;
        p[4 * i + 2] = v.z;
        p[4 * i + 3] = v.w;
    }
}

__global__ __launch_bounds__(256) void k_fill(DBatch B)
{
    fill_item(B, blockIdx.x);
}

/* ------------------------------------------------------------------------ */
/* Compact coefficients (a frame staged through the drop-in boundary): the  */
/* parser's pool arrives as a nonzero mask per chunk of 64 entries, each    */
/* chunk's first position among the packed values, and the nonzero values; */
/* this stage writes the dense pool the transforms read. One wave per four  */
/* chunks, a lane 4 entries (one 8-byte store): its values sit at the       */
/* chunk's start plus the nonzero count below its first entry, and its     */
/* four loads are issued together at clamped places.                        */

__device__ __forceinline__ void expand_item(const DBatch &B, uint32_t w)
{
    const uint32_t j = find_job_wave(B.pre_exp, B.njobs, w);
    const DJob &J = B.jobs[j];
    const uint32_t lane = __lane_id();
    const uint32_t c = (w - B.pre_exp[j]) * 4 + (lane >> 4);
    const uint32_t cc = min(c, J.n_chunks - 1);
    const uint64_t m = gm(J.cmask)[cc];
    const uint32_t st = gm(J.cstart)[cc];
    const uint32_t b = (lane & 15) * 4;
    const uint32_t nib = (uint32_t)(m >> b) & 15;
    uint32_t k = st + (uint32_t)__builtin_popcountll(m & ((1ull << b) - 1));
    const GMEM int16_t *vals = gm(J.cvals);
    uint32_t v[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const bool nz = (nib >> e) & 1;
        const int16_t x = vals[nz ? k : 0];
        v[e] = nz ? (uint32_t)(uint16_t)x : 0u;
        k += nz;
    }
    if (c < J.n_chunks) {
        GMEM uint32_t *d = (GMEM uint32_t *)gm((int16_t *)J.coeffs + (size_t)c * CVP9_COEF_CHUNK + b);
        d[0] = v[0] | v[1] << 16;
        d[1] = v[2] | v[3] << 16;
    }
}

__global__ __launch_bounds__(256) void k_expand(DBatch B)
{
    const uint32_t w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (w >= B.pre_exp[B.njobs]) return; /* uniform per wave */
    expand_item(B, w);
}

/* ------------------------------------------------------------------------ */
/* Owner map: cell (4x4 luma) -> local index of the block covering it.      */

/* block wv of the batch on 16 lanes (lane = 0..15); jhint: the job of the
 * wave's first block */
__device__ __forceinline__ void owner_item(const DBatch &B, uint32_t wv, uint32_t lane, uint32_t jhint)
{
    const uint32_t j = wv < B.pre_blocks[jhint + 1] ? jhint : find_job(B.pre_blocks, B.njobs, wv);
    const DJob &J = B.jobs[j];
    const uint32_t b = wv - B.pre_blocks[j];
    const DBlock k = load_block(J.blocks + b);
    const uint32_t cx0 = k.x >> 2, cy0 = k.y >> 2;
    const uint32_t ncx = ((k.x + k.bw - 1) >> 2) - cx0 + 1;
    const uint32_t ncy = ((k.y + k.bh - 1) >> 2) - cy0 + 1;
    GMEM uint32_t *const owner = gm(J.owner); /* held in registers across the stores */
    const uint32_t ow = J.ow;
    for (uint32_t i = lane; i < ncx * ncy; i += 16) {
        const uint32_t r = i / ncx, c = i - r * ncx;
        owner[(cy0 + r) * ow + cx0 + c] = b;
    }
}

/* workgroup wg of the owner-map stage: 16 blocks, 16 lanes each */
__device__ __forceinline__ void owner_wg(const DBatch &B, uint32_t wg)
{
    if (wg == 0) assign_epochs(B);
    const uint32_t total = B.pre_blocks[B.njobs];
    /* the wave's first block, wave-uniform */
    const uint32_t w0 = __builtin_amdgcn_readfirstlane(wg * 16 + (threadIdx.x >> 6) * 4);
    if (w0 >= total) return;
    const uint32_t jh = find_job_wave(B.pre_blocks, B.njobs, w0);
    const uint32_t wv = w0 + ((threadIdx.x & 63) >> 4);
    if (wv < total) owner_item(B, wv, threadIdx.x & 15, jh);
}

__global__ __launch_bounds__(256) void k_owner(DBatch B)
{
    owner_wg(B, blockIdx.x);
}

/* ------------------------------------------------------------------------ */
/* Motion compensation (vp9_reconstruct.c:94-131, :248-264, :312-334).      */
/* 16 lanes per unit of an inter block (cvp9_mc_unit_rows(bw) luma rows and */
/* the chroma rows under them). A lane computes one column of 4-sample      */
/* quads over a run of rows: each reference row of the run is loaded and    */
/* filtered horizontally once and serves the output rows above and below    */
/* it. A source window of 5 samples inside the reference row is read with   */
/* two aligned dword loads; others take each tap with the edge clamp.       */

/* Source windows of one lane's quad column in one plane, for the rows of
 * its run: reference row y0 + t (clamped to the plane) for t = 0 .. m, each
 * as the 8 bytes at the 4-aligned column b, and the byte of that window each
 * of the 5 taps sx + e takes, edge clamp included (vp9_reconstruct.c:94-131).
 * b is sx clamped to [0, max(rw - 5, 0)] and aligned down, so the window
 * holds every clamped tap position; it may pass the row end by 3 bytes,
 * inside the row pitch (rows are 64-byte aligned) or the frame's tail slack.
 * Rows past m repeat row y0 + m. The loads are issued with no branch around
 * them and used only after all of them (and the residual's) are in flight:
 * a load inside a branch is waited for before the branch ends, which made
 * each row one memory round trip. */
struct McWin {
    uint32_t lo[CVP9_MC_RUN + 1], hi[CVP9_MC_RUN + 1]; /* bytes 0-3, 4-7 of each row */
    uint32_t sel[4]; /* byte selectors of taps (0,1), (2,3), (1,2), (3,4), as 16-bit pairs */
};

/* Samples in packed 16-bit lanes: the filters run two samples per
 * instruction (v_pk_mad_u16 and friends; every intermediate fits 16 bits:
 * a * (8 - f) + b * f + 4 <= 2044). */
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u16x2 as_u16x2(uint32_t x) { return __builtin_bit_cast(u16x2, x); }

__device__ __forceinline__ void mc_win(McWin &q, const uint8_t *ref, uint32_t rs, int rw, int rh,
                                       int sx, int y0, int m)
{
    const int b = clampi(sx, 0, max(rw - 5, 0)) & ~3;
    uint32_t o[5];
#pragma unroll
    for (int e = 0; e < 5; e++) o[e] = (uint32_t)(clampi(sx + e, 0, rw - 1) - b);
    /* v_perm: selector byte i picks byte i of the result from the 8-byte
     * window (0-3 the low dword, 4-7 the high one), 0x0c gives 0 */
    q.sel[0] = o[0] | 0x0cu << 8 | o[1] << 16 | 0x0cu << 24;
    q.sel[1] = o[2] | 0x0cu << 8 | o[3] << 16 | 0x0cu << 24;
    q.sel[2] = o[1] | 0x0cu << 8 | o[2] << 16 | 0x0cu << 24;
    q.sel[3] = o[3] | 0x0cu << 8 | o[4] << 16 | 0x0cu << 24;
    const GMEM uint8_t *base = gm(ref) + b;
#pragma unroll
    for (int t = 0; t <= CVP9_MC_RUN; t++) {
        const int y = clampi(y0 + min(t, m), 0, rh - 1);
        const GMEM uint32_t *p = (const GMEM uint32_t *)(base + (size_t)y * rs);
        q.lo[t] = p[0];
        q.hi[t] = p[1];
    }
}

/* horizontal pass of window row t: samples 0,1 and 2,3 in 16-bit lanes */
__device__ __forceinline__ void mc_hpass(const McWin &q, int t, u16x2 fx, u16x2 gx, u16x2 &h01,
                                         u16x2 &h23)
{
    const u16x2 l01 = as_u16x2(__builtin_amdgcn_perm(q.hi[t], q.lo[t], q.sel[0]));
    const u16x2 l23 = as_u16x2(__builtin_amdgcn_perm(q.hi[t], q.lo[t], q.sel[1]));
    const u16x2 r01 = as_u16x2(__builtin_amdgcn_perm(q.hi[t], q.lo[t], q.sel[2]));
    const u16x2 r23 = as_u16x2(__builtin_amdgcn_perm(q.hi[t], q.lo[t], q.sel[3]));
    h01 = (l01 * gx + (r01 * fx + (u16x2)4)) >> (u16x2)3;
    h23 = (l23 * gx + (r23 * fx + (u16x2)4)) >> (u16x2)3;
}

__device__ __forceinline__ void store_quad(uint8_t *d, uint32_t v, int n)
{
    GMEM uint8_t *g = gm(d);
    if (n >= 4 && ((uintptr_t)d & 3) == 0) {
        *(GMEM uint32_t *)g = v;
    } else {
        for (int e = 0; e < n && e < 4; e++) g[e] = (uint8_t)(v >> (8 * e));
    }
}

/* The m filtered rows of a run (flat 127 for an empty slot,
 * vp9_reconstruct.c:258-263, :328-333), the residual rows res added when
 * add (saturating 16-bit add, then the clamp to 0..255: the same as the
 * reference's int sum clamped), stored at d with pitch ds, nv samples of the
 * quad valid; nothing is stored by a lane with on false, whose values are
 * computed all the same (so that the loads they use stay ahead of every
 * branch). */
__device__ __forceinline__ void mc_emit(const McWin &q, bool on, int m, bool flat, uint32_t fx,
                                        uint32_t fy, const uint64_t *res, bool add, uint8_t *d,
                                        uint32_t ds, int nv)
{
    const u16x2 fx2 = (u16x2)(unsigned short)fx, gx2 = (u16x2)(unsigned short)(8 - fx);
    const u16x2 fy2 = (u16x2)(unsigned short)fy, gy2 = (u16x2)(unsigned short)(8 - fy);
    u16x2 t01, t23;
    mc_hpass(q, 0, fx2, gx2, t01, t23);
#pragma unroll
    for (int t = 0; t < CVP9_MC_RUN; t++) {
        u16x2 b01, b23;
        mc_hpass(q, t + 1, fx2, gx2, b01, b23);
        u16x2 v01 = (t01 * gy2 + (b01 * fy2 + (u16x2)4)) >> (u16x2)3;
        u16x2 v23 = (t23 * gy2 + (b23 * fy2 + (u16x2)4)) >> (u16x2)3;
        t01 = b01;
        t23 = b23;
        if (flat) v01 = v23 = (u16x2)127;
        if (add) {
            i16x2 s01 = __builtin_elementwise_add_sat(__builtin_bit_cast(i16x2, v01),
                                                      __builtin_bit_cast(i16x2, (uint32_t)res[t]));
            i16x2 s23 = __builtin_elementwise_add_sat(__builtin_bit_cast(i16x2, v23),
                                                      __builtin_bit_cast(i16x2, (uint32_t)(res[t] >> 32)));
            s01 = __builtin_elementwise_min(__builtin_elementwise_max(s01, (i16x2)0), (i16x2)255);
            s23 = __builtin_elementwise_min(__builtin_elementwise_max(s23, (i16x2)0), (i16x2)255);
            v01 = __builtin_bit_cast(u16x2, s01);
            v23 = __builtin_bit_cast(u16x2, s23);
        }
        /* low bytes of the four 16-bit lanes */
        const uint32_t v = __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, v23),
                                                 __builtin_bit_cast(uint32_t, v01), 0x06040200u);
        if (on && t < m) store_quad(d + (size_t)t * ds, v, nv);
    }
}

/* Spreads ncol (<= 16) columns by nrows rows over the 16 lanes of a unit:
 * 16 / ncol groups of lanes, each a run of rows. False for an idle lane. */
__device__ __forceinline__ bool mc_lane(int li, int ncol, int nrows, int &col, int &ra, int &n)
{
    const int groups = CVP9_MC_UNIT / ncol;
    const int g = li / ncol;
    if (g >= groups) return false;
    col = li - g * ncol;
    const int per = (nrows + groups - 1) / groups;
    ra = g * per;
    n = min(nrows - ra, per);
    return n > 0;
}

/* One motion-compensation unit of job J on 16 lanes (li = 0..15): e is its
 * unit-table entry, R the reference frame of its block (R.y == nullptr: the
 * slot is empty or the block names none). A lane's luma and chroma runs are
 * at most CVP9_MC_RUN rows: with blocks of at most 64x64, a unit of bw wide
 * has 64 / next_pow2(ceil(bw/4)) luma rows (4 from bw > 32 on) spread over
 * floor(16 / ceil(bw/4)) lane groups, and its chroma rows likewise. */
__device__ __forceinline__ void mc_unit(const DJob &J, uint4 e, const DFrame &R, int li,
                                        const int16_t *resid)
{
    const int bx = (int)(e.x & 0xffff), by = (int)(e.x >> 16);
    const int bw = (int)(e.y & 0xff), bh = (int)((e.y >> 8) & 0xff);
    const int mvx = (int16_t)(e.z & 0xffff), mvy = (int16_t)(e.z >> 16);
    const int rows = (int)cvp9_mc_unit_rows((uint32_t)bw);
    const int r0 = (int)((e.y >> 20) & 15) * rows;
    /* the block's transforms, for the residual add (resid != nullptr) */
    const uint32_t tf = e.y >> 24;
    const int lt = (int)(tf & 3) + 2, rnx = (int)(tf >> 3);
    const bool radd = resid && !(tf & 4);
    const int rrows = radd ? (bh >> lt) << lt : 0, rcols = rnx << lt;
    const int r1 = min(r0 + rows, bh);
    const int cbx = bx >> 1, cby = by >> 1, cbw = (bw + 1) >> 1;
    const int cr0 = (r0 + 1) >> 1, cr1 = (r1 + 1) >> 1; /* chroma rows under [r0, r1) */
    const int nq = (bw + 3) >> 2, ncq = (cbw + 3) >> 2;   /* nq <= 16, ncq <= 8 */
    /* held in registers: a store could alias the descriptor as far as the
     * compiler knows, and would make it reload it */
    const DFrame out = J.out;
    int lcol = 0, lra = 0, ln = 0, ccol = 0, cra = 0, cn = 0;
    const bool lon = mc_lane(li, nq, r1 - r0, lcol, lra, ln);
    const bool con = mc_lane(li, 2 * ncq, cr1 - cr0, ccol, cra, cn);
    const int lm = min(ln, CVP9_MC_RUN), cm = min(cn, CVP9_MC_RUN);
    const int lpx8 = bx * 8 + mvx, lpy8 = by * 8 + mvy;
    /* chroma: MV halved toward zero, bounds from half the reference luma
     * size; the unit's U columns, then its V columns */
    const int cpx8 = cbx * 8 + mvx / 2, cpy8 = cby * 8 + mvy / 2;
    const int cp = ccol >= ncq, cq = ccol - cp * ncq;
    /* an empty slot predicts flat 127; its loads read the output frame's
     * first row and are discarded */
    const bool flat = !R.y;
    const uint8_t *lref = flat ? out.y : R.y;
    const uint8_t *cref = flat ? out.u : (cp ? R.v : R.u);
    const uint32_t lrs = flat ? 0 : R.sy, crs = flat ? 0 : R.suv;
    const int lrw = flat ? 1 : (int)R.w, lrh = flat ? 1 : (int)R.h;
    const int crw = flat ? 1 : (int)R.w / 2, crh = flat ? 1 : (int)R.h / 2;
    const int lx = (lpx8 >> 3) + 4 * lcol, ly = (lpy8 >> 3) + r0 + lra;
    const int cx = (cpx8 >> 3) + 4 * cq, cy = (cpy8 >> 3) + cr0 + cra;
    uint8_t *ld = out.y + (size_t)(by + r0 + lra) * out.sy + bx + 4 * lcol;
    uint8_t *cd = (cp ? out.v : out.u) + (size_t)(cby + cr0 + cra) * out.suv + cbx + 4 * cq;

    McWin lw, cw;
    mc_win(lw, lref, lrs, lrw, lrh, lx, ly, lm);
    mc_win(cw, cref, crs, crw, crh, cx, cy, cm);
    /* the luma residual under the run (transform stage in resid_all mode):
     * the 4 int16 of each row r0 + lra + t at quad column 4 * lcol of a block
     * whose first transform's coefficients start at e.w, transforms of
     * 4 << (lt - 2) samples, rnx to a row; rows / columns past the transform
     * grid have none (vp9_reconstruct.c:336-363). A row is one 8-byte word
     * when e.w is a multiple of 4 (the transforms' offsets then all are);
     * otherwise it is read at use. */
    const bool ral = (e.w & 3) == 0;
    const int16_t *rbase = resid ? resid : (const int16_t *)out.y;
    uint64_t res[CVP9_MC_RUN];
    bool rok[CVP9_MC_RUN];
    const GMEM int16_t *rptr[CVP9_MC_RUN];
#pragma unroll
    for (int t = 0; t < CVP9_MC_RUN; t++) {
        const int rr = r0 + lra + t, c0 = 4 * lcol;
        const int T = 1 << lt;
        rok[t] = lon && radd && t < lm && rr < rrows && c0 < rcols;
        rptr[t] = gm(rbase) + e.w + (size_t)(((rr >> lt) * rnx + (c0 >> lt)) << (2 * lt)) +
                  ((rr & (T - 1)) << lt) + (c0 & (T - 1));
        const GMEM uint64_t *pw =
            rok[t] && ral ? (const GMEM uint64_t *)rptr[t] : (const GMEM uint64_t *)gm(out.y);
        res[t] = *pw;
    }
#pragma unroll
    for (int t = 0; t < CVP9_MC_RUN; t++) {
        if (!rok[t]) {
            res[t] = 0;
        } else if (!ral) {
            uint64_t v = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) v |= (uint64_t)(uint16_t)rptr[t][k] << (16 * k);
            res[t] = v;
        }
    }
    mc_emit(lw, lon, lm, flat, lpx8 & 7, lpy8 & 7, res, radd, ld, out.sy, bw - 4 * lcol);
    /* the reference adds no chroma residual */
    mc_emit(cw, con, cm, flat, cpx8 & 7, cpy8 & 7, res, false, cd, out.suv, cbw - 4 * cq);
}

/* unit u of the batch list (the fused kernel's grid-stride form) */
__device__ __forceinline__ void mc_unit_at(const DBatch &B, uint32_t u, int li)
{
    const uint32_t j = find_job(B.pre_inter, B.njobs, u);
    const DJob &J = B.jobs[j];
    const uint32_t lu = u - B.pre_inter[j];
    if (lu >= J.n_mc_a) return; /* padding of the job's count */
    const uint4 e = gload4(J.mc_tab + lu);
    const uint32_t ref = (e.y >> 16) & 15;
    DFrame R = {};
    if (ref < 8) R = J.ref[ref];
    mc_unit(J, e, R, li, B.resid_all ? J.resid : nullptr);
}

/* A workgroup takes CVP9_MC_WG_UNITS units of one job (job counts are
 * padded to whole workgroups): the job descriptor is read with scalar loads
 * and its reference table is staged in LDS once. */
/* late: the units [n_mc_a, n_mc) of each job, else [0, n_mc_a) */
__device__ __forceinline__ void mc_wg(const DBatch &B, uint32_t wg, uint32_t nwg, bool late)
{
    static_assert(CVP9_MC_WG_UNITS * CVP9_MC_UNIT == 256, "one unit per 16 lanes");
    __shared__ DFrame s_ref[8];
    const uint32_t *const pre = late ? B.pre_inter2 : B.pre_inter;
    const uint32_t u0 = xcd_block(wg, nwg) * CVP9_MC_WG_UNITS;
    const uint32_t j = __builtin_amdgcn_readfirstlane(find_job_wave(pre, B.njobs, u0));
    const DJob &J = B.jobs[j];
    constexpr uint32_t ref_words = sizeof(s_ref) / 4;
    if (threadIdx.x < ref_words)
        ((uint32_t *)s_ref)[threadIdx.x] = ((const uint32_t *)J.ref)[threadIdx.x];
    const uint32_t lo = late ? J.n_mc_a : 0, hi = late ? J.n_mc : J.n_mc_a;
    const uint32_t lu = lo + u0 - pre[j] + threadIdx.x / CVP9_MC_UNIT;
    const bool live = lu < hi;
    uint4 e = make_uint4(0, 0, 0, 0);
    if (live) e = gload4(J.mc_tab + lu);
    __syncthreads();
    if (!live) return;
    const uint32_t ref = (e.y >> 16) & 15;
    DFrame R = {};
    if (ref < 8) R = s_ref[ref];
    mc_unit(J, e, R, threadIdx.x % CVP9_MC_UNIT, B.resid_all ? J.resid : nullptr);
}

/* (held to 6 waves per SIMD: 70 registers instead of 88, no spills; the
 * latency-bound units then have 7 waves per SIMD in flight instead of 5) */
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_mc(DBatch B, int late)
{
    mc_wg(B, blockIdx.x, gridDim.x, late != 0);
}

/* The owner map and motion compensation in one launch (they are
 * independent): the first owner_wgs workgroups build the owner map. */
__global__ __launch_bounds__(256) void k_front(DBatch B, uint32_t owner_wgs)
{
    if (blockIdx.x < owner_wgs) owner_wg(B, blockIdx.x);
    else mc_wg(B, blockIdx.x - owner_wgs, gridDim.x - owner_wgs, false);
}

/* ------------------------------------------------------------------------ */
/* FP64 inverse DCT + residual add (vp9_reconstruct.c:13-60, :336-363).      */
/*                                                                          */
/* S waves (1, 2 or 4) transform G = 64/N transforms of size N together. In */
/* every wave lane (t, j) stands for column j of transform t; wave g owns   */
/* rows [g*R, g*R+R) in pass 1 and columns [g*R, g*R+R) in pass 2, R = N/S. */
/*   pass 1: temp[i][j] = sum_k M[k][i] * in[k][j] for the wave's rows i,   */
/*           k ascending; M[k][i] is wave-uniform (scalar loads).           */
/*   temp is shared through LDS (row pitch N+1 doubles).                    */
/*   pass 2: lane (t, r): out[r][c] = sum_k temp[r][k] * M[k][c] for the    */
/*           wave's columns c.                                              */
/* Every product and sum is rounded separately, in the reference's order.   */
/* A k-step is skipped when its input row (pass 1) or column (pass 2) is    */
/* zero in every transform of the group. Adding a product with an exact     */
/* zero operand leaves a nonzero sum unchanged and can only flip the sign   */
/* of a zero sum, which the final round-half-away conversion ignores, so    */
/* any superset of the nonzero terms taken in ascending k is bit-exact.     */

__device__ __forceinline__ int round_clamp16(double v)
{
    const int32_t r = (int32_t)__dadd_rn(v, v >= 0.0 ? 0.5 : -0.5);
    return clampi(r, -32768, 32767);
}

template <int N, int S>
__device__ __forceinline__ void idct_group(const int16_t *coef, uint8_t *dst, uint32_t stride,
                                           int16_t *rout, double *lds, int wave, int lane)
{
    constexpr int R = N / S;
    const int t = lane / N;
    const int j = lane % N;
    const int i0 = wave * R;
    /* the prediction this lane's pass-2 outputs add to (row j, the wave's
     * columns), loaded now so it arrives during the transform */
    constexpr int RD = R >= 4 ? R / 4 : 1;
    uint32_t pred[RD];
    const bool inter = coef && !rout;
    if (inter) {
        const GMEM uint8_t *row = gm(dst) + (size_t)j * stride + i0;
        if constexpr (R >= 4) {
#pragma unroll
            for (int c4 = 0; c4 < RD; c4++) pred[c4] = *(const GMEM uint32_t *)(row + 4 * c4);
        } else {
            pred[0] = 0;
#pragma unroll
            for (int c = 0; c < R; c++) pred[0] |= (uint32_t)row[c] << (8 * c);
        }
    }
    /* the lane's whole input column, every load issued before the first is
     * used (a load per k-step behind the step's skip test would make the
     * transform one memory round trip per row) */
    int xs[N];
    if (coef) {
        const GMEM int16_t *cp = gm(coef) + j;
#pragma unroll
        for (int k = 0; k < N; k++) xs[k] = cp[k * N];
    } else {
#pragma unroll
        for (int k = 0; k < N; k++) xs[k] = 0;
    }
    double acc[R];
#pragma unroll
    for (int i = 0; i < R; i++) acc[i] = 0.0;
    bool col_nz = false;

#pragma unroll
    for (int k = 0; k < N; k++) {
        const int x = xs[k];
        col_nz |= (x != 0);
        if (wave_ballot(x != 0)) {
            const double xd = (double)x;
#pragma unroll
            for (int i = 0; i < R; i++)
                acc[i] = __dadd_rn(acc[i], __dmul_rn(mat<N>(k * N + i0 + i), xd));
        }
    }
    double *tt = lds + t * (N * (N + 1));
#pragma unroll
    for (int i = 0; i < R; i++) tt[(i0 + i) * (N + 1) + j] = acc[i];

    /* nonzero input columns over all transforms of the group, folded to N bits */
    uint64_t m = wave_ballot(col_nz);
#pragma unroll
    for (int s = 32; s >= N; s >>= 1) m |= m >> s;
    if constexpr (S == 1) {
        /* the group's temp is this wave's own */
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }

    const int r = j;
#pragma unroll
    for (int c = 0; c < R; c++) acc[c] = 0.0;
    /* temp row r in chunks of KC columns, each chunk's LDS reads issued
     * together */
    constexpr int KC = N < 8 ? N : 8;
#pragma unroll
    for (int k0 = 0; k0 < N; k0 += KC) {
        if (((m >> k0) & ((1ull << KC) - 1)) == 0) continue;
        double u[KC];
#pragma unroll
        for (int kk = 0; kk < KC; kk++) u[kk] = tt[r * (N + 1) + k0 + kk];
#pragma unroll
        for (int kk = 0; kk < KC; kk++) {
            const int k = k0 + kk;
            if ((m >> k) & 1) {
#pragma unroll
                for (int c = 0; c < R; c++)
                    acc[c] = __dadd_rn(acc[c], __dmul_rn(u[kk], mat<N>(k * N + i0 + c)));
            }
        }
    }

    if (coef && rout) {
        /* intra transform: keep the int16 residual for the intra stage */
        GMEM int16_t *o = gm(rout) + r * N + i0;
#pragma unroll
        for (int c = 0; c < R; c++) o[c] = (int16_t)round_clamp16(acc[c]);
    } else if (inter) {
        GMEM uint8_t *row = gm(dst) + (size_t)r * stride + i0;
        if constexpr (R >= 4) {
#pragma unroll
            for (int c4 = 0; c4 < R; c4 += 4) {
                const uint32_t px = pred[c4 / 4];
                uint32_t o = 0;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int p = (int)((px >> (8 * e)) & 0xff);
                    o |= (uint32_t)clampi(p + round_clamp16(acc[c4 + e]), 0, 255) << (8 * e);
                }
                *(GMEM uint32_t *)(row + c4) = o;
            }
        } else {
#pragma unroll
            for (int c = 0; c < R; c++)
                row[c] = (uint8_t)clampi((int)((pred[0] >> (8 * c)) & 0xff) + round_clamp16(acc[c]),
                                         0, 255);
        }
    }
}

/* lane's transform in a group of a job's size class c: coefficient pointer
 * and destination origin, or coef = nullptr past the end */
template <int N>
__device__ __forceinline__ void locate_tx(const DJob &J, int cls, uint32_t txi, bool resid_all,
                                          const int16_t *&coef, uint8_t *&dst, int16_t *&rout)
{
    coef = nullptr;
    dst = nullptr;
    rout = nullptr;
    if (txi >= J.cls_ntx[cls]) return;
    const uint2 t = gload2((const uint2 *)J.tx_tab[cls] + txi);
    coef = J.coeffs + t.x;
    if ((t.y >> 31) || resid_all)
        rout = J.resid + t.x;
    else
        dst = J.out.y + (size_t)((t.y >> 14) & 0x3fff) * J.out.sy + (t.y & 0x3fff);
}

template <int N, int S>
struct ResShape {
    static constexpr int C = N == 4 ? 0 : N == 8 ? 1 : N == 16 ? 2 : 3;
    static constexpr int G = 64 / N;
    static constexpr int GPW = 4 / S;                         /* groups per workgroup */
    static constexpr int LDS_DOUBLES = GPW * G * N * (N + 1); /* temp of all groups */
};

/* one workgroup unit: GPW transform groups of size N, S waves per group */
template <int N, int S>
__device__ __forceinline__ void residual_unit(const DBatch &B, uint32_t unit, double *lds)
{
    using RS = ResShape<N, S>;
    constexpr int C = RS::C, G = RS::G, GPW = RS::GPW;
    /* the wave index made wave-uniform, so the matrix columns a wave owns come
     * through scalar loads */
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gi = wave / S, sub = wave - gi * S;
    const uint32_t grp = unit * GPW + gi;
    const uint32_t *pre = B.pre_grp[C];
    /* every wave reaches the barrier inside idct_group: a group past the
     * end runs with no transform */
    const int16_t *coef = nullptr;
    uint8_t *dst = nullptr;
    int16_t *rout = nullptr;
    uint32_t stride = 0;
    if (grp < pre[B.njobs]) {
        const uint32_t j = find_job_wave(pre, B.njobs, grp); /* grp is wave-uniform */
        const DJob &J = B.jobs[j];
        stride = J.out.sy;
        locate_tx<N>(J, C, (grp - pre[j]) * G + lane / N, B.resid_all != 0, coef, dst, rout);
    }
    idct_group<N, S>(coef, dst, stride, rout, lds + gi * (G * N * (N + 1)), sub, lane);
}

template <int N, int S>
__global__ __launch_bounds__(256) void k_residual(DBatch B)
{
    __shared__ double lds[ResShape<N, S>::LDS_DOUBLES];
    residual_unit<N, S>(B, blockIdx.x, lds);
}

/* 4x4 transforms, one per lane: a wave takes four groups of 16 (each job's
 * group count is padded to a multiple of four, so a wave's groups belong to
 * one job), a lane its whole transform. Both passes run in registers and
 * take every term, in the reference's order (the same sums the skipping
 * form computes, see above); there is no LDS and no barrier, and a wave
 * holds four times the transforms of the column-per-lane form, so it pays
 * its memory round trips once for 64 transforms. */
__global__ __launch_bounds__(256) void k_res4_lanes(DBatch B)
{
    const uint32_t *pre = B.pre_grp[0];
    const uint32_t g0 =
        __builtin_amdgcn_readfirstlane((blockIdx.x * 4 + (threadIdx.x >> 6)) * 4);
    if (g0 >= pre[B.njobs]) return; /* uniform per wave */
    const uint32_t j = find_job_wave(pre, B.njobs, g0);
    const DJob &J = B.jobs[j];
    const uint32_t txi = (g0 - pre[j]) * 16 + __lane_id();
    if (txi >= J.cls_ntx[0]) return; /* no barrier below */
    const uint2 t = gload2((const uint2 *)J.tx_tab[0] + txi);
    const bool to_res = (t.y >> 31) || B.resid_all;
    const GMEM int16_t *cp = gm(J.coeffs) + t.x;
    int in[16];
#pragma unroll
    for (int i = 0; i < 16; i++) in[i] = cp[i];
    /* the prediction an inter transform adds to, loaded with the
     * coefficients by every lane (a transform's rows lie inside the frame,
     * rows are 64-byte aligned and x is a multiple of 4; a transform whose
     * residual is kept ignores it) */
    GMEM uint8_t *dst = gm(J.out.y) + (size_t)((t.y >> 14) & 0x3fff) * J.out.sy + (t.y & 0x3fff);
    const uint32_t sy = J.out.sy;
    uint32_t pred[4];
#pragma unroll
    for (int r = 0; r < 4; r++) pred[r] = *(const GMEM uint32_t *)(dst + r * sy);
    /* pass 1: temp[i][c] = sum_k M[k][i] * in[k][c] */
    double tmp[16];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++)
                acc = __dadd_rn(acc, __dmul_rn(mat<4>(k * 4 + i), (double)in[k * 4 + c]));
            tmp[i * 4 + c] = acc;
        }
    /* pass 2: out[r][c] = sum_k temp[r][k] * M[k][c] */
    int out[16];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) acc = __dadd_rn(acc, __dmul_rn(tmp[r * 4 + k], mat<4>(k * 4 + c)));
            out[r * 4 + c] = round_clamp16(acc);
        }
    if (to_res) {
        GMEM int16_t *o = gm(J.resid) + t.x;
#pragma unroll
        for (int i = 0; i < 16; i++) o[i] = (int16_t)out[i];
    } else {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            uint32_t w = 0;
#pragma unroll
            for (int c = 0; c < 4; c++)
                w |= (uint32_t)clampi((int)((pred[r] >> (8 * c)) & 0xff) + out[r * 4 + c], 0, 255)
                     << (8 * c);
            *(GMEM uint32_t *)(dst + r * sy) = w;
        }
    }
}

/* Every transform size in one launch, for batches too small to fill the
 * device with one size at a time: workgroups [0, wg.x) take 4x4, then
 * [wg.x, wg.y) 8x8, [wg.y, wg.z) 16x16 and the rest 32x32, each with the
 * wave split whose registers and LDS suit a small batch (1, 1, 2, 4). The
 * sizes write disjoint samples, so their order does not matter. */
constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
constexpr size_t RES_SMALL_LDS =
    cmax(cmax(ResShape<4, 1>::LDS_DOUBLES, ResShape<8, 1>::LDS_DOUBLES),
         cmax(ResShape<16, 2>::LDS_DOUBLES, ResShape<32, 4>::LDS_DOUBLES));

__device__ __forceinline__ void res_small_wg(const DBatch &B, uint32_t b, uint3 wg, double *lds)
{
    if (b < wg.x) residual_unit<4, 1>(B, b, lds);
    else if (b < wg.y) residual_unit<8, 1>(B, b - wg.x, lds);
    else if (b < wg.z) residual_unit<16, 2>(B, b - wg.y, lds);
    else residual_unit<32, 4>(B, b - wg.z, lds);
}

__global__ __launch_bounds__(256) void k_res_small(DBatch B, uint3 wg)
{
    __shared__ double lds[RES_SMALL_LDS];
    res_small_wg(B, blockIdx.x, wg, lds);
}

/* ------------------------------------------------------------------------ */
/* Intra blocks (vp9_reconstruct.c:63-91, :224-247, :272-311, :336-363).     */
/*                                                                          */
/* The reference predicts blocks strictly in list order, so a neighbour     */
/* sample holds its final value only if its owner block comes earlier in    */
/* the list; otherwise it still holds the initial 128. Inter blocks are     */
/* final before this kernel starts, and the residuals of intra blocks were  */
/* computed by the residual stage (they do not depend on neighbours), so    */
/* the ordered part is prediction + residual add only.                      */
/*                                                                          */
/* Intra blocks of a job are taken in the order of the job's intra list    */
/* (dependency level, then list index) through the job's atomic ticket. A   */
/* worker starts on the job that holds its share of the batch's             */
/* intra blocks and moves on to the next job when that one's tickets run    */
/* out, so the chains of all frames advance together. Each worker publishes */
/* its block's final bottom row and right column of Y, U, V                 */
/* as 64-bit words (4 samples + the frame epoch), written and read with     */
/* device-scope atomics: a word is valid exactly when its epoch matches, so */
/* no fences or flags are needed. A neighbour sample of block b owned by an */
/* earlier intra block o always lies on o's bottom row or right column:     */
/* o covers the sample and not the sample's neighbour inside b, and blocks  */
/* are disjoint rectangles. A block waits (bounded) only on earlier intra   */
/* blocks of its own job at lower dependency levels; the host orders each  */
/* job's intra list by level, so their tickets were handed out earlier, to  */
/* workers that are running, and every wait ends.                           */

/* Neighbour stage: one wave per intra block, one lane per neighbour cell,
 * from the owner map alone (so it can run before motion compensation). A
 * cell owned by no block or by a later block holds the initial 128; a cell
 * of an inter block is marked (CVP9_NB_IMASK) for the intra stage to read
 * from the frame, where it is final by then; a cell of an earlier intra
 * block becomes a pointer into that block's edge record. A sample of block b owned by an earlier intra block o always lies
 * on o's bottom row or right column: o covers the sample and not the
 * sample's neighbour inside b, and blocks are disjoint rectangles. The
 * chroma samples next to b lie in the same 4x4 luma cells as the luma ones
 * (x, y are multiples of 4). */
/* Intra block wv of the batch on 32 lanes (li = 0..31), one cell per lane
 * and pass (at most 33 cells: a second pass only for 64-wide or -high
 * blocks). */
__device__ __forceinline__ void nbr_item(const DBatch &B, uint32_t wv, uint32_t li, uint32_t jhint)
{
    const uint32_t j = wv < B.pre_intra[jhint + 1] ? jhint : find_job(B.pre_intra, B.njobs, wv);
    const DJob &J = B.jobs[j];
    const uint32_t tk = wv - B.pre_intra[j];
    const uint32_t b = gm(J.intra)[tk];
    const DBlock k = load_block(J.blocks + b);
    const int x = k.x, y = k.y, bw = k.bw, bh = k.bh;
    const int cbx = x >> 1, cby = y >> 1;
    const int ncx = (bw + 3) >> 2, ncy = (bh + 3) >> 2;
    GMEM uint64_t *rec = gm(J.nbr) + (size_t)tk * CVP9_NB_WORDS;
    /* fields held in registers across the record stores */
    const GMEM uint32_t *const owner = gm(J.owner);
    const GMEM DBlock *const blocks = gm(J.blocks);
    const uint32_t ow = J.ow;
    /* cells the block's modes do not read are no dependency: left as 128 */
    const uint32_t reads = cvp9_mode_reads(k.ymode, k.uvmode);
    uint64_t mask = 0, imask = 0;
    for (int c = (int)li; c <= ncx + ncy; c += 32) {
        bool dep, inter = false;
        int kind, q = 0;
        bool present;
        if (c < ncx) { kind = 0; q = c; present = y > 0 && (reads & CVP9_READS_ABOVE); }
        else if (c < ncx + ncy) { kind = 1; q = c - ncx; present = x > 0 && (reads & CVP9_READS_LEFT); }
        else { kind = 2; present = x > 0 && y > 0 && (reads & CVP9_READS_CORNER); }
        uint64_t val = 0;
        dep = false;
        if (present) {
            const int cx = kind == 0 ? (x >> 2) + q : (x >> 2) - 1;
            const int cy = kind == 1 ? (y >> 2) + q : (y >> 2) - 1;
            const uint32_t o = owner[(uint32_t)cy * ow + (uint32_t)cx];
            if (o == CVP9_OWNER_NONE || o >= b) {
                val = 0x8080808080808080ull; /* still the initial 128 */
            } else {
                DBlock n;
                {
                    const GMEM uint64_t *q = (const GMEM uint64_t *)gm(blocks + o);
                    const uint64_t w[3] = {q[0], q[1], q[2]};
                    memcpy(&n, w, sizeof(n));
                }
                if (!(n.flags & DB_INTRA)) {
                    inter = true; /* read by the intra stage */
                } else {
                    const int ox = n.x, oy = n.y;
                    const bool bottom_row = y - 1 == oy + n.bh - 1, right_col = x - 1 == ox + n.bw - 1;
                    int side, lw, ci; /* edge side, luma word, first chroma index */
                    if (kind == 0 && bottom_row) {
                        side = 0; lw = (x + 4 * q - ox) >> 2; ci = cbx + 2 * q - (ox >> 1);
                    } else if (kind == 1 && right_col) {
                        side = 1; lw = (y + 4 * q - oy) >> 2; ci = cby + 2 * q - (oy >> 1);
                    } else if (kind == 2 && bottom_row) {
                        side = 0; lw = (x - 4 - ox) >> 2; ci = cbx - 2 - (ox >> 1);
                    } else if (kind == 2 && right_col) {
                        side = 1; lw = (y - 4 - oy) >> 2; ci = cby - 2 - (oy >> 1);
                    } else {
                        flag_error(J.err, J.err_host, 1); /* not on an edge: impossible for disjoint blocks */
                        side = 0; lw = 0; ci = 0;
                    }
                    val = (uint64_t)o | (uint64_t)side << 32 | (uint64_t)(lw & 15) << 33 |
                          (uint64_t)(ci & 31) << 40;
                    dep = true;
                }
            }
        }
        rec[c] = val;
        if (dep) mask |= 1ull << c;
        if (inter) imask |= 1ull << c;
    }
    /* OR of the 32 lanes' dependency and inter bits */
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) {
        mask |= __shfl_xor(mask, off, 32);
        imask |= __shfl_xor(imask, off, 32);
    }
    if (li == 0) rec[CVP9_NB_MASK] = mask;
    if (li == 1) rec[CVP9_NB_IMASK] = imask;
    if (li == 2) rec[CVP9_NB_INDEX] = b;
    if (li >= 3 && li < 3 + CVP9_NB_WORDS - CVP9_NB_BLOCK)
        rec[CVP9_NB_BLOCK + li - 3] = ((const GMEM uint64_t *)gm(blocks + b))[li - 3];
}

/* workgroup wg of nwg: 8 intra blocks, each XCD a contiguous run */
__device__ __forceinline__ void nbr_wg(const DBatch &B, uint32_t wg, uint32_t nwg)
{
    const uint32_t total = B.pre_intra[B.njobs];
    const uint32_t w0 = __builtin_amdgcn_readfirstlane(xcd_block(wg, nwg) * 8 + (threadIdx.x >> 6) * 2);
    if (w0 >= total) return;
    const uint32_t jh = find_job_wave(B.pre_intra, B.njobs, w0); /* the wave's first block */
    const uint32_t wv = w0 + ((threadIdx.x & 63) >> 5);
    if (wv < total) nbr_item(B, wv, threadIdx.x & 31, jh);
}

__global__ __launch_bounds__(256) void k_nbr(DBatch B)
{
    nbr_wg(B, blockIdx.x, gridDim.x);
}

/* The pipelined first phase in two launches instead of four: the owner map
 * beside every transform size (k_owner_res: they touch disjoint data), then
 * the neighbour records beside the early motion compensation (k_nbr_mc:
 * the records read the owner map, motion compensation the residuals). A
 * launch costs the host a few microseconds, and one stream's frame rate is
 * bound by the host's enqueue time. Each launch's first part is padded to
 * whole groups of 8 workgroups, so that workgroup b of its second part
 * still runs on XCD b % 8. */
__global__ __launch_bounds__(256) void k_owner_res(DBatch B, uint32_t owner_wgs, uint3 wg)
{
    __shared__ double lds[RES_SMALL_LDS];
    if (blockIdx.x < owner_wgs) owner_wg(B, blockIdx.x);
    else res_small_wg(B, blockIdx.x - owner_wgs, wg, lds);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_nbr_mc(DBatch B, uint32_t nbr_wgs, uint32_t nbr_used)
{
    if (blockIdx.x < nbr_wgs) {
        if (blockIdx.x < nbr_used) nbr_wg(B, blockIdx.x, nbr_used);
    } else {
        mc_wg(B, blockIdx.x - nbr_wgs, gridDim.x - nbr_wgs, false);
    }
}

/* The samples of an intra cell in the resolved layout (bytes 0-3 luma,
 * 4-5 U, 6-7 V), read from the owner's edge record: its three words (Y, U,
 * V) are loaded together and re-read, bounded, until each carries this
 * frame's epoch. The poll sleeps 64 clocks between reads: a ping-pong probe
 * found thousands of other workgroups polling add only 0.1-0.2 us to a
 * hand-off, and a longer sleep adds its length to every link of the chain. */
__device__ __forceinline__ uint64_t cell_from_edges(const uint64_t *edge, uint32_t epoch, uint32_t *err,
                                                    uint32_t *err_host, uint64_t d, uint32_t spin_limit)
{
    const uint32_t o = (uint32_t)d;
    const uint32_t side = (uint32_t)(d >> 32) & 1, lw = (uint32_t)(d >> 33) & 15;
    const uint32_t c = (uint32_t)(d >> 40) & 31;
    const GMEM uint64_t *rec = gm(edge) + (size_t)o * CVP9_EDGE_WORDS;
    const GMEM uint64_t *py = rec + (side ? 16 : 0) + lw;
    const GMEM uint64_t *pu = rec + (side ? 40 : 32) + (c >> 2);
    const GMEM uint64_t *pv = rec + (side ? 56 : 48) + (c >> 2);
    uint64_t wy = __hip_atomic_load(py, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t wu = __hip_atomic_load(pu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t wv = __hip_atomic_load(pv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t spins = 0;
    for (;;) {
        const bool sy = (uint32_t)(wy >> 32) != epoch, su = (uint32_t)(wu >> 32) != epoch,
                   sv = (uint32_t)(wv >> 32) != epoch;
        if (!(sy || su || sv)) break;
        __builtin_amdgcn_s_sleep(1); /* 64 clocks */
        if (++spins > spin_limit) {
            flag_error(err, err_host, 0);
            break;
        }
        if ((spins & 255) == 0 && read_error(err)) break;
        if (sy) wy = __hip_atomic_load(py, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (su) wu = __hip_atomic_load(pu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (sv) wv = __hip_atomic_load(pv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const uint32_t sh = 8 * (c & 3);
    return (wy & 0xffffffffull) | ((wu >> sh) & 0xffff) << 32 | ((wv >> sh) & 0xffff) << 48;
}



/* x / n for x < 2^16 and 1 <= n <= 128, in a few instructions: the
 * truncated product with a float reciprocal (v_rcp_f32, within 1 ulp) is
 * within one of the quotient, and one step each way corrects it (exact over
 * the whole range for reciprocal errors up to 2 ulp, checked exhaustively) */
__device__ __forceinline__ uint32_t udiv_small_r(uint32_t x, uint32_t n, float rcp_n)
{
    const uint32_t q = (uint32_t)((float)x * rcp_n);
    const int r = (int)x - (int)(q * n);
    return q + (r >= (int)n ? 1u : 0u) - (r < 0 ? 1u : 0u);
}
__device__ __forceinline__ uint32_t udiv_small(uint32_t x, uint32_t n)
{
    return udiv_small_r(x, n, __builtin_amdgcn_rcpf((float)n));
}

/* Lane src's value of x (a byte address src * 4 taken modulo 64 lanes) */
__device__ __forceinline__ uint32_t lane_read(uint32_t byte_addr, uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)byte_addr, (int)x);
}

/* Every candidate is computed and one selected, with no branch: the
 * neighbour samples come from LDS, and a branch per mode would move each
 * read into its own branch and wait for it there. */
__device__ __forceinline__ int intra_pred(int mode, int a, int l, int tl, int dc)
{
    const int tm = clampi(a + l - tl, 0, 255);
    return mode == 0 ? dc : (mode == 1 ? a : (mode == 2 ? l : tm));
}

/* One neighbour cell of an intra block (thread c of the gather; kind 0
 * above, cell q, 1 left, cell q, 2 above-left), in the layout of a resolved
 * cell (4 Y, 2 U, 2 V samples):
 *   pend  (an earlier intra block): from the owner's edge record, one word
 *         each of Y, U and V; stale when a word does not carry the frame's
 *         epoch yet (the owner has not published it);
 *   inter (an inter block): from the frame, final once motion compensation
 *         and the inter residuals are done;
 *   else  the resolved value d.
 * Every load is issued for every thread, at an address inside the edge
 * records or the frame, and the selection follows the loads: a load inside
 * a branch would be waited for before the branch ends, making the gather
 * one memory round trip per kind of cell instead of one in all. Block
 * positions are multiples of 4, so each sample lies in the dword named
 * below; rows are 64-byte aligned, so the dword lies inside the row pitch. */
/* The loads of cell_gather, and the selection that follows them: split so
 * that a chain worker can issue the loads of its next block before it
 * finishes the current one. */
struct CellRaw {
    uint64_t wy, wu, wv;             /* edge words (Y, U, V) */
    uint32_t fy[4], fu[2], fv[2];    /* frame dwords */
};

__device__ __forceinline__ CellRaw cell_load(const uint64_t *edge, uint32_t self, const DFrame &F,
                                             int x, int y, int bh, int kind, int q, bool pend,
                                             bool inter, uint64_t d)
{
    CellRaw r;
    /* edge words (a thread without an edge cell reads the block's own
     * record, and discards it) */
    const uint64_t ed = pend ? d : (uint64_t)self;
    const uint32_t o = (uint32_t)ed;
    const uint32_t side = (uint32_t)(ed >> 32) & 1, lw = (uint32_t)(ed >> 33) & 15;
    const uint32_t cc = (uint32_t)(ed >> 40) & 31;
    const GMEM uint64_t *rec = gm(edge) + (size_t)o * CVP9_EDGE_WORDS;
    r.wy = __hip_atomic_load(rec + (side ? 16 : 0) + lw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    r.wu = __hip_atomic_load(rec + (side ? 40 : 32) + (cc >> 2), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    r.wv = __hip_atomic_load(rec + (side ? 56 : 48) + (cc >> 2), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);

    /* frame dwords, four of luma and two of each chroma plane, and the byte
     * each contributes (shift, and whether it is one of the cell's samples):
     *   above (kind 0): row y-1, the cell's 4 Y samples and 2 U / V samples,
     *     every luma dword the same one, every chroma dword likewise;
     *   left (kind 1): rows of the cell, the sample left of the block in each
     *     (rows past the block's end have none: 0);
     *   above-left (kind 2): row y-1, the single samples in bytes 3, 5, 7.
     * All loads are used the same way whatever the kind, so none is moved
     * into a branch of its own. A thread without an inter cell reads row 0,
     * column 0 and keeps nothing. */
    const int cbx = x >> 1, cby = y >> 1, cbh = (bh + 1) >> 1;
    const int colY = kind == 0 ? x + 4 * q : x - 4;
    const int cC = kind == 0 ? cbx + 2 * q : cbx - 1;
    const int colC = cC & ~3;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int rr = kind == 1 ? y + min(4 * q + e, bh - 1) : y - 1;
        const size_t off = inter ? (size_t)rr * F.sy + colY : 0;
        r.fy[e] = *(const GMEM uint32_t *)(gm(F.y) + off);
    }
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const int rr = kind == 1 ? cby + min(2 * q + e, cbh - 1) : cby - 1;
        const size_t off = inter ? (size_t)rr * F.suv + colC : 0;
        r.fu[e] = *(const GMEM uint32_t *)(gm(F.u) + off);
        r.fv[e] = *(const GMEM uint32_t *)(gm(F.v) + off);
    }
    return r;
}

__device__ __forceinline__ uint64_t cell_value(const CellRaw &r, uint32_t epoch, int x, int bh,
                                               int kind, int q, bool pend, bool inter, uint64_t d,
                                               bool &stale)
{
    const int cbx = x >> 1, cbh = (bh + 1) >> 1;
    const int cC = kind == 0 ? cbx + 2 * q : cbx - 1;
    const int shC = 8 * (cC & 3);
    uint64_t v_inter = 0;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const uint32_t sh = kind == 0 ? 8 * e : 24;
        const bool keep = kind == 0 || (kind == 1 ? 4 * q + e < bh : e == 3);
        v_inter |= (uint64_t)(keep ? (r.fy[e] >> sh) & 0xff : 0u) << (8 * e);
    }
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const uint32_t sh = kind == 0 ? shC + 8 * e : shC;
        const bool keep = kind == 0 || (kind == 1 ? 2 * q + e < cbh : e == 1);
        v_inter |= (uint64_t)(keep ? (r.fu[e] >> sh) & 0xff : 0u) << (32 + 8 * e);
        v_inter |= (uint64_t)(keep ? (r.fv[e] >> sh) & 0xff : 0u) << (48 + 8 * e);
    }

    const uint32_t cc = (uint32_t)((pend ? d : 0) >> 40) & 31;
    const uint32_t sh = 8 * (cc & 3);
    const uint64_t v_edge =
        (r.wy & 0xffffffffull) | ((r.wu >> sh) & 0xffff) << 32 | ((r.wv >> sh) & 0xffff) << 48;
    stale = pend && ((uint32_t)(r.wy >> 32) != epoch || (uint32_t)(r.wu >> 32) != epoch ||
                     (uint32_t)(r.wv >> 32) != epoch);
    return pend ? v_edge : (inter ? v_inter : d);
}

__device__ __forceinline__ uint64_t cell_gather(const uint64_t *edge, uint32_t epoch, uint32_t self,
                                                const DFrame &F, int x, int y, int bh, int kind,
                                                int q, bool pend, bool inter, uint64_t d,
                                                bool &stale)
{
    const CellRaw r = cell_load(edge, self, F, x, y, bh, kind, q, pend, inter, d);
    return cell_value(r, epoch, x, bh, kind, q, pend, inter, d, stale);
}

/* Synchronises the threads of one intra worker: a whole workgroup, or a
 * single wave (whose LDS accesses execute in order; the fences keep the
 * compiler from reordering them). Only LDS is ordered: threads of a worker
 * share nothing through global memory (other blocks read a block's samples
 * only through its edge words, each validated by its epoch), so the
 * workgroup barrier does not wait for the global stores in flight. */
template <int W>
__device__ __forceinline__ void worker_sync()
{
    if constexpr (W == 64) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }
}

struct IntraShared {
    uint8_t nb[3][2][64];                 /* above / left samples per plane */
    __align__(4) uint8_t edge[3][2][64];  /* final bottom row / right column */
    uint32_t sum[3];
    uint32_t ticket;
    uint32_t job;
    uint8_t tl[3];
};

/* One intra block: entry tk of job J's intra list, W threads (tid) sharing
 * S. EARLY publishes the edge record before writing the block (for workers
 * that others wait on); otherwise the edges are collected while the block
 * is written and published at the end. DEFER (with EARLY) writes no sample
 * of the block: after publishing the edge record it leaves the block's
 * resolved neighbour cells and DC values in the block's neighbour record,
 * from which k_intra_px later writes every intra block at once. */
/* A job's frame epoch. The per-stage kernels read it from the descriptor
 * (written by the batch's first kernel, an earlier launch); inside the
 * fused kernel another workgroup of the same launch writes it before a grid
 * barrier, so that kernel reads it with a vector load that bypasses the
 * scalar cache. */
template <bool FUSED>
__device__ __forceinline__ uint32_t job_epoch(const DJob &J)
{
    if constexpr (FUSED)
        return __hip_atomic_load(&((DJob &)J).epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return J.epoch;
}

/* An intra block's neighbour record as one lane holds it: its cell
 * (threads past the cells read the last one, and never use it), the
 * dependency masks, the block's local index and its block record. */
struct IntraRec {
    uint64_t cell, dmask, imask;
    uint64_t kw[3];
    uint32_t b;
};

/* The job fields the intra stage reads, held in registers by a chain
 * worker for its whole chain (through the descriptor, every store would make
 * the compiler load them again: a store could alias it as far as it knows). */
struct JobView {
    DFrame out;
    int16_t *resid;
    uint64_t *edge, *nbr;
    uint32_t *err, *err_host;
};

__device__ __forceinline__ JobView job_view(const DJob &J)
{
    JobView v;
    v.out = J.out;
    v.resid = J.resid;
    v.edge = J.edge;
    v.nbr = J.nbr;
    v.err = J.err;
    v.err_host = J.err_host;
    return v;
}

template <typename JT>
__device__ __forceinline__ IntraRec load_intra_rec(const JT &J, uint32_t tk, int tid)
{
    const GMEM uint64_t *rec = gm(J.nbr) + (size_t)tk * CVP9_NB_WORDS;
    IntraRec R;
    R.cell = rec[min(tid, CVP9_NB_CELLS - 1)];
    R.dmask = rec[CVP9_NB_MASK];
    R.imask = rec[CVP9_NB_IMASK];
    R.b = (uint32_t)rec[CVP9_NB_INDEX];
    R.kw[0] = rec[CVP9_NB_BLOCK];
    R.kw[1] = rec[CVP9_NB_BLOCK + 1];
    R.kw[2] = rec[CVP9_NB_BLOCK + 2];
    return R;
}

/* Where this lane's cell of a block lies and what it is (thread c of the
 * gathering wave): kind 0 above, cell q, 1 left, cell q, 2 above-left;
 * pend: an earlier intra block's (edge record), inter: an inter block's
 * (the frame), else the resolved value d. */
struct CellPos {
    int kind, q;
    bool pend, inter;
    uint64_t d;
};

__device__ __forceinline__ CellPos cell_pos(const IntraRec &R, const DBlock &k, int c)
{
    const int x = k.x, y = k.y, ncx = (k.bw + 3) >> 2, ncy = (k.bh + 3) >> 2;
    CellPos p;
    p.kind = c < ncx ? 0 : (c < ncx + ncy ? 1 : 2);
    p.q = p.kind == 0 ? c : c - ncx;
    const bool present = p.kind == 0 ? y > 0 : (p.kind == 1 ? x > 0 : (x > 0 && y > 0));
    const bool incell = c <= ncx + ncy && present;
    p.pend = incell && ((R.dmask >> c) & 1);
    p.inter = incell && !p.pend && ((R.imask >> c) & 1);
    p.d = present ? R.cell : 0x7f7f7f7f7f7f7f7full;
    return p;
}

/* The luma residual of edge sample i of a block (bottom row, then right
 * column; 0 where the block has none), read at a clamped place when there
 * is none so that the load needs no branch. */
__device__ __forceinline__ int edge_resid(const GMEM int16_t *jres, const DBlock &k, int i)
{
    const int bw = k.bw, bh = k.bh, lt = k.tx + 2, T = 1 << lt;
    const bool res = !(k.flags & DB_SKIP);
    const int rrows = res ? (bh >> lt) << lt : 0, rcols = res ? (int)k.nx << lt : 0;
    const int er = i < bw ? bh - 1 : i - bw, ec = i < bw ? i : bw - 1;
    const bool ok = i < bw + bh && er < rrows && ec < rcols;
    const GMEM int16_t *p = ok ? jres + k.coeff + (size_t)(((er >> lt) * (int)k.nx + (ec >> lt)) << (2 * lt)) +
                                     ((er & (T - 1)) << lt) + (ec & (T - 1))
                               : jres;
    return ok ? (int)p[0] : 0;
}

/* What a chain worker loads for its next block before it finishes the
 * current one: the record, this lane's two luma edge residuals and its
 * cell's loads (the cells of the current block, the next one's chain
 * predecessor, are taken from registers instead). */
struct BlockPre {
    IntraRec R;
    int eres0, eres1;
    CellRaw raw;
};

template <typename JT>
__device__ __forceinline__ BlockPre issue_block(const JT &J, const IntraRec &R, int tid)
{
    BlockPre P;
    P.R = R;
    DBlock k;
    memcpy(&k, R.kw, sizeof(k));
    const GMEM int16_t *jres = gm(J.resid);
    P.eres0 = edge_resid(jres, k, tid);
    P.eres1 = edge_resid(jres, k, tid + 64);
    const CellPos cp = cell_pos(R, k, tid);
    P.raw = cell_load(J.edge, R.b, J.out, k.x, k.y, k.bh, cp.kind, cp.q, cp.pend, cp.inter, cp.d);
    return P;
}

struct NoHook {
    __device__ void operator()() const {}
};

template <int W, bool EARLY, bool DEFER = false, bool CHAIN = false, typename JT = DJob,
          typename HOOK = NoHook>
__device__ __forceinline__ void intra_block(const DBatch &B, const JT &J, uint32_t tk,
                                            IntraShared &S, int tid, uint32_t epoch,
                                            const BlockPre *pre = nullptr,
                                            uint32_t prev_b = 0xffffffffu, uint32_t prev_word = 0,
                                            uint32_t *my_word = nullptr, uint32_t *my_b = nullptr,
                                            const HOOK &after_gather = HOOK())
{
    static_assert(EARLY || !DEFER, "a deferred block publishes its edges first");
    static_assert(!CHAIN || (W == 64 && EARLY), "chain workers are single waves");
    uint32_t myw = 0; /* CHAIN: edge word tid of this block (its 4 samples) */
    /* the job's fields, loaded once, together: after any store the
     * compiler would otherwise reload them (a store could alias the
     * descriptor as far as it knows), one memory round trip each */
    const DFrame out = J.out;
    const GMEM int16_t *const jres = gm(J.resid);
    uint64_t *const jedge = J.edge;
    uint32_t *const jerr = J.err, *const jerr_host = J.err_host;
    /* block index, block record, dependency mask and this thread's cell
     * all come from the neighbour record, in one round of loads (a chain
     * worker loaded it one block ahead) */
    const GMEM uint64_t *rec = gm(J.nbr) + (size_t)tk * CVP9_NB_WORDS;
    IntraRec R;
    if constexpr (CHAIN) R = pre->R;
    else R = load_intra_rec(J, tk, tid);
    const uint64_t cell = R.cell, dmask = R.dmask, imask = R.imask;
    const uint32_t b = R.b;
    DBlock k;
    memcpy(&k, R.kw, sizeof(k));
    const int x = k.x, y = k.y, bw = k.bw, bh = k.bh;
    const int cbx = x >> 1, cby = y >> 1, cbw = (bw + 1) >> 1, cbh = (bh + 1) >> 1;

    /* residual geometry; this thread's first luma quad of residual is
     * loaded now, so it arrives while the neighbours are gathered. A
     * quad with residual lies inside one transform row (c0 is a multiple
     * of 4, transforms are at least 4 wide). */
    const int lt = k.tx + 2, T = 1 << lt;
    const bool res = !(k.flags & DB_SKIP);
    const int rrows = res ? (bh >> lt) << lt : 0, rcols = res ? (int)k.nx << lt : 0;
    const int nq = (bw + 3) >> 2;
    auto resid_quad = [&](int i) -> const GMEM int16_t * {
        const int r = (int)udiv_small((uint32_t)i, (uint32_t)nq), c0 = (i - r * nq) * 4;
        if (r >= rrows || c0 >= rcols) return nullptr;
        return jres + k.coeff + (size_t)(((r >> lt) * (int)k.nx + (c0 >> lt)) << (2 * lt)) +
               ((r & (T - 1)) << lt) + (c0 & (T - 1));
    };
    /* this thread's first luma quad of residual, loaded with no branch
     * around it (a thread without residual reads the neighbour record and
     * ignores it): a branch would wait for it */
    const GMEM int16_t *rs0 = !DEFER && tid < nq * bh ? resid_quad(tid) : nullptr;
    const GMEM int16_t *rp0 = rs0 ? rs0 : (const GMEM int16_t *)rec;
    int16_t rq0[4];
#pragma unroll
    for (int e = 0; e < 4; e++) rq0[e] = rp0[e];
    /* residual at one luma sample (0 where the block has none) */
    auto resid_at = [&](int r, int c) -> int {
        if (r >= rrows || c >= rcols) return 0;
        return jres[k.coeff + (size_t)(((r >> lt) * (int)k.nx + (c >> lt)) << (2 * lt)) +
                       ((r & (T - 1)) << lt) + (c & (T - 1))];
    };
    /* and this thread's first luma edge sample: bottom row, then right
     * column */
    const int er = tid < bw ? bh - 1 : tid - bw, ec = tid < bw ? tid : bw - 1;
    const bool eres_ok = EARLY && tid < bw + bh && er < rrows && ec < rcols;
    const int eres0 =
        CHAIN ? pre->eres0 : !EARLY ? 0
               : (int)(eres_ok ? jres + k.coeff +
                                     (size_t)(((er >> lt) * (int)k.nx + (ec >> lt)) << (2 * lt)) +
                                     ((er & (T - 1)) << lt) + (ec & (T - 1))
                               : (const GMEM int16_t *)rec)[0];
    /* a 64-thread worker has a second luma edge sample (sample tid + 64 of
     * at most 128) */
    const int i1 = tid + 64;
    const int er1 = i1 < bw ? bh - 1 : i1 - bw, ec1 = i1 < bw ? i1 : bw - 1;
    const bool eres1_ok = EARLY && W == 64 && i1 < bw + bh && er1 < rrows && ec1 < rcols;
    const int eres1 =
        CHAIN ? pre->eres1 : !(EARLY && W == 64) ? 0
               : (int)(eres1_ok ? jres + k.coeff +
                                      (size_t)(((er1 >> lt) * (int)k.nx + (ec1 >> lt)) << (2 * lt)) +
                                      ((er1 & (T - 1)) << lt) + (ec1 & (T - 1))
                                : (const GMEM int16_t *)rec)[0];

    /* 1. neighbours: 127 outside the frame; resolved cells as recorded by
     *    the neighbour stage; cells of earlier intra blocks from their
     *    edge records. ncx + ncy + 1 <= 33 cells: one per thread. */
    const int ncx = (bw + 3) >> 2, ncy = (bh + 3) >> 2;
    /* edge samples of Y, U and V; a block with at most 64 of them (up to
     * 16x16) computes and publishes them inside the gathering wave, from
     * its registers, with no barrier in between */
    const int ne = bw + bh + 2 * (cbw + cbh);
    const bool fast = EARLY && ne <= 64;
    if (tid < 64) { /* the worker's first wave, whole */
        const int c = tid;
        const int kind = c < ncx ? 0 : (c < ncx + ncy ? 1 : 2);
        const bool present = kind == 0 ? y > 0 : (kind == 1 ? x > 0 : (x > 0 && y > 0));
        const bool incell = c <= ncx + ncy && present;
        const bool pend = incell && ((dmask >> c) & 1);
        const bool inter = incell && !pend && ((imask >> c) & 1);
        bool stale;
        uint64_t v;
        if constexpr (CHAIN)
            v = cell_value(pre->raw, epoch, x, bh, kind, kind == 0 ? c : c - ncx, pend, inter,
                           present ? cell : 0x7f7f7f7f7f7f7f7full, stale);
        else
            v = cell_gather(jedge, epoch, b, out, x, y, bh, kind, kind == 0 ? c : c - ncx, pend,
                            inter, present ? cell : 0x7f7f7f7f7f7f7f7full, stale);
        if constexpr (CHAIN) {
            /* cells of the block this wave finished just before (its chain
             * predecessor): from the lanes holding that block's edge words
             * (lane t word t), with no memory access and no wait */
            const uint32_t side = (uint32_t)(cell >> 32) & 1, lw = (uint32_t)(cell >> 33) & 15,
                           cc = (uint32_t)(cell >> 40) & 31;
            const uint32_t ly = lane_read(4u * ((side ? 16u : 0u) + lw), prev_word);
            const uint32_t lu = lane_read(4u * ((side ? 40u : 32u) + (cc >> 2)), prev_word);
            const uint32_t lv = lane_read(4u * ((side ? 56u : 48u) + (cc >> 2)), prev_word);
            if (pend && (uint32_t)cell == prev_b) {
                const uint32_t sh = 8 * (cc & 3);
                v = (uint64_t)ly | (uint64_t)((lu >> sh) & 0xffff) << 32 |
                    (uint64_t)((lv >> sh) & 0xffff) << 48;
                stale = false;
            }
        }
        /* the edge residuals are in registers before any wait: issued at
         * the block's start, they would otherwise be loaded where they are
         * used, a memory round trip after the last neighbour arrives */
        asm volatile("" ::"v"(eres0), "v"(eres1));

        /* Everything of the DC sums and of the edge samples that does not
         * depend on the neighbour values is worked out here, before any
         * wait: after the last neighbour arrives, a block on a dependency
         * chain then runs only the arithmetic on the values (a wave issues
         * one instruction per four clocks, and the index work had doubled
         * the length of every chain link).
         * DC sums over the w above and h left samples of each plane
         * (vp9_reconstruct.c:67-90): byte masks of the samples this lane's
         * cell contributes, luma in the low word, U and V in the high word
         * (bytes 0-1 and 2-3). */
        uint32_t mY = 0, mUV = 0;
        if (c <= ncx + ncy && kind != 2) {
            const int q = kind == 0 ? c : c - ncx;
            const int ly = kind == 0 ? bw : bh, lc = kind == 0 ? cbw : cbh;
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (4 * q + e < ly) mY |= 0xffu << (8 * e);
#pragma unroll
            for (int e = 0; e < 2; e++)
                if (2 * q + e < lc) mUV |= 0x00ff00ffu << (8 * e);
        }
        /* edge sample i = tid (fast blocks), numbered as in the edge loop
         * below: Y bottom row, Y right column, then U and V likewise. Its
         * above, left and above-left samples come from the lanes holding
         * those cells (an above or left cell is 4 luma or 2 chroma samples
         * wide; bytes 0-3 luma, 4-5 U, 6-7 V; the above-left cell's samples
         * in bytes 3, 5, 7): their lanes as byte addresses, the shifts that
         * pick the sample, the plane. Edge word w = tid: the lanes holding
         * its four samples and the mask of those inside the edge. */
        uint32_t f_adr = 0, f_sh = 0, w_adr = 0, w_msk = 0;
        int f_mode = 0, f_add = 0;
        if (EARLY && fast) {
            const int i = tid;
            int p, dir, e;
            if (i < bw + bh) { p = 0; dir = i >= bw; e = dir ? i - bw : i; }
            else {
                const int q = i - bw - bh;
                p = q < cbw + cbh ? 1 : 2;
                const int u = q - (p - 1) * (cbw + cbh);
                dir = u >= cbw;
                e = dir ? u - cbw : u;
            }
            const int r = p == 0 ? (dir ? e : bh - 1) : (dir ? e : cbh - 1);
            const int cc = p == 0 ? (dir ? bw - 1 : e) : (dir ? cbw - 1 : e);
            const uint32_t qa = (uint32_t)min(p == 0 ? cc >> 2 : cc >> 1, 63);
            const uint32_t ql = (uint32_t)min(ncx + (p == 0 ? r >> 2 : r >> 1), 63);
            const uint32_t qt = (uint32_t)(ncx + ncy);
            const uint32_t sa = p == 0 ? 8 * (cc & 3) : 8 * (cc & 1) + (p == 2 ? 16 : 0);
            const uint32_t sl = p == 0 ? 8 * (r & 3) : 8 * (r & 1) + (p == 2 ? 16 : 0);
            const uint32_t st = p == 1 ? 8 : 24;
            f_adr = qa << 2 | ql << 10 | qt << 18;
            f_sh = sa | sl << 8 | st << 16 | (uint32_t)p << 24;
            f_mode = p == 0 ? k.ymode : k.uvmode;
            f_add = p == 0 && eres_ok ? eres0 : 0;
            const int w = tid;
            int wp, wdir, wi, len;
            if (w < 32) { wp = 0; wdir = w >= 16; wi = w & 15; len = wdir ? bh : bw; }
            else { wp = (w < 48) ? 1 : 2; wdir = (w & 15) >= 8; wi = w & 7; len = wdir ? cbh : cbw; }
            const int base = wp == 0 ? (wdir ? bw : 0) : bw + bh + (wp - 1) * (cbw + cbh) + (wdir ? cbw : 0);
            w_adr = (uint32_t)(base + 4 * wi) << 2;
            const int n = len - 4 * wi;
            w_msk = n >= 4 ? ~0u : (n > 0 ? (1u << (8 * n)) - 1 : 0u);
        }
        const bool dc_y = k.ymode == 0, dc_uv = k.uvmode == 0;
        const int nY = bw + bh, nC = cbw + cbh;
        const float rY = __builtin_amdgcn_rcpf((float)nY), rC = __builtin_amdgcn_rcpf((float)nC);
        asm volatile("" ::"v"(mY), "v"(mUV), "v"(f_adr), "v"(f_sh), "v"(w_adr), "v"(w_msk), "v"(rY),
                     "v"(rC));

        /* a neighbour still running: poll the cell's words until it has
         * published them */
        if (stale) v = cell_from_edges(jedge, epoch, jerr, jerr_host, cell, B.spin_limit);
        /* (a chain worker issues its next block's loads here: the loads it
         * waits for above were issued a block earlier, and what it loads
         * now is in flight while this block is computed and published) */
        after_gather();

        /* the sums, reduced across the wave (only for a DC mode: the other
         * predictors do not read them) */
        const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
        const uint32_t huv = hi & mUV;
        uint32_t syu = __builtin_amdgcn_sad_u8(lo & mY, 0u, 0u) |
                       __builtin_amdgcn_sad_u8(huv & 0xffffu, 0u, 0u) << 16;
        uint32_t sv = __builtin_amdgcn_sad_u8(huv >> 16, 0u, 0u);
        if (dc_y || dc_uv) {
            syu = wave_sum(syu);
            sv = wave_sum(sv);
        }
        if (EARLY && fast) {
            const uint32_t p = f_sh >> 24;
            /* both words of each source lane: which one a lane needs
             * depends on its own plane */
            const uint32_t a_at = f_adr & 0xff, l_at = (f_adr >> 8) & 0xff, t_at = (f_adr >> 16) & 0xff;
            const uint32_t alo = lane_read(a_at, lo), ahi = lane_read(a_at, hi);
            const uint32_t llo = lane_read(l_at, lo), lhi = lane_read(l_at, hi);
            const uint32_t tlo = lane_read(t_at, lo), thi = lane_read(t_at, hi);
            const uint32_t aw = p ? ahi : alo, lw = p ? lhi : llo, tw = p ? thi : tlo;
            const int a = (int)((aw >> (f_sh & 31)) & 0xff);
            const int l = (int)((lw >> ((f_sh >> 8) & 31)) & 0xff);
            const int t = (int)((tw >> ((f_sh >> 16) & 31)) & 0xff);
            int dc = 0;
            if (p == 0 ? dc_y : dc_uv) {
                const uint32_t sum = p == 0 ? syu & 0xffff : (p == 1 ? syu >> 16 : sv);
                const uint32_t n = p == 0 ? (uint32_t)nY : (uint32_t)nC;
                dc = (int)(uint8_t)udiv_small_r(sum + (n >> 1), n, p == 0 ? rY : rC);
            }
            const int es = clampi(intra_pred(f_mode, a, l, t, dc) + f_add, 0, 255);
            uint32_t word = 0;
#pragma unroll
            for (int e2 = 0; e2 < 4; e2++)
                word |= (lane_read(w_adr + 4 * e2, (uint32_t)es) & 0xff) << (8 * e2);
            if (w_msk)
                __hip_atomic_store(gm(jedge) + (size_t)b * CVP9_EDGE_WORDS + tid,
                                   (uint64_t)epoch << 32 | (word & w_msk), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            myw = word & w_msk;
        }
        if constexpr (DEFER) {
            /* the samples of the block are written later, by k_intra_px,
             * from its resolved cells and DC values, kept in the block's
             * neighbour record (which nothing else reads after this) */
            GMEM uint64_t *recw = gm(J.nbr) + (size_t)tk * CVP9_NB_WORDS;
            if (c <= ncx + ncy) recw[c] = v;
            if (c == 0) {
                const uint32_t dY = dc_y ? udiv_small_r((syu & 0xffff) + (uint32_t)(nY >> 1), (uint32_t)nY, rY) : 0u;
                const uint32_t dU = dc_uv ? udiv_small_r((syu >> 16) + (uint32_t)(nC >> 1), (uint32_t)nC, rC) : 0u;
                const uint32_t dV = dc_uv ? udiv_small_r(sv + (uint32_t)(nC >> 1), (uint32_t)nC, rC) : 0u;
                recw[CVP9_NB_MASK] = (uint64_t)((dY & 0xff) | (dU & 0xff) << 8 | (dV & 0xff) << 16);
            }
        }
        if (c == 0 && !(DEFER && fast)) {
            S.sum[0] = syu & 0xffff;
            S.sum[1] = syu >> 16;
            S.sum[2] = sv;
        }
        if (c > ncx + ncy || (DEFER && fast)) {
            /* no cell, or no LDS copy needed (the edges are published) */
        } else if (kind == 2) {
            S.tl[0] = (uint8_t)(v >> 24);
            S.tl[1] = (uint8_t)(v >> 40);
            S.tl[2] = (uint8_t)(v >> 56);
        } else {
            const int q = kind == 0 ? c : c - ncx;
#pragma unroll
            for (int e = 0; e < 4; e++) S.nb[0][kind][4 * q + e] = (uint8_t)(v >> (8 * e));
#pragma unroll
            for (int e = 0; e < 2; e++) {
                S.nb[1][kind][2 * q + e] = (uint8_t)(v >> (32 + 8 * e));
                S.nb[2][kind][2 * q + e] = (uint8_t)(v >> (48 + 8 * e));
            }
        }
    } else {
        asm volatile("" ::"v"(eres0)); /* likewise, while the first wave gathers */
    }
    worker_sync<W>();

    const int dc0 = (int)(uint8_t)udiv_small(S.sum[0] + (uint32_t)((bw + bh) >> 1), (uint32_t)(bw + bh));
    if (EARLY && !fast) {
        /* 2. with EARLY, the edge samples first (bottom row and right
         *    column of Y, U and V), published at once, so blocks waiting on
         *    this one go on while its interior is written */
        for (int i = tid; i < ne; i += W) {
            int p, dir, e;
            if (i < bw + bh) { p = 0; dir = i >= bw; e = dir ? i - bw : i; }
            else {
                const int q = i - bw - bh;
                p = q < cbw + cbh ? 1 : 2;
                const int u = q - (p - 1) * (cbw + cbh);
                dir = u >= cbw;
                e = dir ? u - cbw : u;
            }
            int v;
            if (p == 0) {
                const int r = dir ? e : bh - 1, c = dir ? bw - 1 : e;
                v = intra_pred(k.ymode, S.nb[0][0][c], S.nb[0][1][r], S.tl[0], dc0);
                const int rv = i == tid ? (eres_ok ? eres0 : 0)
                               : (W == 64 && i == i1 ? (eres1_ok ? eres1 : 0) : resid_at(r, c));
                v = clampi(v + rv, 0, 255);
            } else {
                const int r = dir ? e : cbh - 1, c = dir ? cbw - 1 : e;
                const int dcc = (int)(uint8_t)udiv_small(S.sum[p] + (uint32_t)((cbw + cbh) >> 1),
                                                         (uint32_t)(cbw + cbh));
                v = intra_pred(k.uvmode, S.nb[p][0][c], S.nb[p][1][r], S.tl[p], dcc);
            }
            S.edge[p][dir][e] = (uint8_t)v;
        }
        worker_sync<W>();

        /* 3. publish the edge record: Y bottom, Y right, U bottom, U right,
         *    V bottom, V right */
        for (int t = tid; t < CVP9_EDGE_WORDS; t += W) {
            int p, dir, wi, len;
            if (t < 32) { p = 0; dir = t >= 16; wi = t & 15; len = dir ? bh : bw; }
            else { p = (t < 48) ? 1 : 2; dir = (t & 15) >= 8; wi = t & 7; len = dir ? cbh : cbw; }
            if (4 * wi < len) {
                /* four samples in one LDS read (the edge rows are 4-aligned),
                 * those past the edge's length zeroed */
                const int n = len - 4 * wi;
                const uint32_t q = *(const uint32_t *)&S.edge[p][dir][4 * wi];
                const uint64_t d = (uint64_t)epoch << 32 | (n >= 4 ? q : q & ((1u << (8 * n)) - 1));
                __hip_atomic_store(gm(jedge) + (size_t)b * CVP9_EDGE_WORDS + t, d, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                myw = (uint32_t)d;
            }
        }
    }
    if constexpr (CHAIN) {
        *my_word = myw;
        *my_b = b;
    }

    if constexpr (DEFER) return; /* k_intra_px writes the block */

    /* 4. the whole block: prediction + residual, 4 luma samples per item,
     *    four items' residuals loaded before any is used (so a 64-thread
     *    worker on a large block waits once per four items, not per item);
     *    chroma prediction only (the reference adds no chroma residual), 4
     *    samples per item. Without EARLY the edge samples are collected
     *    here and published after. */
    {
        const int tl = S.tl[0], mode = k.ymode;
        uint8_t *Y = out.y + (size_t)y * out.sy + x;
        const int nitems = nq * bh;
        for (int i0 = tid; i0 < nitems; i0 += 4 * W) {
            int16_t rq[4][4];
            bool has[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int i = i0 + u * W;
                const bool first = i == tid; /* loaded before the gather */
                const GMEM int16_t *rs = first ? rs0 : (i < nitems ? resid_quad(i) : nullptr);
                has[u] = rs != nullptr;
#pragma unroll
                for (int e = 0; e < 4; e++) rq[u][e] = first ? rq0[e] : (rs ? rs[e] : (int16_t)0);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int i = i0 + u * W;
                if (i >= nitems) break;
                const int r = (int)udiv_small((uint32_t)i, (uint32_t)nq), c0 = (i - r * nq) * 4;
                const int l = S.nb[0][1][r];
                uint32_t word = 0;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int c = c0 + e;
                    if (c < bw) {
                        int v = intra_pred(mode, S.nb[0][0][c], l, tl, dc0);
                        if (has[u]) v = clampi(v + rq[u][e], 0, 255);
                        word |= (uint32_t)v << (8 * e);
                        if constexpr (!EARLY) {
                            if (r == bh - 1) S.edge[0][0][c] = (uint8_t)v;
                            if (c == bw - 1) S.edge[0][1][r] = (uint8_t)v;
                        }
                    }
                }
                store_quad(Y + (size_t)r * out.sy + c0, word, bw - c0);
            }
        }
        const int nqc = (cbw + 3) >> 2;
        for (int i = tid; i < 2 * nqc * cbh; i += W) {
            const int p = i < nqc * cbh ? 1 : 2;
            const int q = i - (p - 1) * nqc * cbh;
            const int r = (int)udiv_small((uint32_t)q, (uint32_t)nqc), c0 = (q - r * nqc) * 4;
            const int dcc = (int)(uint8_t)udiv_small(S.sum[p] + (uint32_t)((cbw + cbh) >> 1),
                                                     (uint32_t)(cbw + cbh));
            const int l = S.nb[p][1][r];
            uint32_t word = 0;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int c = c0 + e;
                if (c < cbw) {
                    const int v = intra_pred(k.uvmode, S.nb[p][0][c], l, S.tl[p], dcc);
                    word |= (uint32_t)v << (8 * e);
                    if constexpr (!EARLY) {
                        if (r == cbh - 1) S.edge[p][0][c] = (uint8_t)v;
                        if (c == cbw - 1) S.edge[p][1][r] = (uint8_t)v;
                    }
                }
            }
            uint8_t *pl = p == 1 ? out.u : out.v;
            store_quad(pl + (size_t)(cby + r) * out.suv + cbx + c0, word, cbw - c0);
        }
    }
    if constexpr (!EARLY) {
        worker_sync<W>();

        /* 5. publish the edge record: Y bottom, Y right, U bottom, U right,
         *    V bottom, V right */
        for (int t = tid; t < CVP9_EDGE_WORDS; t += W) {
            int p, dir, wi, len;
            if (t < 32) { p = 0; dir = t >= 16; wi = t & 15; len = dir ? bh : bw; }
            else { p = (t < 48) ? 1 : 2; dir = (t & 15) >= 8; wi = t & 7; len = dir ? cbh : cbw; }
            if (4 * wi < len) {
                /* four samples in one LDS read (the edge rows are 4-aligned),
                 * those past the edge's length zeroed */
                const int n = len - 4 * wi;
                const uint32_t q = *(const uint32_t *)&S.edge[p][dir][4 * wi];
                const uint64_t d = (uint64_t)epoch << 32 | (n >= 4 ? q : q & ((1u << (8 * n)) - 1));
                __hip_atomic_store(gm(jedge) + (size_t)b * CVP9_EDGE_WORDS + t, d, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

/* Persistent workers are spread over the batch's jobs in proportion to
 * their intra blocks: worker w starts on the job holding intra block
 * floor(w * total / nwk) of the batch, so job j's workers are
 * [first(j), first(j + 1)) with first(j) = ceil(pre_intra[j] * nwk / total). */
__device__ __forceinline__ uint32_t job_first_worker(const DBatch &B, uint32_t j, uint32_t nwk,
                                                     uint32_t total)
{
    return (uint32_t)(((uint64_t)B.pre_intra[j] * nwk + total - 1) / total);
}

/* A job of the batch that no worker started on and that has intra tickets
 * left, looked for by one whole wave from job j + 1 on, round robin;
 * B.njobs when there is none. Each lane reads one job's ticket counter: a
 * plain load, so workers that run out of work do not queue up on the
 * counters' atomics. */
__device__ __forceinline__ uint32_t find_open_job(const DBatch &B, uint32_t j, uint32_t nwk,
                                                  uint32_t total)
{
    const uint32_t lane = __lane_id();
    for (uint32_t k = 0; k < B.njobs; k += 64) {
        uint32_t idx = j + 1 + k + lane;
        while (idx >= B.njobs) idx -= B.njobs;
        bool open = false;
        if (k + lane < B.njobs)
            open = job_first_worker(B, idx, nwk, total) == job_first_worker(B, idx + 1, nwk, total) &&
                   B.jstart[idx] +
                           __hip_atomic_load(B.jticket + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                       B.jobs[idx].n_intra;
        const uint64_t m = __ballot(open);
        if (m) return __shfl(idx, __builtin_ctzll(m));
    }
    return B.njobs;
}

/* Persistent workers, W threads each: 256 (one workgroup per block) or 64
 * (one wave per block, four independent workers per workgroup); this is
 * worker wk of nwk.
 *
 * A job's remaining intra-list entries [jstart, n_intra) are dealt out to
 * its c workers by stride: worker r of them takes entries jstart + r,
 * jstart + r + c, ... in that order. Nothing is fetched through an atomic
 * counter: a counter on one address hands out only some tens of entries
 * per microsecond, which held the intra stage's first levels back. The
 * lowest unfinished entry of a job always runs: its worker has finished
 * every earlier entry it owns, and its dependencies are lower entries,
 * finished too (every worker of the launch becomes resident: the launch is
 * held to what the device holds, see intra_resident_wgs). A job that no
 * worker started on (its share of the batch's intra blocks is below one
 * worker) is served through its ticket counter by the workers that have
 * finished their own entries: the tickets a worker holds are its current
 * one and a later one of the same job. */
template <int W, bool FUSED, bool DEFER = false>
__device__ __forceinline__ void intra_worker(const DBatch &B, IntraShared *shm, uint32_t wk,
                                             uint32_t nwk)
{
    IntraShared &S = shm[threadIdx.x / W];
    const int tid = threadIdx.x % W;
    const uint32_t total = B.pre_intra[B.njobs];

    /* Job, entry and stride are made wave-uniform with readfirstlane, so
     * every loop test is a scalar branch. With an LDS value as the exit
     * test the compiler treats the exit as divergent and restructures the
     * loop so that thread 0's ticket fetch runs outside the loop holding the
     * barriers: the other lanes then repeat one ticket forever. */
    uint32_t j = __builtin_amdgcn_readfirstlane(
        find_job(B.pre_intra, B.njobs, (uint32_t)((uint64_t)wk * total / nwk)));
    const uint32_t f0 = job_first_worker(B, j, nwk, total);
    uint32_t stride = __builtin_amdgcn_readfirstlane(job_first_worker(B, j + 1, nwk, total) - f0);
    uint32_t tk = __builtin_amdgcn_readfirstlane(B.jstart[j] + (wk - f0));
    bool counted = false; /* entries from the job's ticket counter */
    for (;;) {
        const DJob &J = B.jobs[j];
        if (tk >= J.n_intra) {
            /* this worker's entries of the job are done: go on with a job
             * that no worker started on and has tickets left (the first
             * sync: every thread has read S.ticket) */
            worker_sync<W>();
            if (tid < 64) {
                const uint32_t nj = find_open_job(B, j, nwk, total);
                if (tid == 0) {
                    S.job = nj;
                    if (nj < B.njobs) S.ticket = atomicAdd(B.jticket + nj, 1u);
                }
            }
            worker_sync<W>();
            j = __builtin_amdgcn_readfirstlane(S.job);
            if (j >= B.njobs) break;
            tk = __builtin_amdgcn_readfirstlane(B.jstart[j] + S.ticket);
            counted = true;
            continue;
        }
        /* with a counted job, the next ticket is fetched while this block
         * runs (atomicInc: the compiler's wave aggregation of atomicAdd
         * would wait for the result right here; tickets never reach the
         * wrap limit) */
        uint32_t next = 0;
        if (counted && tid == 0) next = atomicInc(B.jticket + j, 0xffffffffu);
        intra_block<W, true, DEFER>(B, J, tk, S, tid, job_epoch<FUSED>(J));

        /* next entry; the sync also orders LDS reuse */
        if (tid == 0) S.ticket = next;
        worker_sync<W>();
        tk = counted ? __builtin_amdgcn_readfirstlane(B.jstart[j] + S.ticket) : tk + stride;
    }
}

template <int W, bool DEFER>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_intra(DBatch B)
{
    __shared__ IntraShared shm[256 / W];
    intra_worker<W, false, DEFER>(B, shm, blockIdx.x * (256 / W) + threadIdx.x / W,
                                  gridDim.x * (256 / W));
}

/* Chain workers (a batch of one job whose frame lists chains of intra
 * blocks): worker w < K takes chain w in order, each block right after its
 * chain predecessor on the same wave, which hands that predecessor's edge
 * words over in registers (no memory round trip and no poll on the link);
 * a chain block's samples are left to k_intra_px (deferred form), so the
 * next link does not wait for them. The other workers deal the remaining
 * entries by stride and write their blocks, as the persistent stage does. A chain rises in level block by block and the
 * remaining entries are in list order, so every worker meets levels in
 * non-decreasing order and, with every worker resident, every wait ends. */
__device__ __forceinline__ void intra_chain_worker(const DBatch &B, IntraShared *shm, uint32_t wk,
                                                   uint32_t nwk)
{
    const DJob &J = B.jobs[0];
    IntraShared &S = shm[threadIdx.x / 64];
    const int tid = threadIdx.x % 64;
    const uint32_t epoch = J.epoch;
    const GMEM uint32_t *ct = gm(J.chain);
    const uint32_t K = J.n_chains;
    const GMEM uint32_t *ent = ct + K + 1;
    if (wk < K) {
        const uint32_t a = __builtin_amdgcn_readfirstlane(ct[wk]);
        const uint32_t z = __builtin_amdgcn_readfirstlane(ct[wk + 1]);
        /* the job's fields in registers for the whole chain, and the
         * chain's entries one per lane (64 at a time) */
        const JobView V = job_view(J);
        const uint32_t lane = __lane_id();
        uint32_t base = a; /* ents holds entries [base, base + 64) */
        uint32_t ents = ent[min(base + lane, z - 1)];
        auto entry = [&](uint32_t e) -> uint32_t {
            return __builtin_amdgcn_readlane(ents, (int)((min(e, z - 1) - base) & 63));
        };
        /* software pipeline: block e's loads (record, edge residuals, cell
         * loads) were issued while block e - 1 ran, and block e + 1's
         * record while block e - 1 ran too, so a link waits only for the
         * predecessor's arithmetic (its edge words arrive in registers).
         * The loads past the chain's end repeat its last entry: no branch
         * around a load. */
        uint32_t tk = entry(a);
        BlockPre P = issue_block(V, load_intra_rec(V, tk, tid), tid);
        IntraRec Rn = load_intra_rec(V, entry(a + 1), tid);
        uint32_t prev_b = 0xffffffffu, word = 0;
        for (uint32_t e = a; e < z; e++) {
            const uint32_t tkn = entry(e + 1);
            uint32_t tkn2;
            if (e + 2 >= base + 64 && e + 2 < z) {
                /* a chain longer than 64: its next 64 entries (a uniform
                 * branch, taken once per 64 blocks) */
                ents = ent[min(e + 2 + lane, z - 1)];
                base = e + 2;
                tkn2 = __builtin_amdgcn_readlane(ents, 0);
            } else {
                tkn2 = entry(e + 2);
            }
            BlockPre Pn;
            auto next_loads = [&]() {
                Pn = issue_block(V, Rn, tid);
                Rn = load_intra_rec(V, tkn2, tid);
            };
            uint32_t nw = 0, nb = 0;
            intra_block<64, true, true, true>(B, V, tk, S, tid, epoch, &P, prev_b, word, &nw, &nb,
                                              next_loads);
            word = nw;
            prev_b = __builtin_amdgcn_readfirstlane(nb);
            P = Pn;
            tk = tkn;
            worker_sync<64>(); /* orders the LDS reuse of the next block */
        }
    } else {
        const uint32_t r0 = __builtin_amdgcn_readfirstlane(ct[K]);
        for (uint32_t e = r0 + (wk - K); e < J.n_intra; e += nwk - K) {
            const uint32_t tk = __builtin_amdgcn_readfirstlane(ent[e]);
            intra_block<64, true>(B, J, tk, S, tid, epoch);
            worker_sync<64>();
        }
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_intra_chain(DBatch B)
{
    __shared__ IntraShared shm[4];
    intra_chain_worker(B, shm, blockIdx.x * 4 + threadIdx.x / 64, gridDim.x * 4);
}

/* Leading dependency levels: one launch per level, one worker per intra
 * block of that level across all jobs. Every block a block depends on lies
 * on a lower level, finished by an earlier launch, so no worker waits (its
 * edge words already carry the frame's epoch). The first small_wgs
 * workgroups take the level's small blocks (at most 16x16: one pass of a
 * 64-thread worker), four to a workgroup; the rest take one large block
 * each on all 256 threads. */
template <bool DEFER>
__global__ __launch_bounds__(256) void k_intra_level(DBatch B, uint32_t level, uint32_t small_wgs)
{
    __shared__ IntraShared shm[4];
    const size_t row = (size_t)level * (B.njobs + 1);
    const uint32_t *ps = B.lvl_pre + row, *pl = B.lvl_pre2 + row, *start = B.lvl_start + row;
    if (blockIdx.x < small_wgs) {
        /* wave-uniform, so the job lookup runs on the whole wave */
        const uint32_t g = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + threadIdx.x / 64);
        if (g >= ps[B.njobs]) return; /* uniform per worker */
        const uint32_t j = find_job_wave(ps, B.njobs, g);
        intra_block<64, DEFER, DEFER>(B, B.jobs[j], start[j] + (g - ps[j]), shm[threadIdx.x / 64],
                                      threadIdx.x % 64, job_epoch<false>(B.jobs[j]));
    } else {
        const uint32_t g = blockIdx.x - small_wgs; /* < pl[njobs] by the grid */
        const uint32_t j = find_job_wave(pl, B.njobs, g);
        intra_block<256, DEFER, DEFER>(B, B.jobs[j], start[j] + (ps[j + 1] - ps[j]) + (g - pl[j]),
                                       shm[0], threadIdx.x, job_epoch<false>(B.jobs[j]));
    }
}

/* Samples of the intra blocks in the deferred form (Tune.intra_defer). The
 * ordered stage published each block's edge record and left its resolved
 * neighbour cells (bytes 0-3 luma, 4-5 U, 6-7 V; the above-left cell's
 * samples in bytes 3, 5, 7) and its DC values (Y, U, V in bytes 0-2 of word
 * CVP9_NB_MASK) in the block's neighbour record. Here every intra block of
 * the batch is written at once, one wave each: prediction
 * (vp9_reconstruct.c:63-91) plus the luma residual, clamped (:336-363);
 * chroma prediction only (:266-311). The other blocks of the frame read an
 * intra block's samples only through its edge record, so nothing waits for
 * this pass but deblocking. A sample's above and left neighbours come from
 * the lanes holding their cells (ds_bpermute, with the whole wave active);
 * the above-left samples and DC values are wave-uniform. */
/* Intra-list entry tk of job J on L lanes: L = 64 (a wave, any block) or
 * 16 (blocks of at most 8x8: four of them to a wave); the L lanes start at
 * lane gb of the wave and hold the block's neighbour cells. */
template <int L>
__device__ __forceinline__ void intra_px_entry(const DJob &J, uint32_t tk, uint32_t gb = 0)
{
    const DFrame out = J.out;
    const GMEM int16_t *const jres = gm(J.resid);
    const GMEM uint64_t *rec = gm(J.nbr) + (size_t)tk * CVP9_NB_WORDS;
    const uint32_t lane = __lane_id() - gb; /* within the group */
    const uint64_t cell = rec[min(lane, (uint32_t)CVP9_NB_CELLS - 1)];
    const uint32_t dcw = (uint32_t)rec[CVP9_NB_MASK];
    DBlock k;
    {
        const uint64_t kw[3] = {rec[CVP9_NB_BLOCK], rec[CVP9_NB_BLOCK + 1], rec[CVP9_NB_BLOCK + 2]};
        memcpy(&k, kw, sizeof(k));
    }
    const int x = k.x, y = k.y, bw = k.bw, bh = k.bh;
    const int cbx = x >> 1, cby = y >> 1, cbw = (bw + 1) >> 1, cbh = (bh + 1) >> 1;
    const int ncx = (bw + 3) >> 2, ncy = (bh + 3) >> 2;
    const uint32_t lo = (uint32_t)cell, hi = (uint32_t)(cell >> 32);
    uint32_t tlo, thi;
    if constexpr (L == 64) {
        const int tc = __builtin_amdgcn_readfirstlane(ncx + ncy);
        tlo = (uint32_t)__builtin_amdgcn_readlane((int)lo, tc);
        thi = (uint32_t)__builtin_amdgcn_readlane((int)hi, tc);
    } else {
        tlo = lane_read(4u * (gb + (uint32_t)(ncx + ncy)), lo);
        thi = lane_read(4u * (gb + (uint32_t)(ncx + ncy)), hi);
    }

    /* luma: 4 samples per item, four items' residuals loaded before any
     * is used; lanes past the block compute and store nothing useful */
    const int lt = k.tx + 2, T = 1 << lt;
    const bool res = !(k.flags & DB_SKIP);
    const int rrows = res ? (bh >> lt) << lt : 0, rcols = res ? (int)k.nx << lt : 0;
    const int nq = (bw + 3) >> 2, nitems = nq * bh;
    const int ymode = k.ymode, tlY = (int)(tlo >> 24), dcY = (int)(dcw & 0xff);
    uint8_t *const Y = out.y + (size_t)y * out.sy + x;
    /* a quad's 4 residuals are one 8-byte word when the block's coefficients
     * start on a multiple of 4 (its transforms' offsets then all do) */
    const bool ral = (k.coeff & 3) == 0;
    for (int i0 = 0; i0 < nitems; i0 += 4 * L) {
        int16_t rq[4][4];
        bool has[4];
        int rr[4], cc[4];
        const GMEM int16_t *rs[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = min(i0 + u * L + (int)lane, nitems - 1);
            const int r = (int)udiv_small((uint32_t)i, (uint32_t)nq), c0 = (i - r * nq) * 4;
            rr[u] = r;
            cc[u] = c0;
            has[u] = r < rrows && c0 < rcols;
            /* (a quad with residual lies inside one transform row) */
            rs[u] = has[u] ? jres + k.coeff + (size_t)(((r >> lt) * (int)k.nx + (c0 >> lt)) << (2 * lt)) +
                                 ((r & (T - 1)) << lt) + (c0 & (T - 1))
                           : jres;
        }
        /* all four items' loads inside one branch (a branch's loads are
         * waited for where it ends) */
        if (ral) {
            uint64_t w[4];
#pragma unroll
            for (int u = 0; u < 4; u++) w[u] = *(const GMEM uint64_t *)rs[u];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int e = 0; e < 4; e++) rq[u][e] = (int16_t)(uint16_t)(w[u] >> (16 * e));
        } else {
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int e = 0; e < 4; e++) rq[u][e] = rs[u][has[u] ? e : 0];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int r = rr[u], c0 = cc[u];
            const uint32_t aw = lane_read(4u * (gb + (uint32_t)(c0 >> 2)), lo);
            const uint32_t lw = lane_read(4u * (gb + (uint32_t)(ncx + (r >> 2))), lo);
            const int l = (int)((lw >> (8 * (r & 3))) & 0xff);
            uint32_t word = 0;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                int v = intra_pred(ymode, (int)((aw >> (8 * e)) & 0xff), l, tlY, dcY);
                if (has[u]) v = clampi(v + rq[u][e], 0, 255);
                word |= (uint32_t)v << (8 * e);
            }
            if (i0 + u * L + (int)lane < nitems) store_quad(Y + (size_t)r * out.sy + c0, word, bw - c0);
        }
    }

    /* chroma: U items then V items, 4 samples each */
    const int nqc = (cbw + 3) >> 2, nci = nqc * cbh;
    const int uvmode = k.uvmode;
    for (int i0 = 0; i0 < 2 * nci; i0 += L) {
        const int i = min(i0 + (int)lane, 2 * nci - 1);
        const bool pu = i < nci;
        const int q = pu ? i : i - nci;
        const int r = (int)udiv_small((uint32_t)q, (uint32_t)nqc), c0 = (q - r * nqc) * 4;
        const uint32_t sh = pu ? 0u : 16u;
        const uint32_t a0 = lane_read(4u * (gb + (uint32_t)(c0 >> 1)), hi);
        const uint32_t a1 = lane_read(4u * (gb + (uint32_t)min((c0 >> 1) + 1, L - 1)), hi);
        const uint32_t lw = lane_read(4u * (gb + (uint32_t)(ncx + (r >> 1))), hi);
        const int l = (int)((lw >> (sh + 8 * (r & 1))) & 0xff);
        const int tl = pu ? (int)((thi >> 8) & 0xff) : (int)(thi >> 24);
        const int dc = pu ? (int)((dcw >> 8) & 0xff) : (int)((dcw >> 16) & 0xff);
        const uint32_t av = ((a0 >> sh) & 0xffff) | ((a1 >> sh) & 0xffff) << 16;
        uint32_t word = 0;
#pragma unroll
        for (int e = 0; e < 4; e++)
            word |= (uint32_t)intra_pred(uvmode, (int)((av >> (8 * e)) & 0xff), l, tl, dc) << (8 * e);
        if (i0 + (int)lane < 2 * nci) {
            uint8_t *pl = pu ? out.u : out.v;
            store_quad(pl + (size_t)(cby + r) * out.suv + cbx + c0, word, cbw - c0);
        }
    }
}

/* Wave w of the batch's sample pass: a job's first ceil(n_px_small / 4)
 * waves take its blocks of at most 8x8 four at a time (16 lanes each, the
 * entries J.px lists first), its other waves one entry each. */
__global__ __launch_bounds__(256) void k_intra_px(DBatch B)
{
    const uint32_t w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (w >= B.pre_px[B.njobs]) return; /* uniform per wave */
    const uint32_t j = find_job_wave(B.pre_px, B.njobs, w);
    const DJob &J = B.jobs[j];
    const uint32_t lw = w - B.pre_px[j], ns = J.n_px_small, G = (ns + 3) / 4;
    if (lw < G) {
        const uint32_t sub = __lane_id() >> 4, e = 4 * lw + sub;
        if (e < ns) intra_px_entry<16>(J, gm(J.px)[e], 16 * sub);
    } else {
        intra_px_entry<64>(J, __builtin_amdgcn_readfirstlane(gm(J.px)[ns + lw - G]));
    }
}

/* The same for the chained entries of a one-job batch (chain workers leave
 * their blocks' samples to it): wave w takes chained entry w. */
__global__ __launch_bounds__(256) void k_intra_px_chain(DBatch B, uint32_t n)
{
    const uint32_t w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (w >= n) return; /* uniform per wave */
    const DJob &J = B.jobs[0];
    const uint32_t tk = __builtin_amdgcn_readfirstlane(gm(J.chain)[J.n_chains + 1 + w]);
    intra_px_entry<64>(J, tk);
}

/* ------------------------------------------------------------------------ */
/* Deblocking (vp9_reconstruct.c:134-191), luma only, in place.             */
/*                                                                          */
/* The horizontal pass filters rows y-2..y+1 around each edge y = 8, 16, ..  */
/* (< H), column by column; the vertical pass then filters columns x-2..x+1 */
/* around each edge x = 8, 16, .. (< W), row by row, on the horizontal      */
/* pass's output. Split the plane into units of 8 rows x 8 columns at       */
/* (8k - 2, 8m - 4): the rows the horizontal pass touches are the first     */
/* four of a unit (8k-2 .. 8k+1), the columns the vertical pass touches its */
/* middle four (8m-2 .. 8m+1). So every sample a unit's filters read or     */
/* write is the unit's own: one lane filters its unit in registers, both    */
/* passes in the reference's order, and no two lanes share a byte. Nothing  */
/* goes through LDS and nothing waits; each byte of the plane is read and   */
/* written once.                                                            */

__device__ __forceinline__ void lf_edge(int &p1, int &p0, int &q0, int &q1, int thr)
{
    const int a = p1, b = p0, c = q0, d = q1;
    if (abs(b - c) * 2 + abs(a - d) / 2 <= thr) {
        const int delta = clampi((3 * (c - b) + 4) >> 3, -thr, thr);
        p0 = clampi(b + delta, 0, 255);
        q0 = clampi(c - delta, 0, 255);
        p1 = clampi(a + delta / 2, 0, 255);
        q1 = clampi(d - delta / 2, 0, 255);
    }
}

__device__ __forceinline__ int byte_of(uint32_t w, int e)
{
    return (int)((w >> (8 * e)) & 0xff);
}
__device__ __forceinline__ uint32_t with_byte(uint32_t w, int e, int v)
{
    return (w & ~(0xffu << (8 * e))) | (uint32_t)v << (8 * e);
}

/* Units per job: lf_wy bands of 8 rows, each lf_wx waves of 64 units (see
 * the job descriptor); wave w of the batch takes 64 units side by side in
 * one band, lane = unit, a unit row being two dwords (columns 8m-4 .. 8m-1
 * and 8m .. 8m+3). */
__device__ __forceinline__ void lf_units(const DBatch &B, uint32_t w)
{
    const uint32_t j = find_job_wave(B.pre_lf, B.njobs, w);
    const DJob &J = B.jobs[j];
    const uint32_t l = w - B.pre_lf[j];
    const int k = (int)(l / J.lf_wx);
    const int m = (int)(l - (uint32_t)k * J.lf_wx) * 64 + (int)__lane_id();
    const int W = (int)J.out.w, H = (int)J.out.h, thr = J.lf_thr;
    const int nu = (W + 3) / 8 + 1, dmax = (W - 1) >> 2;
    GMEM uint32_t *const Y = (GMEM uint32_t *)gm(J.out.y);
    const size_t sw = J.out.sy >> 2;
    /* dwords 2m-1 and 2m of each row; the first holds a sample of the plane
     * for every unit m >= 1 up to nu - 1 (8m - 4 <= W - 1), the second when
     * 8m < W; loads at clamped places, stores only where they belong */
    const int d0 = max(2 * m - 1, 0), d1 = min(2 * m, dmax);
    const bool ok0 = m >= 1 && m < nu, ok1 = m < nu && 2 * m <= dmax;
    const int y0 = 8 * k - 2;
    uint32_t lo[8], hi[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const size_t row = (size_t)clampi(y0 + r, 0, H - 1) * sw;
        lo[r] = Y[row + d0];
        hi[r] = Y[row + d1];
    }
    /* horizontal edge y = 8k (rows 8k-2 .. 8k+1 lie inside the plane for an
     * accepted frame: H is not 1 mod 8), all 8 columns */
    if (k >= 1 && 8 * k < H) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            int a = byte_of(lo[0], e), b = byte_of(lo[1], e), c = byte_of(lo[2], e), d = byte_of(lo[3], e);
            lf_edge(a, b, c, d, thr);
            lo[0] = with_byte(lo[0], e, a);
            lo[1] = with_byte(lo[1], e, b);
            lo[2] = with_byte(lo[2], e, c);
            lo[3] = with_byte(lo[3], e, d);
            a = byte_of(hi[0], e), b = byte_of(hi[1], e), c = byte_of(hi[2], e), d = byte_of(hi[3], e);
            lf_edge(a, b, c, d, thr);
            hi[0] = with_byte(hi[0], e, a);
            hi[1] = with_byte(hi[1], e, b);
            hi[2] = with_byte(hi[2], e, c);
            hi[3] = with_byte(hi[3], e, d);
        }
    }
    /* vertical edge x = 8m, all 8 rows (rows outside the plane are never
     * stored) */
    if (m >= 1 && 8 * m < W) {
#pragma unroll
        for (int r = 0; r < 8; r++) {
            int a = byte_of(lo[r], 2), b = byte_of(lo[r], 3), c = byte_of(hi[r], 0), d = byte_of(hi[r], 1);
            lf_edge(a, b, c, d, thr);
            lo[r] = (lo[r] & 0xffffu) | (uint32_t)a << 16 | (uint32_t)b << 24;
            hi[r] = (hi[r] & 0xffff0000u) | (uint32_t)c | (uint32_t)d << 8;
        }
    }
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int y = y0 + r;
        if (y >= 0 && y < H) {
            const size_t row = (size_t)y * sw;
            if (ok0) Y[row + d0] = lo[r];
            if (ok1) Y[row + d1] = hi[r];
        }
    }
}

/* Four waves per workgroup, each its own item; workgroups mapped so that
 * each XCD takes one contiguous run of the bands. */
__global__ __launch_bounds__(256) void k_lf(DBatch B)
{
    const uint32_t w = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * 4 +
                                                      (threadIdx.x >> 6));
    if (w >= B.pre_lf[B.njobs]) return; /* uniform per wave */
    lf_units(B, w);
}

/* ------------------------------------------------------------------------ */
/* I420 -> NV12 into the job's render target, pitch W (vp9_nv12_pack and    */
/* vaapi_entrypoint.c:131-153 layout). One workgroup per 64x64 luma tile.    */

__device__ __forceinline__ void nv12_item(const DBatch &B, uint32_t w)
{
    const uint32_t j = find_job_wave(B.pre_nv12, B.njobs, w);
    const DJob &J = B.jobs[j];
    const uint32_t l = w - B.pre_nv12[j];
    const uint32_t W = J.out.w, H = J.out.h, cw = (W + 1) / 2, ch = (H + 1) / 2;
    const uint32_t tx = (W + 63) / 64;
    const uint32_t ty0 = (l / tx) * 64, tx0 = (l % tx) * 64;
    GMEM uint8_t *const dst = gm(J.nv12);
    const DFrame F = J.out; /* held in registers across the stores */
    /* luma: 4 samples per thread and pass, one aligned load (rows are
     * 64-byte aligned); one 4-byte store where the target row allows it */
    const bool y4 = (W & 3) == 0;
    for (uint32_t i = threadIdx.x; i < 64 * 16; i += 256) {
        const uint32_t r = ty0 + (i >> 4), c = tx0 + 4 * (i & 15);
        if (r >= H || c >= W) continue;
        const uint32_t q = *(const GMEM uint32_t *)(gm(F.y) + (size_t)r * F.sy + c);
        GMEM uint8_t *o = dst + (size_t)r * W + c;
        if (y4 && c + 4 <= W) {
            *(GMEM uint32_t *)o = q;
        } else {
            for (uint32_t e = 0; e < 4 && c + e < W; e++) o[e] = (uint8_t)(q >> (8 * e));
        }
    }
    /* chroma: 2 samples of each plane per thread and pass, interleaved
     * U V U V into one 4-byte store where the target row allows it */
    GMEM uint8_t *const uv = dst + (size_t)W * H;
    const bool c4 = (cw & 1) == 0;
    for (uint32_t i = threadIdx.x; i < 32 * 16; i += 256) {
        const uint32_t r = ty0 / 2 + (i >> 4), c = tx0 / 2 + 2 * (i & 15);
        if (r >= ch || c >= cw) continue;
        const uint32_t u = *(const GMEM uint16_t *)(gm(F.u) + (size_t)r * F.suv + c);
        const uint32_t v = *(const GMEM uint16_t *)(gm(F.v) + (size_t)r * F.suv + c);
        const uint32_t pair0 = (u & 0xff) | (v & 0xff) << 8, pair1 = (u >> 8) | (v & 0xff00);
        GMEM uint8_t *o = uv + (size_t)r * 2 * cw + 2 * c;
        if (c + 2 <= cw) {
            if (c4) {
                *(GMEM uint32_t *)o = pair0 | pair1 << 16;
            } else {
                *(GMEM uint16_t *)o = (uint16_t)pair0;
                *(GMEM uint16_t *)(o + 2) = (uint16_t)pair1;
            }
        } else {
            *(GMEM uint16_t *)o = (uint16_t)pair0;
        }
    }
}

__global__ __launch_bounds__(256) void k_nv12(DBatch B)
{
    nv12_item(B, xcd_block(blockIdx.x, gridDim.x)); /* neighbouring tiles share row lines */
}

/* ------------------------------------------------------------------------ */
/* Fused persistent reconstruction: every stage of a batch in one launch.   */
/* All workgroups are co-resident (the grid is sized from the occupancy     */
/* query); each phase spreads its work items over the grid and ends with a  */
/* grid barrier. Waits are bounded: a timeout sets error byte 2 of the      */
/* first job's stream and falls through.                                    */

/* The grid barrier's per-workgroup state, held by thread 0. */
struct BarState {
    uint32_t grp;    /* this workgroup's group: the id of the XCD it runs on */
    uint32_t n;      /* workgroups in the group (0: not read yet) */
    uint32_t ngrp;   /* groups with any workgroup */
};

__device__ __forceinline__ uint32_t xcd_id()
{
    /* HW_REG_XCC_ID (hardware register 20), bits 3:0 */
    return __builtin_amdgcn_s_getreg(20 | (0 << 6) | (31 << 11)) & (CVP9_BAR_GROUPS - 1);
}

__device__ __forceinline__ bool bar_wait(const DBatch &B, const uint32_t *w, uint32_t target)
{
    uint32_t spins = 0;
    while (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > B.spin_limit) {
            flag_error(B.jobs[0].err, B.jobs[0].err_host, 2);
            return false;
        }
    }
    return true;
}

/* At kernel start: thread 0 registers its workgroup in its XCD's group. */
__device__ __forceinline__ void bar_register(const DBatch &B, BarState &st)
{
    st.grp = xcd_id();
    st.n = 0;
    __hip_atomic_fetch_add(B.bar + CVP9_BAR_COUNT + st.grp, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    /* release: the count is in before the registration is */
    __hip_atomic_fetch_add(B.bar + CVP9_BAR_STARTED, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

/* Barrier `ph` (1, 2, ...). Every wave first waits until its own stores are
 * acknowledged by the L2 of its XCD, which all workgroups of the group share.
 * The last workgroup of a group to arrive makes the whole L2 visible device
 * wide with one agent-scope release (an L2 write-back: it covers the stores
 * of every workgroup of the group, as they are all in that L2) before it
 * arrives on the global counter; so only it waits there, and the others
 * wait on their group's release word: 1/ngroups of the pollers per word, and
 * one L2 write-back per XCD instead of one per workgroup. Each workgroup
 * then acquires at agent scope itself. */
__device__ __forceinline__ void grid_barrier(const DBatch &B, BarState &st, uint32_t ph)
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        if (st.n == 0) {
            /* the first barrier: every workgroup has registered once the
             * grid has, and the group sizes are final */
            if (bar_wait(B, B.bar + CVP9_BAR_STARTED, gridDim.x)) {
                st.ngrp = 0;
                for (uint32_t g = 0; g < CVP9_BAR_GROUPS; g++) {
                    const uint32_t c = __hip_atomic_load(B.bar + CVP9_BAR_COUNT + g, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
                    st.ngrp += c ? 1u : 0u;
                    if (g == st.grp) st.n = c;
                }
            }
            if (st.n == 0) st.n = st.ngrp = 1; /* timed out: the error is set; do not divide the wait */
        }
        uint32_t *arrive = B.bar + CVP9_BAR_ARRIVE + 32 * st.grp;
        uint32_t *release = B.bar + CVP9_BAR_RELEASE + 32 * st.grp;
        const uint32_t old = __hip_atomic_fetch_add(arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1 == st.n * ph) {
            /* the group's last: its members' arrivals precede this point */
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_fetch_add(B.bar + CVP9_BAR_GLOBAL, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            bar_wait(B, B.bar + CVP9_BAR_GLOBAL, st.ngrp * ph);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(release, ph, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            bar_wait(B, release, ph);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

/* transform split used inside the fused kernel, per size 4, 8, 16, 32 */
constexpr int FS4 = 1, FS8 = 1, FS16 = 2, FS32 = 4;
constexpr size_t FUSED_LDS =
    cmax(8 * cmax(cmax(ResShape<4, FS4>::LDS_DOUBLES, ResShape<8, FS8>::LDS_DOUBLES),
                  cmax(ResShape<16, FS16>::LDS_DOUBLES, ResShape<32, FS32>::LDS_DOUBLES)),
         cmax(4 * sizeof(IntraShared), 64 * 65));

template <int N, int S>
__device__ __forceinline__ void residual_phase(const DBatch &B, double *lds)
{
    using RS = ResShape<N, S>;
    const uint32_t units = (B.pre_grp[RS::C][B.njobs] + RS::GPW - 1) / RS::GPW;
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        residual_unit<N, S>(B, u, lds);
        __syncthreads(); /* the next unit reuses the LDS */
    }
}

__global__ __launch_bounds__(256) void k_fused(DBatch B)
{
    __shared__ __align__(16) uint8_t smem[FUSED_LDS];
    const uint32_t nwg = gridDim.x;
    const int wave = threadIdx.x >> 6;
    uint32_t phase = 0;
    BarState bar{0u, 0u, 0u};
    if (threadIdx.x == 0) bar_register(B, bar);

    if (blockIdx.x == 0) assign_epochs(B); /* read by the intra phase, after a grid barrier */
    const uint32_t nfill = B.pre_fill[B.njobs];
    if (nfill) {
        for (uint32_t w = blockIdx.x; w < nfill; w += nwg) fill_item(B, w);
        grid_barrier(B, bar, ++phase);
    }
    /* owner map and motion compensation (independent) */
    const uint32_t nblk = B.pre_blocks[B.njobs], nmc = B.pre_inter[B.njobs];
    for (uint32_t wv = blockIdx.x * 16 + (threadIdx.x >> 4); wv < nblk; wv += nwg * 16)
        owner_item(B, wv, threadIdx.x & 15, find_job(B.pre_blocks, B.njobs, wv));
    for (uint32_t u = blockIdx.x * 16 + threadIdx.x / CVP9_MC_UNIT; u < nmc; u += nwg * 16)
        mc_unit_at(B, u, threadIdx.x % CVP9_MC_UNIT);
    grid_barrier(B, bar, ++phase);
    /* transforms: inter ones add to the prediction, intra ones go aside */
    double *lds = (double *)smem;
    residual_phase<4, FS4>(B, lds);
    residual_phase<8, FS8>(B, lds);
    residual_phase<16, FS16>(B, lds);
    residual_phase<32, FS32>(B, lds);
    grid_barrier(B, bar, ++phase);
    /* neighbour records, then intra blocks in ticket order on the tuned
     * number and width of persistent workers (the other workgroups go on
     * to the barrier: more workers than the chains can feed only poll) */
    const uint32_t nint = B.pre_intra[B.njobs];
    if (nint) {
        for (uint32_t wv = blockIdx.x * 8 + (threadIdx.x >> 5); wv < nint; wv += nwg * 8)
            nbr_item(B, wv, threadIdx.x & 31, find_job(B.pre_intra, B.njobs, wv));
        grid_barrier(B, bar, ++phase);
        const uint32_t nwk = B.fused_workers < 1 ? 1 : B.fused_workers;
        if (B.fused_lanes == 64) {
            const uint32_t wgs = (nwk + 3) / 4 < nwg ? (nwk + 3) / 4 : nwg;
            if (blockIdx.x < wgs)
                intra_worker<64, true>(B, (IntraShared *)smem, blockIdx.x * 4 + wave, wgs * 4);
        } else {
            const uint32_t wgs = nwk < nwg ? nwk : nwg;
            if (blockIdx.x < wgs) intra_worker<256, true>(B, (IntraShared *)smem, blockIdx.x, wgs);
        }
        grid_barrier(B, bar, ++phase);
    }
    /* deblocking: one item of 64 units per wave */
    const uint32_t nlf = B.pre_lf[B.njobs];
    for (uint32_t w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wave); w < nlf; w += nwg * 4)
        lf_units(B, w);
    /* NV12 */
    const uint32_t nnv = B.pre_nv12[B.njobs];
    if (nnv) {
        grid_barrier(B, bar, ++phase);
        for (uint32_t w = blockIdx.x; w < nnv; w += nwg) nv12_item(B, w);
    }
}

hipError_t cvp9_launch_fused(const DBatch &b, uint32_t wgs, hipStream_t s)
{
    hipLaunchKernelGGL(k_fused, dim3(wgs), dim3(256), 0, s, b);
    return hipGetLastError();
}

template <int W, bool DEFER>
static uint32_t intra_capacity()
{
    static uint32_t cache[16] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) dev = 0;
    if (!cache[dev]) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_intra<W, DEFER>, 256, 0) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            per_cu < 1 || cus < 1)
            return 256;
        cache[dev] = (uint32_t)(per_cu * cus);
    }
    return cache[dev];
}

/* The workgroups of k_intra<W> a launch may have. Persistent workers own
 * their intra entries by stride, and a worker may wait on an entry of any
 * other worker of its job, so every workgroup of a launch must become
 * resident while the others wait. The kernels that run beside it never
 * wait, so they end and give their slots up; the other launches that may
 * wait are the intra stages on the other queue of the same engine and on
 * the queues of the other engines of the process. So a launch is held to
 * the device's resident capacity for the kernel (when nothing else runs;
 * cached per device) divided over two queues per live engine. Engines of
 * other processes on the same device are not counted: a wait that then
 * cannot end gives up at its bound and is reported as a device error. */
template <int W, bool DEFER>
static uint32_t intra_resident_wgs()
{
    const int live = cvp9_live_engines();
    const uint32_t per = intra_capacity<W, DEFER>() / (2u * (uint32_t)(live > 1 ? live : 1));
    return per ? per : 1u;
}

static uint32_t chain_resident_wgs()
{
    static uint32_t cache[16] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) dev = 0;
    if (!cache[dev]) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_intra_chain, 256, 0) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            per_cu < 1 || cus < 1)
            return 128;
        cache[dev] = (uint32_t)(per_cu * cus);
    }
    const int live = cvp9_live_engines();
    const uint32_t per = cache[dev] / (2u * (uint32_t)(live > 1 ? live : 1));
    return per ? per : 1u;
}

uint32_t cvp9_fused_max_wgs(int device)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fused, 256, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    return (uint32_t)(per_cu * cus);
}

/* ------------------------------------------------------------------------ */
/* Plain streaming copy: the reference kernel for bandwidth and counter      */
/* calibration (bytes moved are known exactly).                              */

/* One pass over the buffer (no grid stride), each thread moving four
 * 16-byte words with nontemporal loads and stores, all four loads issued
 * before the first store: on MI355X this shape streams about 6.3 TB/s,
 * against 4.6 TB/s for a grid-strided copy of one word per thread. */
typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));
constexpr int COPY_UNROLL = 4;

__global__ __launch_bounds__(256) void k_copy_probe(const v4u32 *__restrict__ src,
                                                    v4u32 *__restrict__ dst, size_t n)
{
    const size_t base = (size_t)blockIdx.x * 256 * COPY_UNROLL + threadIdx.x;
    v4u32 v[COPY_UNROLL];
#pragma unroll
    for (int u = 0; u < COPY_UNROLL; u++) {
        const size_t i = base + (size_t)u * 256;
        v[u] = i < n ? __builtin_nontemporal_load(src + i) : (v4u32){0, 0, 0, 0};
    }
#pragma unroll
    for (int u = 0; u < COPY_UNROLL; u++) {
        const size_t i = base + (size_t)u * 256;
        if (i < n) __builtin_nontemporal_store(v[u], dst + i);
    }
}

hipError_t cvp9_launch_copy_probe(const void *src, void *dst, size_t bytes, hipStream_t s)
{
    const size_t n = bytes / 16;
    const size_t per = 256 * COPY_UNROLL;
    hipLaunchKernelGGL(k_copy_probe, dim3((uint32_t)((n + per - 1) / per)), dim3(256), 0, s,
                       (const v4u32 *)src, (v4u32 *)dst, n);
    return hipGetLastError();
}

/* Snapshot of a stream's (device) error word into a per-frame host word,
 * in queue order (a copy into host memory through a kernel: the host's
 * word is read without any further synchronisation). */
__global__ void k_copy_word(const uint32_t *src, uint32_t *dst)
{
    __hip_atomic_store(dst, read_error(src), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

hipError_t cvp9_launch_copy_word(const uint32_t *src, uint32_t *dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_word, dim3(1), dim3(1), 0, s, src, dst);
    return hipGetLastError();
}

/* ------------------------------------------------------------------------ */
/* host-side launchers                                                       */

hipError_t cvp9_set_idct_tables(const double *m4, const double *m8, const double *m16,
                                const double *m32)
{
    hipError_t e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_M4), m4, 16 * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_M8), m8, 64 * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_M16), m16, 256 * sizeof(double))) != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_M32), m32, 1024 * sizeof(double));
}

static inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

Tune cvp9_default_tune(int cus)
{
    Tune t;
    t.res_split[0] = 1;
    t.res_split[1] = 1;
    t.res_split[2] = 2;
    t.res_split[3] = 4;
    t.intra_lanes = 256;
    t.intra_workers = 4u * (uint32_t)(cus > 0 ? cus : 256);
    t.intra_level_min = 0;
    t.intra_defer = 0;
    return t;
}

/* The intra stage: leading levels one launch each, the persistent workers
 * for the rest, and with DEFER the pass that writes the blocks' samples. */
template <bool DEFER>
static hipError_t launch_intra(const DBatch &b, const StageCounts &n, hipStream_t s)
{
    /* chain workers: a one-job batch that lists chains, no level launches,
     * more workers than chains */
    if (!DEFER && n.tune.intra_chains && n.chains && n.levels == 0 && n.intra_tail) {
        const uint32_t cap = chain_resident_wgs();
        uint32_t g = div_up(n.tune.intra_workers, 4);
        if (g > cap) g = cap;
        if (g * 4 > n.chains) {
            hipLaunchKernelGGL(k_intra_chain, dim3(g), dim3(256), 0, s, b);
            if (n.chained)
                hipLaunchKernelGGL(k_intra_px_chain, dim3(div_up(n.chained, 4)), dim3(256), 0, s, b,
                                   n.chained);
            return hipGetLastError();
        }
    }
    for (uint32_t l = 0; l < n.levels; l++) {
        const uint32_t sw = div_up(n.level_ns[l], 4);
        hipLaunchKernelGGL(k_intra_level<DEFER>, dim3(sw + n.level_n[l]), dim3(256), 0, s, b, l, sw);
    }
    if (n.intra_tail) {
        const uint32_t w = n.intra_tail < n.tune.intra_workers ? n.intra_tail : n.tune.intra_workers;
        if (n.tune.intra_lanes == 64) {
            const uint32_t g = div_up(w, 4), cap = intra_resident_wgs<64, DEFER>();
            hipLaunchKernelGGL((k_intra<64, DEFER>), dim3(g < cap ? g : cap), dim3(256), 0, s, b);
        } else {
            const uint32_t cap = intra_resident_wgs<256, DEFER>();
            hipLaunchKernelGGL((k_intra<256, DEFER>), dim3(w < cap ? w : cap), dim3(256), 0, s, b);
        }
    }
    if (DEFER && n.px) hipLaunchKernelGGL(k_intra_px, dim3(div_up(n.px, 4)), dim3(256), 0, s, b);
    return hipGetLastError();
}

template <int N>
static hipError_t launch_residual(const DBatch &b, const StageCounts &n, int c, hipStream_t s)
{
    const uint32_t g = n.grp[c];
    if (!g) return hipSuccess;
    if constexpr (N == 4) { /* one transform per lane, 16 groups per workgroup */
        hipLaunchKernelGGL(k_res4_lanes, dim3(div_up(g, 16)), dim3(256), 0, s, b);
        return hipGetLastError();
    }
    switch (n.tune.res_split[c]) {
    case 1: hipLaunchKernelGGL((k_residual<N, 1>), dim3(div_up(g, 4)), dim3(256), 0, s, b); break;
    case 2: hipLaunchKernelGGL((k_residual<N, 2>), dim3(div_up(g, 2)), dim3(256), 0, s, b); break;
    default: hipLaunchKernelGGL((k_residual<N, 4>), dim3(g), dim3(256), 0, s, b); break;
    }
    return hipGetLastError();
}

/* workgroups of the one-launch transform stage: per size, then the total */
static inline uint3 res_small_starts(const StageCounts &n, uint32_t *total)
{
    const uint32_t a = div_up(n.grp[0], 4), b = a + div_up(n.grp[1], 4),
                   c = b + div_up(n.grp[2], 2);
    *total = c + n.grp[3];
    return make_uint3(a, b, c);
}

/* Launches kernel K; with *done set, the event is recorded at the kernel's
 * completion by its own dispatch (hipExtLaunchKernel's stop event) and
 * *done is cleared. */
#define CVP9_LAUNCH(done, K, grid, block, s, ...)                                              \
    do {                                                                                       \
        if ((done) && *(done)) {                                                               \
            hipExtLaunchKernelGGL(K, grid, block, 0, s, nullptr, *(done), 0, __VA_ARGS__);     \
            *(done) = nullptr;                                                                 \
        } else {                                                                               \
            hipLaunchKernelGGL(K, grid, block, 0, s, __VA_ARGS__);                             \
        }                                                                                      \
    } while (0)

static inline uint32_t pad8(uint32_t x) { return (x + 7) & ~7u; }

int cvp9_stage_span(int stage, const DBatch &b, const StageCounts &n)
{
    /* owner map + motion compensation (with resid_all, motion compensation
     * follows the transforms instead) */
    if (stage == STAGE_OWNER && n.inter && !b.resid_all) return 2;
    /* with resid_all (the pipelined first phase): owner map + the four
     * transform sizes, then neighbour records + early motion compensation */
    if (stage == STAGE_OWNER && b.resid_all) {
        uint32_t total;
        res_small_starts(n, &total);
        return total && total <= CVP9_RES_SMALL_MAX_WGS ? 5 : 1;
    }
    if (stage == STAGE_NBR && b.resid_all && n.intra && n.inter) return 2;
    if (stage == STAGE_RES4) {
        uint32_t total;
        res_small_starts(n, &total);
        if (total <= CVP9_RES_SMALL_MAX_WGS) return 4;
    }
    return 1;
}

hipError_t cvp9_launch_stages(int stage, int span, const DBatch &b, const StageCounts &n,
                              hipStream_t s, hipEvent_t *done)
{
    if (span == 2 && stage == STAGE_OWNER) {
        const uint32_t ow = n.blocks ? div_up(n.blocks, 16) : 1;
        hipLaunchKernelGGL(k_front, dim3(ow + n.inter / CVP9_MC_WG_UNITS), dim3(256), 0, s, b, ow);
        return hipGetLastError();
    }
    if (span == 5 && stage == STAGE_OWNER) {
        uint32_t total;
        const uint3 wg = res_small_starts(n, &total);
        const uint32_t ow = pad8(n.blocks ? div_up(n.blocks, 16) : 1);
        hipLaunchKernelGGL(k_owner_res, dim3(ow + total), dim3(256), 0, s, b, ow, wg);
        return hipGetLastError();
    }
    if (span == 2 && stage == STAGE_NBR) {
        const uint32_t used = div_up(n.intra, 8), nw = pad8(used);
        hipLaunchKernelGGL(k_nbr_mc, dim3(nw + n.inter / CVP9_MC_WG_UNITS), dim3(256), 0, s, b, nw, used);
        return hipGetLastError();
    }
    if (span == 4 && stage == STAGE_RES4) {
        uint32_t total;
        const uint3 wg = res_small_starts(n, &total);
        if (total) hipLaunchKernelGGL(k_res_small, dim3(total), dim3(256), 0, s, b, wg);
        return hipGetLastError();
    }
    return span == 1 ? cvp9_launch_stage(stage, b, n, s, done) : hipErrorInvalidValue;
}

hipError_t cvp9_launch_mc_late(const DBatch &b, const StageCounts &n, hipStream_t s, hipEvent_t *done)
{
    if (n.inter2) CVP9_LAUNCH(done, k_mc, dim3(n.inter2 / CVP9_MC_WG_UNITS), dim3(256), s, b, 1);
    return hipGetLastError();
}


hipError_t cvp9_launch_stage(int stage, const DBatch &b, const StageCounts &n, hipStream_t s,
                             hipEvent_t *done)
{
    switch (stage) {
    case STAGE_FILL:
        if (n.fill) hipLaunchKernelGGL(k_fill, dim3(n.fill), dim3(256), 0, s, b);
        break;
    case STAGE_OWNER:
        hipLaunchKernelGGL(k_owner, dim3(n.blocks ? div_up(n.blocks, 16) : 1), dim3(256), 0, s, b);
        break;
    case STAGE_MC:
        if (n.inter) CVP9_LAUNCH(done, k_mc, dim3(n.inter / CVP9_MC_WG_UNITS), dim3(256), s, b, 0);
        break;
    case STAGE_RES4:
        return launch_residual<4>(b, n, 0, s);
    case STAGE_RES8:
        return launch_residual<8>(b, n, 1, s);
    case STAGE_RES16:
        return launch_residual<16>(b, n, 2, s);
    case STAGE_RES32:
        return launch_residual<32>(b, n, 3, s);
    case STAGE_NBR:
        if (n.intra) hipLaunchKernelGGL(k_nbr, dim3(div_up(n.intra, 8)), dim3(256), 0, s, b);
        break;
    case STAGE_INTRA:
        return n.tune.intra_defer ? launch_intra<true>(b, n, s) : launch_intra<false>(b, n, s);
    case STAGE_LF:
        if (n.lf) CVP9_LAUNCH(done, k_lf, dim3(div_up(n.lf, 4)), dim3(256), s, b);
        break;
    case STAGE_NV12:
        if (n.nv12) CVP9_LAUNCH(done, k_nv12, dim3(n.nv12), dim3(256), s, b);
        break;
    case STAGE_EXPAND:
        if (n.exp) hipLaunchKernelGGL(k_expand, dim3(div_up(n.exp, 4)), dim3(256), 0, s, b);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
