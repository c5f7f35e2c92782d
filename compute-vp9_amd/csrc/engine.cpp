/*
 * engine.cpp — frame planning, device stream state and batch launch.
 */
#include "engine.hpp"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

namespace cvp9 {

static inline uint64_t align256(uint64_t v) { return (v + 255) & ~255ull; }
static inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

/* ---- IDCT tables (vp9_reconstruct.c:19-28) ------------------------------ */

void idct_matrix(int n, double *m)
{
    /* Same expression and evaluation order as the reference; built with
     * -ffp-contract=off and the C library's cos, so every entry is
     * bit-identical to the table the CPU path builds. */
    const double s0 = sqrt(1.0 / n);
    const double s1 = sqrt(2.0 / n);
    for (int i = 0; i < n; i++) {
        const double s = i == 0 ? s0 : s1;
        for (int j = 0; j < n; j++)
            m[i * n + j] = s * cos(((2.0 * j + 1.0) * i * 3.141592653589793) / (2.0 * n));
    }
}

hipError_t init_device(int device)
{
    static std::mutex mu;
    static bool done[64];
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(mu);
    if (done[device]) return hipSuccess;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return e;
    static double m4[16], m8[64], m16[256], m32[1024];
    idct_matrix(4, m4);
    idct_matrix(8, m8);
    idct_matrix(16, m16);
    idct_matrix(32, m32);
    e = cvp9_set_idct_tables(m4, m8, m16, m32);
    if (e == hipSuccess) done[device] = true;
    return e;
}

/* ---- planning ---------------------------------------------------------- */

namespace {

/* sets bits [start, start+len) of a bitmap; false if any was already set */
bool claim_bits(std::vector<uint64_t> &bits, uint64_t start, uint64_t len)
{
    while (len) {
        const uint64_t w = start >> 6, o = start & 63;
        const uint64_t n = len < 64 - o ? len : 64 - o;
        const uint64_t mask = (n == 64 ? ~0ull : ((1ull << n) - 1)) << o;
        if (bits[w] & mask) return false;
        bits[w] |= mask;
        start += n;
        len -= n;
    }
    return true;
}

struct Clamped {
    int bw, bh;
};

inline Clamped clamp_block(const vp9_macroblock_info_t &b, int W, int H)
{
    /* vp9_reconstruct.c:218-221 */
    Clamped c;
    c.bw = b.x + b.width > W ? W - b.x : b.width;
    c.bh = b.y + b.height > H ? H - b.y : b.height;
    return c;
}

inline uint8_t mode_class(uint8_t m)
{
    /* vp9_reconstruct.c:67-90: 0 DC, 1 V, 2 H, anything else TM */
    return m <= 2 ? m : 3;
}

} // namespace

cvp9_err_t plan_measure(const vp9_parsed_frame_t *pf, PrepLayout &L)
{
    if (!pf) return CVP9_ERR_INVALID_DATA;
    L = PrepLayout();
    const uint32_t W = pf->hdr.width, H = pf->hdr.height;
    /* Sizes the kernels take. The CPU deblock reads one row (column) past
     * the plane when H (W) is 1 mod 8, and the chroma bilinear clamp
     * breaks on a 1-pixel reference; those go to the CPU path. */
    if (W < 2 || H < 2 || W > 16384 || H > 16384) return CVP9_ERR_UNSUPPORTED;
    if ((W & 7) == 1 || (H & 7) == 1) return CVP9_ERR_UNSUPPORTED;
    if (pf->num_blocks && !pf->blocks) return CVP9_ERR_INVALID_DATA;
    L.W = W;
    L.H = H;
    L.key = pf->hdr.frame_type == VP9_FRAME_KEY;
    if (pf->hdr.filter_level) {
        uint32_t thr = pf->hdr.filter_level; /* vp9_reconstruct.c:139-144 */
        const uint32_t sh = pf->hdr.sharpness_level;
        if (sh > 0) {
            thr >>= (sh > 4) ? 2 : 1;
            if (thr > 9 - sh) thr = 9 - sh;
        }
        if (thr == 0) thr = 1;
        L.lf_thr = (int)thr;
    }

    const uint32_t ow = (W + 3) / 4, oh = (H + 3) / 4;
    thread_local std::vector<uint64_t> bits;
    bits.assign(((uint64_t)ow * oh + 63) / 64 + 1, 0);
    uint64_t covered = 0;

    for (uint32_t i = 0; i < pf->num_blocks; i++) {
        const vp9_macroblock_info_t &b = pf->blocks[i];
        const Clamped c = clamp_block(b, (int)W, (int)H);
        if (c.bw <= 0 || c.bh <= 0) continue;
        /* blocks on the 4x4 grid: neighbour ownership is exact per cell */
        if ((b.x & 3) || (b.y & 3) || (b.width & 3) || (b.height & 3)) return CVP9_ERR_UNSUPPORTED;
        const bool intra = b.is_intra != 0;
        /* VP9 blocks are at most 64x64; intra blocks publish their edges in
         * fixed 64-sample records, motion compensation spreads a block row
         * over at most 16 lanes */
        if (c.bw > CVP9_MAX_INTRA_DIM || c.bh > CVP9_MAX_INTRA_DIM) return CVP9_ERR_UNSUPPORTED;
        bool res = !b.skip;
        if (res) {
            if (b.tx_size > 3) return CVP9_ERR_UNSUPPORTED;
            const uint32_t T = 4u << b.tx_size;
            const uint64_t ntx = (uint64_t)(c.bw / T) * (c.bh / T);
            if (ntx == 0) res = false;
            else {
                if (!pf->coeffs) return CVP9_ERR_INVALID_DATA;
                if ((uint64_t)b.coeff_offset + ntx * T * T > pf->num_coeffs) return CVP9_ERR_INVALID_DATA;
                L.cls_nblk[b.tx_size]++;
                L.cls_ntx[b.tx_size] += (uint32_t)ntx;
            }
        }
        const uint32_t cx0 = b.x >> 2, cx1 = (b.x + c.bw - 1) >> 2;
        const uint32_t cy0 = b.y >> 2, cy1 = (b.y + c.bh - 1) >> 2;
        for (uint32_t cy = cy0; cy <= cy1; cy++)
            if (!claim_bits(bits, (uint64_t)cy * ow + cx0, cx1 - cx0 + 1)) return CVP9_ERR_UNSUPPORTED;
        covered += (uint64_t)(cx1 - cx0 + 1) * (cy1 - cy0 + 1);
        L.n_blocks++;
        if (intra) L.n_intra++;
        else {
            L.n_inter++;
            const uint32_t rows = cvp9_mc_unit_rows((uint32_t)c.bw);
            const uint32_t units = ((uint32_t)c.bh + rows - 1) / rows;
            L.n_mc += units;
            if (b.ref_frame[0] != 1) L.n_mc_not0 += units; /* slot ref_frame - 1 */
        }
    }
    L.fill = covered != (uint64_t)ow * oh;
    L.num_coeffs = pf->coeffs ? pf->num_coeffs : 0;

    uint64_t off = 0;
    L.off_blocks = off; off = align256(off + (uint64_t)L.n_blocks * sizeof(DBlock));
    L.off_mc_tab = off; off = align256(off + (uint64_t)L.n_mc * 16);
    L.off_intra = off;  off = align256(off + (uint64_t)L.n_intra * 4);
    for (int k = 0; k < 4; k++) {
        L.off_tx_tab[k] = off; off = align256(off + (uint64_t)L.cls_ntx[k] * 8);
    }
    L.off_coeffs = off; off = align256(off + (uint64_t)L.num_coeffs * 2);
    L.bytes = off;
    return CVP9_OK;
}

/* Blocks that one 64-thread intra worker handles in a single pass. */
static inline bool intra_small(const DBlock &k)
{
    return k.bw <= 16 && k.bh <= 16;
}

/* Orders the intra list by dependency level, then size class (small
 * first), then list index. An intra block depends on the earlier intra
 * blocks owning its above row, left column and above-left sample (for
 * chroma too: a chroma neighbour lies in the same 4x4 luma cells). Level 0
 * has no such block; otherwise the level is one more than the highest
 * level among them. Intra workers take blocks in this order, so every block
 * a worker waits for was handed out earlier, and the blocks of one level
 * can all run at once. Those neighbour samples
 * lie on the owner's bottom row or right column, so only those cells record
 * the owner's level (plus one) in the scratch map, in list order: a cell
 * still at zero when block b reads it is not owned by an earlier intra
 * block. */
static void order_intra(const DBlock *blocks, uint32_t *intra, uint32_t ni, uint32_t W, uint32_t H,
                        std::vector<uint32_t> &levels, std::vector<uint32_t> &level_small)
{
    levels.clear();
    level_small.clear();
    if (ni == 0) return;
    if (ni == 1) {
        levels = {0, 1};
        level_small = {intra_small(blocks[intra[0]]) ? 1u : 0u};
        return;
    }
    const uint32_t ow = (W + 3) / 4, oh = (H + 3) / 4;
    thread_local std::vector<uint32_t> map; /* level + 1 per 4x4 cell, zero between frames */
    thread_local std::vector<uint32_t> lev, cnt, tmp, hgt, pos;
    if (map.size() < (size_t)ow * oh) map.assign((size_t)ow * oh, 0);
    lev.resize(ni);
    uint32_t depth = 0;
    for (uint32_t i = 0; i < ni; i++) {
        const DBlock &k = blocks[intra[i]];
        const uint32_t cx0 = k.x >> 2, cy0 = k.y >> 2;
        const uint32_t cx1 = (k.x + k.bw - 1) >> 2, cy1 = (k.y + k.bh - 1) >> 2;
        uint32_t m = 0;
        if (cy0 > 0) {
            const uint32_t *row = &map[(size_t)(cy0 - 1) * ow];
            for (uint32_t cx = cx0 ? cx0 - 1 : 0; cx <= cx1; cx++) m = std::max(m, row[cx]);
        }
        if (cx0 > 0)
            for (uint32_t cy = cy0; cy <= cy1; cy++) m = std::max(m, map[(size_t)cy * ow + cx0 - 1]);
        lev[i] = m; /* level of this block */
        depth = std::max(depth, m + 1);
        for (uint32_t cx = cx0; cx <= cx1; cx++) map[(size_t)cy1 * ow + cx] = m + 1;
        for (uint32_t cy = cy0; cy <= cy1; cy++) map[(size_t)cy * ow + cx1] = m + 1;
    }
    /* clear what this frame wrote */
    for (uint32_t i = 0; i < ni; i++) {
        const DBlock &k = blocks[intra[i]];
        const uint32_t cx0 = k.x >> 2, cy0 = k.y >> 2;
        const uint32_t cx1 = (k.x + k.bw - 1) >> 2, cy1 = (k.y + k.bh - 1) >> 2;
        for (uint32_t cx = cx0; cx <= cx1; cx++) map[(size_t)cy1 * ow + cx] = 0;
        for (uint32_t cy = cy0; cy <= cy1; cy++) map[(size_t)cy * ow + cx1] = 0;
    }
    /* height: the longest chain of later intra blocks that depend on a
     * block (1 without any), from a pass in reverse list order. A later
     * block leaves its height on the cells it reads (the row above it from
     * the above-left cell on, and the column left of it); a block's
     * dependents are exactly the later blocks that read its bottom row or
     * right column, and blocks earlier in the list write their marks only
     * after it has been scanned. */
    hgt.resize(ni);
    for (uint32_t i = ni; i-- > 0;) {
        const DBlock &k = blocks[intra[i]];
        const uint32_t cx0 = k.x >> 2, cy0 = k.y >> 2;
        const uint32_t cx1 = (k.x + k.bw - 1) >> 2, cy1 = (k.y + k.bh - 1) >> 2;
        uint32_t h = 0;
        for (uint32_t cx = cx0; cx <= cx1; cx++) h = std::max(h, map[(size_t)cy1 * ow + cx]);
        for (uint32_t cy = cy0; cy <= cy1; cy++) h = std::max(h, map[(size_t)cy * ow + cx1]);
        hgt[i] = h + 1;
        if (cy0 > 0) {
            uint32_t *row = &map[(size_t)(cy0 - 1) * ow];
            for (uint32_t cx = cx0 ? cx0 - 1 : 0; cx <= cx1; cx++) row[cx] = std::max(row[cx], h + 1);
        }
        if (cx0 > 0)
            for (uint32_t cy = cy0; cy <= cy1; cy++)
                map[(size_t)cy * ow + cx0 - 1] = std::max(map[(size_t)cy * ow + cx0 - 1], h + 1);
    }
    for (uint32_t i = 0; i < ni; i++) { /* clear the marks */
        const DBlock &k = blocks[intra[i]];
        const uint32_t cx0 = k.x >> 2, cy0 = k.y >> 2;
        const uint32_t cx1 = (k.x + k.bw - 1) >> 2, cy1 = (k.y + k.bh - 1) >> 2;
        if (cy0 > 0)
            for (uint32_t cx = cx0 ? cx0 - 1 : 0; cx <= cx1; cx++) map[(size_t)(cy0 - 1) * ow + cx] = 0;
        if (cx0 > 0)
            for (uint32_t cy = cy0; cy <= cy1; cy++) map[(size_t)cy * ow + cx0 - 1] = 0;
    }
    /* stable counting sort by (level, size class): small blocks of a level
     * first, so a level launch gives them 64-thread workers four to a
     * workgroup and the large ones a workgroup each */
    cnt.assign(2 * depth + 1, 0);
    for (uint32_t i = 0; i < ni; i++)
        cnt[2 * lev[i] + (intra_small(blocks[intra[i]]) ? 0 : 1) + 1]++;
    level_small.resize(depth);
    for (uint32_t l = 0; l < depth; l++) level_small[l] = cnt[2 * l + 1];
    for (uint32_t k = 0; k < 2 * depth; k++) cnt[k + 1] += cnt[k];
    levels.resize(depth + 1);
    for (uint32_t l = 0; l <= depth; l++) levels[l] = cnt[2 * l];
    tmp.resize(ni);
    pos.resize(ni);
    for (uint32_t i = 0; i < ni; i++) {
        const uint32_t at = cnt[2 * lev[i] + (intra_small(blocks[intra[i]]) ? 0 : 1)]++;
        tmp[at] = intra[i];
        pos[at] = i;
    }
    /* within a (level, size class) bucket, the blocks heading the longest
     * chains first (then list order): tickets go out in list order, so the
     * dependency chain that bounds the stage starts at once instead of
     * behind the level's other blocks */
    for (uint32_t k = 0; k < 2 * depth; k++) {
        const uint32_t a = k ? cnt[k - 1] : 0, e = cnt[k];
        std::stable_sort(pos.begin() + a, pos.begin() + e,
                         [&](uint32_t u, uint32_t v) { return hgt[u] > hgt[v]; });
    }
    for (uint32_t i = 0; i < ni; i++) tmp[i] = intra[pos[i]];
    memcpy(intra, tmp.data(), (size_t)ni * 4);
}

void plan_write(const vp9_parsed_frame_t *pf, PrepLayout &L, uint8_t *dst)
{
    const int W = (int)L.W, H = (int)L.H;
    DBlock *blocks = (DBlock *)(dst + L.off_blocks);
    uint32_t *mc_tab = (uint32_t *)(dst + L.off_mc_tab); /* 4 words per unit */
    /* units of blocks predicted from slot 0 go after all others */
    uint32_t nm_not0 = 0, nm_0 = L.n_mc_not0;
    uint32_t *intra = (uint32_t *)(dst + L.off_intra);
    uint32_t *txt[4], cn[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; k++) txt[k] = (uint32_t *)(dst + L.off_tx_tab[k]);
    uint32_t n = 0, ni = 0;
    for (uint32_t i = 0; i < pf->num_blocks; i++) {
        const vp9_macroblock_info_t &b = pf->blocks[i];
        const Clamped c = clamp_block(b, W, H);
        if (c.bw <= 0 || c.bh <= 0) continue;
        DBlock d;
        memset(&d, 0, sizeof(d));
        d.x = b.x;
        d.y = b.y;
        d.bw = (uint8_t)c.bw;
        d.bh = (uint8_t)c.bh;
        const bool is_intra = b.is_intra != 0;
        d.flags = is_intra ? DB_INTRA : 0;
        d.ymode = mode_class(b.y_mode);
        d.uvmode = mode_class(b.uv_mode);
        const int r = (int)b.ref_frame[0] - 1;
        d.ref = (uint8_t)((r >= 0 && r < 8) ? r : 8);
        d.mvx = b.mv[0][0];
        d.mvy = b.mv[0][1];
        d.coeff = b.coeff_offset;
        bool res = !b.skip;
        uint32_t ntx = 0;
        if (res) {
            const uint32_t T = 4u << b.tx_size;
            d.tx = b.tx_size;
            d.nx = (uint32_t)c.bw / T;
            ntx = d.nx * ((uint32_t)c.bh / T);
            if (ntx == 0) res = false;
        }
        if (!res) {
            d.flags |= DB_SKIP;
            d.tx = 0;
            d.nx = 0;
        }
        blocks[n] = d;
        if (is_intra) intra[ni++] = n;
        else {
            const uint32_t rows = cvp9_mc_unit_rows(d.bw);
            const uint32_t units = (d.bh + rows - 1) / rows;
            uint32_t &nm = d.ref == 0 ? nm_0 : nm_not0;
            for (uint32_t u = 0; u < units; u++, nm++) {
                uint32_t *m = mc_tab + 4 * (size_t)nm;
                const uint32_t skip = (d.flags & DB_SKIP) ? 1u : 0u;
                m[0] = (uint32_t)d.x | (uint32_t)d.y << 16;
                m[1] = (uint32_t)d.bw | (uint32_t)d.bh << 8 | ((uint32_t)d.ref | u << 4) << 16 |
                       ((uint32_t)d.tx | skip << 2 | d.nx << 3) << 24;
                m[2] = (uint32_t)(uint16_t)d.mvx | (uint32_t)(uint16_t)d.mvy << 16;
                m[3] = d.coeff;
            }
        }
        if (res) { /* transforms of inter and intra blocks alike, list order */
            const int k = d.tx;
            const uint32_t T = 4u << k;
            for (uint32_t l = 0; l < ntx; l++) {
                const uint32_t ty = l / d.nx, tx = l - ty * d.nx;
                uint32_t *e = txt[k] + 2 * (size_t)cn[k]++;
                e[0] = d.coeff + l * T * T;
                e[1] = (d.x + tx * T) | (d.y + ty * T) << 14 | (is_intra ? 1u << 31 : 0u);
            }
        }
        n++;
    }
    order_intra(blocks, intra, ni, L.W, L.H, L.levels, L.level_small);
    if (L.num_coeffs) memcpy(dst + L.off_coeffs, pf->coeffs, (size_t)L.num_coeffs * 2);
}

/* ---- frames and streams ------------------------------------------------ */

size_t frame_bytes(uint32_t w, uint32_t h, uint32_t *sy, uint32_t *suv)
{
    const uint32_t s0 = (w + 63) & ~63u;
    const uint32_t s1 = (((w + 1) / 2) + 63) & ~63u;
    if (sy) *sy = s0;
    if (suv) *suv = s1;
    /* + 64 bytes of slack: the motion-compensation fast path reads an
     * aligned 8-byte window that may extend 3 bytes past the last sample */
    return (size_t)s0 * h + 2 * (size_t)s1 * ((h + 1) / 2) + 64;
}

DFrame FrameBuf::dframe() const
{
    DFrame f;
    f.y = dev;
    f.u = dev ? dev + (size_t)sy * h : nullptr;
    f.v = dev ? f.u + (size_t)suv * ((h + 1) / 2) : nullptr;
    f.w = w;
    f.h = h;
    f.sy = sy;
    f.suv = suv;
    return f;
}

StreamState::~StreamState()
{
    hipSetDevice(device_);
    for (FrameBuf *f : all_) {
        hipFree(f->dev);
        delete f;
    }
    for (int p = 0; p < 2; p++) {
        hipFree(owner_[p]);
        hipFree(nbr_[p]);
        hipFree(resid_[p]);
    }
    hipFree(edge_);
    hipFree(ectr_);
    hipFree(errd_);
    if (err_) hipHostFree(err_);
}

bool StreamState::scratch_fits(const PrepLayout &L) const
{
    const size_t cells = (((size_t)((L.W + 3) / 4) * ((L.H + 3) / 4)) + 3) & ~(size_t)3;
    return err_ && errd_ && ectr_ && cells <= owner_cap_[par_] && (L.n_blocks ? L.n_blocks : 1) <= edge_cap_ &&
           (L.n_intra ? L.n_intra : 1) <= nbr_cap_[par_] &&
           (L.num_coeffs ? L.num_coeffs : 1) <= resid_cap_[par_];
}

hipError_t StreamState::acquire(uint32_t w, uint32_t h, FrameBuf **out, bool may_free,
                                bool avoid_prev)
{
    uint32_t sy, suv;
    const size_t need = frame_bytes(w, h, &sy, &suv);
    FrameBuf *pick = nullptr;
    for (FrameBuf *f : all_)
        if (f->refcnt == 0 && f->cap >= need && (!pick || f->cap < pick->cap) &&
            !(avoid_prev && read_by_prev(f)))
            pick = f;
    if (!pick) {
        /* drop idle buffers of other sizes before allocating */
        for (size_t i = 0; may_free && i < all_.size();) {
            if (all_[i]->refcnt == 0 && !(avoid_prev && read_by_prev(all_[i]))) {
                hipFree(all_[i]->dev);
                delete all_[i];
                all_[i] = all_.back();
                all_.pop_back();
            } else {
                i++;
            }
        }
        pick = new FrameBuf();
        hipError_t e = hipMalloc(&pick->dev, need);
        if (e != hipSuccess) {
            delete pick;
            return e;
        }
        pick->cap = need;
        all_.push_back(pick);
    }
    pick->w = w;
    pick->h = h;
    pick->sy = sy;
    pick->suv = suv;
    *out = pick;
    return hipSuccess;
}

void StreamState::release(FrameBuf *f)
{
    if (f && f->refcnt > 0) f->refcnt--;
}

void StreamState::update_refs(bool key, FrameBuf *out)
{
    last_written_ = key ? 1 : 0;
    for (int i = 0; i < 8; i++) {
        if (!key && i != 0) continue;
        FrameBuf *old = refs_[i];
        refs_[i] = out;
        out->refcnt++;
        release(old);
    }
}

hipError_t StreamState::reset()
{
    for (int i = 0; i < 8; i++) {
        release(refs_[i]);
        refs_[i] = nullptr;
    }
    last_written_ = 1; /* a frame still in flight may read any slot's old frame */
    if (err_) __atomic_store_n(err_, 0u, __ATOMIC_RELEASE);
    return errd_ ? hipMemset(errd_, 0, 256) : hipSuccess;
}

hipError_t StreamState::ensure_scratch(const PrepLayout &L, hipStream_t s)
{
    hipError_t e;
    const size_t cells = (((size_t)((L.W + 3) / 4) * ((L.H + 3) / 4)) + 3) & ~(size_t)3;
    if (!err_) {
        if ((e = hipHostMalloc((void **)&err_, 64, hipHostMallocCoherent)) != hipSuccess) {
            err_ = nullptr;
            return e;
        }
        memset(err_, 0, 64);
        if ((e = hipHostGetDevicePointer((void **)&err_dev_, err_, 0)) != hipSuccess) return e;
    }
    if (!errd_) {
        if ((e = hipMalloc(&errd_, 256)) != hipSuccess) return e;
        if ((e = hipMemsetAsync(errd_, 0, 256, s)) != hipSuccess) return e;
    }
    if (!ectr_) {
        if ((e = hipMalloc(&ectr_, 256)) != hipSuccess) return e;
        if ((e = hipMemsetAsync(ectr_, 0, 256, s)) != hipSuccess) return e;
    }
    const int p = par_;
    if (cells > owner_cap_[p]) {
        if (owner_[p]) {
            if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
            hipFree(owner_[p]);
            owner_[p] = nullptr;
        }
        if ((e = hipMalloc(&owner_[p], cells * 4)) != hipSuccess) return e;
        owner_cap_[p] = cells;
    }
    const size_t nb = L.n_blocks ? L.n_blocks : 1;
    if (nb > edge_cap_) {
        size_t cap = edge_cap_ ? edge_cap_ : 4096;
        while (cap < nb) cap *= 2;
        if (edge_) {
            if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
            hipFree(edge_);
            edge_ = nullptr;
        }
        const size_t bytes = cap * CVP9_EDGE_WORDS * sizeof(uint64_t);
        if ((e = hipMalloc(&edge_, bytes)) != hipSuccess) return e;
        /* epochs start at 1: zeroed words never look valid */
        if ((e = hipMemsetAsync(edge_, 0, bytes, s)) != hipSuccess) return e;
        edge_cap_ = cap;
    }
    const size_t ni = L.n_intra ? L.n_intra : 1;
    if (ni > nbr_cap_[p]) {
        size_t cap = nbr_cap_[p] ? nbr_cap_[p] : 1024;
        while (cap < ni) cap *= 2;
        if (nbr_[p]) {
            if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
            hipFree(nbr_[p]);
            nbr_[p] = nullptr;
        }
        /* rewritten by the neighbour stage before every read */
        if ((e = hipMalloc(&nbr_[p], cap * CVP9_NB_WORDS * sizeof(uint64_t))) != hipSuccess) return e;
        nbr_cap_[p] = cap;
    }
    const size_t nc = L.num_coeffs ? L.num_coeffs : 1;
    if (nc > resid_cap_[p]) {
        size_t cap = resid_cap_[p] ? resid_cap_[p] : 65536;
        while (cap < nc) cap *= 2;
        if (resid_[p]) {
            if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
            hipFree(resid_[p]);
            resid_[p] = nullptr;
        }
        if ((e = hipMalloc(&resid_[p], cap * sizeof(int16_t))) != hipSuccess) return e;
        resid_cap_[p] = cap;
    }
    return hipSuccess;
}

DJob StreamState::make_job(const PrepLayout &L, const uint8_t *dev, FrameBuf *out, uint8_t *nv12,
                           uint64_t serial)
{
    DJob j;
    memset(&j, 0, sizeof(j));
    j.out = out->dframe();
    for (int i = 0; i < 8; i++) {
        if (refs_[i]) j.ref[i] = refs_[i]->dframe();
        prev_read_[i] = refs_[i];
    }
    j.blocks = (const DBlock *)(dev + L.off_blocks);
    j.coeffs = (const int16_t *)(dev + L.off_coeffs);
    j.mc_tab = (const uint4 *)(dev + L.off_mc_tab);
    j.intra = (const uint32_t *)(dev + L.off_intra);
    for (int k = 0; k < 4; k++) {
        j.tx_tab[k] = (const uint32_t *)(dev + L.off_tx_tab[k]);
        j.cls_ntx[k] = L.cls_ntx[k];
    }
    j.owner = owner_[par_];
    j.edge = edge_;
    j.nbr = nbr_[par_];
    j.resid = resid_[par_];
    reader_[par_] = serial;
    if (two_sets_) par_ ^= 1;
    j.nv12 = nv12;
    j.err = errd_;
    j.err_host = err_dev_;
    j.ectr = ectr_;
    j.n_blocks = L.n_blocks;
    j.n_inter = L.n_inter;
    j.n_mc = L.n_mc;
    j.n_mc_a = L.n_mc; /* all early unless the engine splits them */
    j.n_intra = L.n_intra;
    j.ow = (L.W + 3) / 4;
    j.oh = (L.H + 3) / 4;
    j.epoch = 0; /* set on the device by the batch's first kernel */
    j.lf_thr = L.lf_thr;
    j.lf_wx = (L.W + 2 + 63) / 64;
    j.lf_wy = (L.H + 2 + 63) / 64;
    j.fill = L.fill;
    return j;
}

size_t StreamState::resident_bytes() const
{
    size_t b = (owner_cap_[0] + owner_cap_[1]) * 4 + edge_cap_ * CVP9_EDGE_WORDS * 8 +
               (nbr_cap_[0] + nbr_cap_[1]) * CVP9_NB_WORDS * 8 +
               (resid_cap_[0] + resid_cap_[1]) * 2 + (ectr_ ? 256 : 0);
    for (const FrameBuf *f : all_) b += f->cap;
    return b;
}

/* ---- batches ------------------------------------------------------------ */

size_t BatchBuilder::bytes_for(uint32_t njobs)
{
    return align256((uint64_t)njobs * sizeof(DJob)) + align256(13ull * (njobs + 1) * 4) + 256 +
           3ull * CVP9_MAX_HEAD_LEVELS * (njobs + 1) * 4;
}

void BatchBuilder::build(const DJob *jobs, const PrepLayout *const *layouts, uint32_t njobs,
                         uint8_t *host, uint8_t *dev, const Tune &tune, uint32_t spin_limit)
{
    memcpy(host, jobs, (size_t)njobs * sizeof(DJob));
    const uint64_t pre_off = align256((uint64_t)njobs * sizeof(DJob));
    uint32_t *pre = (uint32_t *)(host + pre_off);
    const uint32_t stride = njobs + 1;
    enum { P_FILL, P_BLK, P_INTER, P_INTER2, P_G0, P_G1, P_G2, P_G3, P_INTRA, P_LF, P_NV12, P_N };
    static_assert(P_N + 2 <= 13, "prefix rows, the ticket row and the start row fit the descriptor");
    for (int a = 0; a < P_N; a++) pre[a * stride] = 0;

    for (uint32_t i = 0; i < njobs; i++) {
        const DJob &J = jobs[i];
        uint32_t c[P_N];
        uint32_t fill = 0;
        if (J.fill) {
            const uint64_t fb = (uint64_t)J.out.sy * J.out.h + 2ull * J.out.suv * ((J.out.h + 1) / 2);
            const uint64_t ob = ((uint64_t)J.ow * J.oh * 4 + 15) & ~15ull;
            fill = (uint32_t)((fb + 65535) >> 16) + (uint32_t)((ob + 65535) >> 16);
        }
        c[P_FILL] = fill;
        c[P_BLK] = J.n_blocks;
        c[P_INTER] = (J.n_mc_a + CVP9_MC_WG_UNITS - 1) / CVP9_MC_WG_UNITS * CVP9_MC_WG_UNITS;
        c[P_INTER2] = (J.n_mc - J.n_mc_a + CVP9_MC_WG_UNITS - 1) / CVP9_MC_WG_UNITS * CVP9_MC_WG_UNITS;
        for (int k = 0; k < 4; k++) c[P_G0 + k] = div_up(J.cls_ntx[k], 64u >> (2 + k));
        /* 4x4: whole waves of four groups (64 transforms, one per lane) per job */
        c[P_G0] = div_up(c[P_G0], 4u) * 4u;
        c[P_INTRA] = J.n_intra;
        c[P_LF] = J.lf_thr > 0 ? J.lf_wx * J.lf_wy : 0;
        c[P_NV12] = J.nv12 ? div_up(J.out.w, 64) * div_up(J.out.h, 64) : 0;
        for (int a = 0; a < P_N; a++) pre[a * stride + i + 1] = pre[a * stride + i] + c[a];
    }
    const uint32_t *dpre = (const uint32_t *)(dev + pre_off);
    b_.jobs = (const DJob *)dev;
    b_.njobs = njobs;
    b_.pre_fill = dpre + P_FILL * stride;
    b_.pre_blocks = dpre + P_BLK * stride;
    b_.pre_inter = dpre + P_INTER * stride;
    b_.pre_inter2 = dpre + P_INTER2 * stride;
    for (int k = 0; k < 4; k++) b_.pre_grp[k] = dpre + (P_G0 + k) * stride;
    b_.pre_intra = dpre + P_INTRA * stride;
    b_.pre_lf = dpre + P_LF * stride;
    b_.pre_nv12 = dpre + P_NV12 * stride;
    b_.jticket = (uint32_t *)(dev + pre_off) + P_N * stride;
    b_.jstart = (const uint32_t *)(dev + pre_off) + (P_N + 1) * stride;
    const uint64_t tk_off = pre_off + align256(13ull * stride * 4);
    b_.bar = (uint32_t *)(dev + tk_off);
    memset(host + tk_off, 0, 256); /* the barrier counter starts at zero */
    b_.spin_limit = spin_limit;
    b_.fused_lanes = tune.intra_lanes;
    b_.fused_workers = tune.intra_workers;
    n_.fill = pre[P_FILL * stride + njobs];
    n_.blocks = pre[P_BLK * stride + njobs];
    n_.inter = pre[P_INTER * stride + njobs];
    n_.inter2 = pre[P_INTER2 * stride + njobs];
    for (int k = 0; k < 4; k++) n_.grp[k] = pre[(P_G0 + k) * stride + njobs];
    n_.intra = pre[P_INTRA * stride + njobs];
    n_.lf = pre[P_LF * stride + njobs];
    n_.nv12 = pre[P_NV12 * stride + njobs];
    n_.tune = tune;
    used_ = tk_off + 256;

    /* Leading levels at least tune.intra_level_min blocks wide (over the
     * batch) are launched one by one; the persistent workers take the rest,
     * each job's tickets starting where its first remaining level starts. */
    uint32_t *tick = pre + P_N * stride;        /* row P_N: per-job intra tickets */
    uint32_t *jst = pre + (P_N + 1) * stride;   /* row P_N + 1: where they start */
    uint32_t head = 0;
    level_n_.clear();
    level_ns_.clear();
    if (tune.intra_level_min > 0) {
        for (; head < CVP9_MAX_HEAD_LEVELS; head++) {
            uint32_t c = 0, cs = 0;
            for (uint32_t i = 0; i < njobs; i++) {
                const std::vector<uint32_t> &lv = layouts[i]->levels;
                if (head + 1 < lv.size()) {
                    c += lv[head + 1] - lv[head];
                    cs += layouts[i]->level_small[head];
                }
            }
            if (c == 0 || c < tune.intra_level_min) break;
            level_n_.push_back(c - cs);
            level_ns_.push_back(cs);
        }
    }
    uint32_t tail = 0;
    for (uint32_t i = 0; i < njobs; i++) {
        const std::vector<uint32_t> &lv = layouts[i]->levels;
        jst[i] = lv.empty() ? 0 : lv[std::min<size_t>(head, lv.size() - 1)];
        tick[i] = 0;
        tail += jobs[i].n_intra - jst[i];
    }
    n_.levels = head;
    n_.level_n = head ? level_n_.data() : nullptr;
    n_.level_ns = head ? level_ns_.data() : nullptr;
    n_.intra_tail = tail;
    b_.lvl_pre = b_.lvl_pre2 = b_.lvl_start = nullptr;
    if (head) {
        uint32_t *ps = (uint32_t *)(host + used_), *pl = ps + (size_t)head * stride,
                 *ls = pl + (size_t)head * stride;
        for (uint32_t l = 0; l < head; l++) {
            uint32_t as = 0, al = 0;
            for (uint32_t i = 0; i < njobs; i++) {
                const std::vector<uint32_t> &lv = layouts[i]->levels;
                const bool has = l + 1 < lv.size();
                const uint32_t cs = has ? layouts[i]->level_small[l] : 0;
                ps[(size_t)l * stride + i] = as;
                pl[(size_t)l * stride + i] = al;
                ls[(size_t)l * stride + i] = has ? lv[l] : 0;
                as += cs;
                al += has ? lv[l + 1] - lv[l] - cs : 0;
            }
            ps[(size_t)l * stride + njobs] = as;
            pl[(size_t)l * stride + njobs] = al;
            ls[(size_t)l * stride + njobs] = 0;
        }
        b_.lvl_pre = (const uint32_t *)(dev + used_);
        b_.lvl_pre2 = b_.lvl_pre + (size_t)head * stride;
        b_.lvl_start = b_.lvl_pre2 + (size_t)head * stride;
        used_ += 3ull * head * stride * 4;
    }
}

hipError_t BatchBuilder::launch(const DBatch &b, const StageCounts &n, hipStream_t s)
{
    for (int st = 0; st < STAGE_FUSED; st++) {
        hipError_t e = cvp9_launch_stage(st, b, n, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

} // namespace cvp9
