/*
 * engine.cpp — frame planning, device stream state and batch launch.
 */
#include "engine.hpp"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <memory>
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

/* ---- host thread pool ---------------------------------------------------- */

struct HostPool::Job {
    const std::function<void(int)> *fn = nullptr;
    int n = 0;
    std::atomic<int> next{0};
};

/* how long an idle worker watches for the next job before it sleeps: frames
 * arrive every few hundred microseconds, and a wake-up through the
 * condition variable costs tens of them */
constexpr auto kPoolSpin = std::chrono::microseconds(300);

HostPool &HostPool::get()
{
    static HostPool pool;
    return pool;
}

HostPool::HostPool()
{
    const unsigned hc = std::thread::hardware_concurrency();
    const unsigned n = hc > 1 ? std::min(7u, hc - 1) : 0u;
    for (unsigned i = 0; i < n; i++) th_.emplace_back([this] { worker(); });
}

HostPool::~HostPool()
{
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (std::thread &t : th_) t.join();
}

void HostPool::worker()
{
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        if (!stop_ && gen_.load(std::memory_order_relaxed) == seen) {
            lk.unlock();
            const auto t0 = std::chrono::steady_clock::now();
            while (gen_.load(std::memory_order_acquire) == seen &&
                   std::chrono::steady_clock::now() - t0 < kPoolSpin)
                __builtin_ia32_pause();
            lk.lock();
        }
        cv_.wait(lk, [&] { return stop_ || (job_ && gen_ != seen); });
        if (stop_) return;
        seen = gen_;
        Job *j = job_;
        users_++;
        lk.unlock();
        for (int i; (i = j->next.fetch_add(1)) < j->n;) (*j->fn)(i);
        lk.lock();
        if (--users_ == 0) idle_cv_.notify_all();
    }
}

void HostPool::run(int n, const std::function<void(int)> &fn)
{
    std::unique_lock<std::mutex> busy(run_mu_, std::try_to_lock);
    if (!busy.owns_lock() || th_.empty() || n <= 1) {
        for (int i = 0; i < n; i++) fn(i);
        return;
    }
    Job j;
    j.fn = &fn;
    j.n = n;
    {
        std::lock_guard<std::mutex> g(mu_);
        job_ = &j;
        gen_++;
    }
    cv_.notify_all();
    for (int i; (i = j.next.fetch_add(1)) < n;) fn(i);
    /* every task has been taken; wait for the workers still running one
     * (workers that had not joined the job by now never will) */
    std::unique_lock<std::mutex> lk(mu_);
    job_ = nullptr;
    idle_cv_.wait(lk, [&] { return users_ == 0; });
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

/* the same on a bitmap that several threads claim from at once */
bool claim_bits(std::atomic<uint64_t> *bits, uint64_t start, uint64_t len)
{
    while (len) {
        const uint64_t w = start >> 6, o = start & 63;
        const uint64_t n = len < 64 - o ? len : 64 - o;
        const uint64_t mask = (n == 64 ? ~0ull : ((1ull << n) - 1)) << o;
        if (bits[w].fetch_or(mask, std::memory_order_relaxed) & mask) return false;
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

/* Validation of blocks [c.b0, c.b1) and their counts. The first error of
 * the range (block index err_at); overlap with a block claimed earlier in
 * the bitmap is `conflict` (in a parallel pass the other block may come
 * later in the list, so the caller then repeats the validation serially). */
struct RangeResult {
    cvp9_err_t err = CVP9_OK;
    uint32_t err_at = UINT32_MAX;
    bool conflict = false;
    uint64_t covered = 0;
};

template <typename Bits>
void measure_range(const vp9_parsed_frame_t *pf, uint32_t W, uint32_t H, uint32_t ow, Bits &&bits,
                   PlanChunk &out, RangeResult &res)
{
    /* counted in locals, stored once: the ranges' records share cache lines */
    PlanChunk c;
    c.b0 = out.b0;
    c.b1 = out.b1;
    c.intra.swap(out.intra);
    c.intra.clear();
    RangeResult r;
    struct Store {
        PlanChunk &c0, &c;
        RangeResult &r0, &r;
        ~Store() { c0 = std::move(c); r0 = r; }
    } store{out, c, res, r};
    for (uint32_t i = c.b0; i < c.b1; i++) {
        const vp9_macroblock_info_t &b = pf->blocks[i];
        const Clamped cl = clamp_block(b, (int)W, (int)H);
        if (cl.bw <= 0 || cl.bh <= 0) continue;
        auto fail = [&](cvp9_err_t e) {
            r.err = e;
            r.err_at = i;
        };
        /* blocks on the 4x4 grid: neighbour ownership is exact per cell */
        if ((b.x & 3) || (b.y & 3) || (b.width & 3) || (b.height & 3)) return fail(CVP9_ERR_UNSUPPORTED);
        const bool intra = b.is_intra != 0;
        /* VP9 blocks are at most 64x64; intra blocks publish their edges in
         * fixed 64-sample records, motion compensation spreads a block row
         * over at most 16 lanes */
        if (cl.bw > CVP9_MAX_INTRA_DIM || cl.bh > CVP9_MAX_INTRA_DIM) return fail(CVP9_ERR_UNSUPPORTED);
        if (!b.skip) {
            if (b.tx_size > 3) return fail(CVP9_ERR_UNSUPPORTED);
            const uint32_t T = 4u << b.tx_size;
            const uint64_t ntx = (uint64_t)(cl.bw / T) * (cl.bh / T);
            if (ntx) {
                if (!pf->coeffs) return fail(CVP9_ERR_INVALID_DATA);
                if ((uint64_t)b.coeff_offset + ntx * T * T > pf->num_coeffs) return fail(CVP9_ERR_INVALID_DATA);
                c.ntx[b.tx_size] += (uint32_t)ntx;
            }
        }
        const uint32_t cx0 = b.x >> 2, cx1 = (b.x + cl.bw - 1) >> 2;
        const uint32_t cy0 = b.y >> 2, cy1 = (b.y + cl.bh - 1) >> 2;
        for (uint32_t cy = cy0; cy <= cy1; cy++)
            if (!claim_bits(bits, (uint64_t)cy * ow + cx0, cx1 - cx0 + 1)) {
                r.conflict = true;
                return fail(CVP9_ERR_UNSUPPORTED);
            }
        r.covered += (uint64_t)(cx1 - cx0 + 1) * (cy1 - cy0 + 1);
        if (intra) {
            c.intra.push_back(IntraGeom{c.n_blocks, (uint16_t)cx0, (uint16_t)cy0, (uint16_t)cx1,
                                        (uint16_t)cy1,
                                        cvp9_mode_reads(mode_class(b.y_mode), mode_class(b.uv_mode))});
            c.n_intra++;
        }
        c.n_blocks++;
        if (!intra) {
            const uint32_t rows = cvp9_mc_unit_rows((uint32_t)cl.bw);
            const uint32_t units = ((uint32_t)cl.bh + rows - 1) / rows;
            if (b.ref_frame[0] != 1) c.n_mc_not0 += units; /* slot ref_frame - 1 */
            else c.n_mc_0 += units;
        }
    }
}

} // namespace

cvp9_err_t plan_measure(const vp9_parsed_frame_t *pf, PrepLayout &L, int threads)
{
    if (!pf) return CVP9_ERR_INVALID_DATA;
    const bool compact = L.compact, chains = L.chains;
    L = PrepLayout();
    L.compact = compact;
    L.chains = chains;
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
    const uint64_t words = ((uint64_t)ow * oh + 63) / 64 + 1;
    const uint32_t nb = pf->num_blocks;
    int P = nb >= kParallelBlocks ? HostPool::get().threads() : 1;
    if (threads > 0 && threads < P) P = threads;
    std::vector<RangeResult> res((size_t)P);
    L.chunks.assign((size_t)P, PlanChunk());
    for (int c = 0; c < P; c++) {
        L.chunks[c].b0 = (uint32_t)((uint64_t)nb * c / P);
        L.chunks[c].b1 = (uint32_t)((uint64_t)nb * (c + 1) / P);
    }
    auto serial = [&]() {
        thread_local std::vector<uint64_t> bits;
        bits.assign(words, 0);
        L.chunks.assign(1, PlanChunk());
        L.chunks[0].b1 = nb;
        res.assign(1, RangeResult());
        measure_range(pf, W, H, ow, bits, L.chunks[0], res[0]);
    };
    if (P == 1) {
        serial();
    } else {
        thread_local std::unique_ptr<std::atomic<uint64_t>[]> abits;
        thread_local uint64_t acap = 0;
        if (acap < words) {
            abits.reset(new std::atomic<uint64_t>[words]);
            acap = words;
        }
        for (uint64_t w = 0; w < words; w++) abits[w].store(0, std::memory_order_relaxed);
        std::atomic<uint64_t> *const bits = abits.get();
        HostPool::get().run(P, [&](int c) { measure_range(pf, W, H, ow, bits, L.chunks[c], res[c]); });
        bool conflict = false;
        for (const RangeResult &r : res) conflict |= r.conflict;
        if (conflict) serial(); /* an overlap: the serial pass names the later block */
    }
    const RangeResult *first = nullptr;
    uint64_t covered = 0;
    for (const RangeResult &r : res) {
        covered += r.covered;
        if (r.err != CVP9_OK && (!first || r.err_at < first->err_at)) first = &r;
    }
    if (first) return first->err;

    for (const PlanChunk &c : L.chunks) {
        L.n_blocks += c.n_blocks;
        L.n_intra += c.n_intra;
        L.n_mc_not0 += c.n_mc_not0;
        L.n_mc += c.n_mc_not0 + c.n_mc_0;
        for (int k = 0; k < 4; k++) L.cls_ntx[k] += c.ntx[k];
    }
    L.n_inter = L.n_blocks - L.n_intra;
    L.fill = covered != (uint64_t)ow * oh;
    L.num_coeffs = pf->coeffs ? pf->num_coeffs : 0;

    uint64_t off = 0;
    L.off_blocks = off; off = align256(off + (uint64_t)L.n_blocks * sizeof(DBlock));
    L.off_mc_tab = off; off = align256(off + (uint64_t)L.n_mc * 16);
    L.off_intra = off;  off = align256(off + (uint64_t)L.n_intra * 4);
    L.off_px = off;     off = align256(off + (uint64_t)L.n_intra * 4);
    if (L.chains) {
        L.off_chain = off;
        off = align256(off + ((uint64_t)L.n_intra + kMaxChains + 1) * 4);
    }
    for (int k = 0; k < 4; k++) {
        L.off_tx_tab[k] = off; off = align256(off + (uint64_t)L.cls_ntx[k] * 8);
    }
    if (L.compact) {
        L.n_chunks = (L.num_coeffs + 63) / 64;
        L.off_cmask = off;  off = align256(off + (uint64_t)L.n_chunks * 8);
        L.off_cstart = off; off = align256(off + (uint64_t)L.n_chunks * 4);
        L.off_cvals = off;  off = align256(off + (uint64_t)L.num_coeffs * 2 + 2);
        L.host_bytes = off;
        L.off_coeffs = off; off = align256(off + (uint64_t)L.n_chunks * 64 * 2);
        L.upload = L.host_bytes; /* until plan_write knows the values' count */
    } else {
        L.n_chunks = 0;
        L.off_coeffs = off; off = align256(off + (uint64_t)L.num_coeffs * 2);
        L.host_bytes = L.upload = off;
    }
    L.bytes = off;
    return CVP9_OK;
}

/* Blocks that one 64-thread intra worker handles in a single pass
 * (at most 16x16, i.e. at most 4x4 cells). */
static inline bool intra_small(const IntraGeom &g)
{
    return g.cx1 - g.cx0 < 4 && g.cy1 - g.cy0 < 4;
}

/* A per-cell scratch map whose entries carry the pass that wrote them
 * (upper 16 bits) beside a 16-bit value, so a pass needs no clearing of what
 * the previous one wrote: an entry of another pass reads as 0. Levels and
 * chain heights stay below 2^16 (a dependency chain climbs one level per
 * 4x4 cell step right or down, at most 8192 steps across a 16384x16384
 * frame). */
struct StampMap {
    std::vector<uint32_t> v;
    uint32_t stamp = 0;
    void next_pass(size_t cells)
    {
        if (v.size() < cells) v.assign(cells, 0);
        if (++stamp == 0x10000) { /* the stamps wrapped: start afresh */
            std::fill(v.begin(), v.end(), 0u);
            stamp = 1;
        }
    }
    uint32_t get(size_t i) const { return (v[i] >> 16) == stamp ? (v[i] & 0xffff) : 0u; }
    void set(size_t i, uint32_t x) { v[i] = stamp << 16 | x; }
};

/* Orders the intra list by dependency level, then size class (small
 * first), then the length of the chain of later blocks depending on the
 * block (longest first), then list index; writes it to `intra` and the
 * position where each level starts (and its small blocks) to levels /
 * level_small. An intra block depends on the earlier intra blocks owning
 * its above row, left column and above-left sample (for chroma too: a
 * chroma neighbour lies in the same 4x4 luma cells). Level 0 has no such
 * block; otherwise the level is one more than the highest level among them.
 * Intra workers take blocks in this order, so every block a worker waits
 * for was handed out earlier, and the blocks of one level can all run at
 * once. Those neighbour samples lie on the owner's bottom row or right
 * column, so only those cells record the owner's level (plus one), in list
 * order: a cell without a mark of this pass when block b reads it is not
 * owned by an earlier intra block. g holds the intra blocks in list
 * order. */

/* Chains over the intra blocks (g in list order; lev, hgt their levels and
 * heights; key the sorted keys, whose low 32 bits are indices into g).
 * Each block continues the chain of one of its parents, the one of highest
 * level (then greatest height) whose chain has no next block yet; a parent
 * owns a cell the block's modes read, and such a cell always lies on the
 * parent's bottom row or right column, so only those cells carry a mark.
 * The chains of at least two blocks, the most critical first (level plus
 * height of the first block, then length), at most kMaxChains of them, are
 * written to `out` as intra-list positions after n_chains + 1 offsets; every
 * other position follows in list order. A worker that takes one chain in
 * order, or a subsequence of the list, meets levels in non-decreasing
 * order, which keeps the persistent stage free of deadlock. */
static void chain_intra(const std::vector<IntraGeom> &g, const std::vector<uint32_t> &lev,
                        const std::vector<uint32_t> &hgt, const std::vector<uint64_t> &key,
                        size_t ow, size_t cells, uint32_t *out, uint32_t &n_chains,
                        uint32_t &n_chained)
{
    constexpr uint32_t NONE = 0xffffffffu;
    const uint32_t ni = (uint32_t)g.size();
    thread_local std::vector<uint64_t> own; /* cell -> pass << 32 | (g index + 1) */
    thread_local uint32_t pass = 0;
    if (own.size() < cells) own.assign(cells, 0);
    if (++pass == 0) {
        std::fill(own.begin(), own.end(), 0ull);
        pass = 1;
    }
    thread_local std::vector<uint32_t> prev, next, pos, heads;
    prev.assign(ni, NONE);
    next.assign(ni, NONE);
    pos.resize(ni);
    for (uint32_t i = 0; i < ni; i++) pos[(uint32_t)key[i]] = i;
    for (uint32_t i = 0; i < ni; i++) {
        const IntraGeom &k = g[i];
        uint32_t best = NONE;
        auto consider = [&](size_t c) {
            const uint64_t v = own[c];
            if ((uint32_t)(v >> 32) != pass) return;
            const uint32_t o = (uint32_t)v - 1;
            if (next[o] != NONE) return;
            if (best == NONE || lev[o] > lev[best] || (lev[o] == lev[best] && hgt[o] > hgt[best]))
                best = o;
        };
        if (k.cy0 > 0) {
            const size_t row = (size_t)(k.cy0 - 1) * ow;
            if (k.cx0 > 0 && (k.reads & CVP9_READS_CORNER)) consider(row + k.cx0 - 1);
            if (k.reads & CVP9_READS_ABOVE)
                for (uint32_t cx = k.cx0; cx <= k.cx1; cx++) consider(row + cx);
        }
        if (k.cx0 > 0 && (k.reads & CVP9_READS_LEFT))
            for (uint32_t cy = k.cy0; cy <= k.cy1; cy++) consider((size_t)cy * ow + k.cx0 - 1);
        if (best != NONE) {
            prev[i] = best;
            next[best] = i;
        }
        const uint64_t mark = (uint64_t)pass << 32 | (i + 1);
        for (uint32_t cx = k.cx0; cx <= k.cx1; cx++) own[(size_t)k.cy1 * ow + cx] = mark;
        for (uint32_t cy = k.cy0; cy <= k.cy1; cy++) own[(size_t)cy * ow + k.cx1] = mark;
    }
    heads.clear();
    for (uint32_t i = 0; i < ni; i++)
        if (prev[i] == NONE && next[i] != NONE) heads.push_back(i);
    auto len = [&](uint32_t h) {
        uint32_t n = 0;
        for (uint32_t i = h; i != NONE; i = next[i]) n++;
        return n;
    };
    std::vector<std::pair<uint64_t, uint32_t>> rank;
    rank.reserve(heads.size());
    for (uint32_t h : heads) {
        const uint64_t crit = (uint64_t)(lev[h] + hgt[h]) << 32 | len(h);
        rank.emplace_back(~crit, h); /* ascending sort = most critical first */
    }
    std::sort(rank.begin(), rank.end());
    n_chains = (uint32_t)std::min<size_t>(rank.size(), kMaxChains);
    uint32_t *off = out, *ent = out + n_chains + 1;
    thread_local std::vector<uint8_t> taken;
    taken.assign(ni, 0);
    uint32_t n = 0;
    for (uint32_t c = 0; c < n_chains; c++) {
        off[c] = n;
        for (uint32_t i = rank[c].second; i != NONE; i = next[i]) {
            ent[n++] = pos[i];
            taken[pos[i]] = 1;
        }
    }
    off[n_chains] = n;
    n_chained = n;
    for (uint32_t p = 0; p < ni; p++)
        if (!taken[p]) ent[n++] = p;
}

/* Entries of the deferred sample pass: blocks of at most 8x8 (2x2 cells)
 * first, each group in list order. */
static uint32_t px_order(const std::vector<IntraGeom> &g, const uint32_t *entry_geom, uint32_t *px)
{
    const uint32_t ni = (uint32_t)g.size();
    uint32_t ns = 0;
    for (uint32_t i = 0; i < ni; i++) {
        const IntraGeom &q = g[entry_geom[i]];
        if (q.cx1 - q.cx0 < 2 && q.cy1 - q.cy0 < 2) px[ns++] = i;
    }
    uint32_t n = ns;
    for (uint32_t i = 0; i < ni; i++) {
        const IntraGeom &q = g[entry_geom[i]];
        if (!(q.cx1 - q.cx0 < 2 && q.cy1 - q.cy0 < 2)) px[n++] = i;
    }
    return ns;
}

static void order_intra(const std::vector<IntraGeom> &g, uint32_t *intra, uint32_t W, uint32_t H,
                        std::vector<uint32_t> &levels, std::vector<uint32_t> &level_small,
                        uint32_t *chain, uint32_t &n_chains, uint32_t &n_chained,
                        uint32_t *px = nullptr, uint32_t *n_px_small = nullptr)
{
    levels.clear();
    level_small.clear();
    n_chains = n_chained = 0;
    const uint32_t ni = (uint32_t)g.size();
    if (n_px_small) *n_px_small = 0;
    if (ni == 0) {
        if (chain) chain[0] = 0;
        return;
    }
    if (ni == 1) {
        intra[0] = g[0].b;
        if (px) {
            const uint32_t at0 = 0;
            *n_px_small = px_order(g, &at0, px);
        }
        levels = {0, 1};
        level_small = {intra_small(g[0]) ? 1u : 0u};
        if (chain) {
            chain[0] = 0; /* no chains; the one entry */
            chain[1] = 0;
        }
        return;
    }
    const size_t ow = (W + 3) / 4, cells = ow * ((H + 3) / 4);
    thread_local StampMap map;
    thread_local std::vector<uint32_t> lev, hgt;
    thread_local std::vector<uint64_t> key, tmp;
    lev.resize(ni);
    hgt.resize(ni);
    uint32_t depth = 0;
    map.next_pass(cells);
    for (uint32_t i = 0; i < ni; i++) {
        const IntraGeom &k = g[i];
        uint32_t m = 0;
        if (k.cy0 > 0) {
            const size_t row = (size_t)(k.cy0 - 1) * ow;
            if (k.cx0 > 0 && (k.reads & CVP9_READS_CORNER)) m = std::max(m, map.get(row + k.cx0 - 1));
            if (k.reads & CVP9_READS_ABOVE)
                for (uint32_t cx = k.cx0; cx <= k.cx1; cx++) m = std::max(m, map.get(row + cx));
        }
        if (k.cx0 > 0 && (k.reads & CVP9_READS_LEFT))
            for (uint32_t cy = k.cy0; cy <= k.cy1; cy++) m = std::max(m, map.get((size_t)cy * ow + k.cx0 - 1));
        lev[i] = m; /* level of this block */
        depth = std::max(depth, m + 1);
        for (uint32_t cx = k.cx0; cx <= k.cx1; cx++) map.set((size_t)k.cy1 * ow + cx, m + 1);
        for (uint32_t cy = k.cy0; cy <= k.cy1; cy++) map.set((size_t)cy * ow + k.cx1, m + 1);
    }
    /* height: the longest chain of later intra blocks that depend on a
     * block (1 without any), from a pass in reverse list order. A later
     * block leaves its height on the cells it reads (the row above it from
     * the above-left cell on, and the column left of it); a block's
     * dependents are exactly the later blocks that read its bottom row or
     * right column, and blocks earlier in the list write their marks only
     * after it has been scanned. */
    map.next_pass(cells);
    for (uint32_t i = ni; i-- > 0;) {
        const IntraGeom &k = g[i];
        uint32_t h = 0;
        for (uint32_t cx = k.cx0; cx <= k.cx1; cx++) h = std::max(h, map.get((size_t)k.cy1 * ow + cx));
        for (uint32_t cy = k.cy0; cy <= k.cy1; cy++) h = std::max(h, map.get((size_t)cy * ow + k.cx1));
        hgt[i] = h + 1;
        auto mark = [&](size_t c) { map.set(c, std::max(map.get(c), h + 1)); };
        if (k.cy0 > 0) {
            const size_t row = (size_t)(k.cy0 - 1) * ow;
            if (k.cx0 > 0 && (k.reads & CVP9_READS_CORNER)) mark(row + k.cx0 - 1);
            if (k.reads & CVP9_READS_ABOVE)
                for (uint32_t cx = k.cx0; cx <= k.cx1; cx++) mark(row + cx);
        }
        if (k.cx0 > 0 && (k.reads & CVP9_READS_LEFT))
            for (uint32_t cy = k.cy0; cy <= k.cy1; cy++) mark((size_t)cy * ow + k.cx0 - 1);
    }
    /* one stable sort on the key (level, size class, longest chain first):
     * three passes of an 11-bit LSD radix sort over entries in list order,
     * so equal keys keep list order */
    key.resize(ni);
    tmp.resize(ni);
    level_small.assign(depth, 0);
    for (uint32_t i = 0; i < ni; i++) {
        const uint32_t small = intra_small(g[i]) ? 1u : 0u;
        level_small[lev[i]] += small;
        key[i] = ((uint64_t)lev[i] << 17 | (uint64_t)(1 - small) << 16 |
                  (uint64_t)(0xffff - std::min(hgt[i], 0xffffu))) << 32 | i;
    }
    for (int pass = 0; pass < 3; pass++) {
        const int sh = 32 + 11 * pass;
        uint32_t count[2049] = {0};
        for (uint32_t i = 0; i < ni; i++) count[((key[i] >> sh) & 2047) + 1]++;
        for (int d = 0; d < 2048; d++) count[d + 1] += count[d];
        for (uint32_t i = 0; i < ni; i++) tmp[count[(key[i] >> sh) & 2047]++] = key[i];
        key.swap(tmp);
    }
    levels.assign(depth + 1, 0);
    for (uint32_t i = 0; i < ni; i++) {
        const uint32_t at = (uint32_t)key[i];
        intra[i] = g[at].b;
        levels[lev[at] + 1]++;
    }
    if (px) {
        thread_local std::vector<uint32_t> at_of;
        at_of.resize(ni);
        for (uint32_t i = 0; i < ni; i++) at_of[i] = (uint32_t)key[i];
        *n_px_small = px_order(g, at_of.data(), px);
    }
    for (uint32_t l = 0; l < depth; l++) levels[l + 1] += levels[l];
    if (chain) chain_intra(g, lev, hgt, key, ow, cells, chain, n_chains, n_chained);
}

namespace {

/* Where a range of the block list writes: its first local block index,
 * intra-list entry, motion-compensation unit of each kind (units of blocks
 * predicted from slot 0 come after all others) and transform of each size
 * class. */
struct RangeBase {
    uint32_t n, nm_not0, nm_0, tx[4];
};

void write_range(const vp9_parsed_frame_t *pf, const PrepLayout &L, uint8_t *dst,
                 const PlanChunk &c, RangeBase at)
{
    const int W = (int)L.W, H = (int)L.H;
    DBlock *blocks = (DBlock *)(dst + L.off_blocks);
    uint32_t *mc_tab = (uint32_t *)(dst + L.off_mc_tab); /* 4 words per unit */
    uint32_t *txt[4];
    for (int k = 0; k < 4; k++) txt[k] = (uint32_t *)(dst + L.off_tx_tab[k]);
    uint32_t n = at.n, nm_not0 = at.nm_not0, nm_0 = at.nm_0;
    uint32_t cn[4] = {at.tx[0], at.tx[1], at.tx[2], at.tx[3]};
    for (uint32_t i = c.b0; i < c.b1; i++) {
        const vp9_macroblock_info_t &b = pf->blocks[i];
        const Clamped cl = clamp_block(b, W, H);
        if (cl.bw <= 0 || cl.bh <= 0) continue;
        DBlock d;
        memset(&d, 0, sizeof(d));
        d.x = b.x;
        d.y = b.y;
        d.bw = (uint8_t)cl.bw;
        d.bh = (uint8_t)cl.bh;
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
            d.nx = (uint32_t)cl.bw / T;
            ntx = d.nx * ((uint32_t)cl.bh / T);
            if (ntx == 0) res = false;
        }
        if (!res) {
            d.flags |= DB_SKIP;
            d.tx = 0;
            d.nx = 0;
        }
        blocks[n] = d;
        if (!is_intra) {
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
            uint32_t *e = txt[k] + 2 * (size_t)cn[k];
            for (uint32_t ty = 0, l = 0; ty < ntx / d.nx; ty++)
                for (uint32_t tx = 0; tx < d.nx; tx++, l++, e += 2) {
                    e[0] = d.coeff + l * T * T;
                    e[1] = (d.x + tx * T) | (d.y + ty * T) << 14 | (is_intra ? 1u << 31 : 0u);
                }
            cn[k] += ntx;
        }
        n++;
    }
}

/* Nonzero mask of 64 pool entries: compares and byte masks of whole
 * vector registers (a boolean vector's bits are its lanes). */
typedef short v64s __attribute__((ext_vector_type(64)));
typedef bool v64b __attribute__((ext_vector_type(64)));
inline uint64_t nonzero_mask64(const int16_t *p)
{
    v64s v;
    memcpy(&v, p, sizeof(v));
    const v64b nz = __builtin_convertvector(v != 0, v64b);
    return __builtin_bit_cast(uint64_t, nz);
}

/* Compact coefficients, first pass: the nonzero mask of each chunk
 * [c0, c1) of 64 pool entries (entries past the pool's end count as
 * zero); returns the chunks' nonzero count. */
uint32_t chunk_masks(const int16_t *pool, uint32_t n, uint32_t c0, uint32_t c1, uint64_t *mask)
{
    uint32_t cnt = 0;
    for (uint32_t c = c0; c < c1; c++) {
        const uint32_t a = c * 64, e = std::min(n, a + 64);
        uint64_t m = 0;
        if (e - a == 64) {
            m = nonzero_mask64(pool + a);
        } else {
            for (uint32_t i = a; i < e; i++)
                if (pool[i]) m |= 1ull << (i - a);
        }
        mask[c] = m;
        cnt += (uint32_t)__builtin_popcountll(m);
    }
    return cnt;
}

/* Second pass: each chunk's position among the packed values, and the
 * values, from position pos on. */
void chunk_values(const int16_t *pool, uint32_t c0, uint32_t c1, const uint64_t *mask, uint32_t pos,
                  uint32_t *start, int16_t *vals)
{
    for (uint32_t c = c0; c < c1; c++) {
        start[c] = pos;
        const int16_t *p = pool + (size_t)c * 64;
        for (uint64_t m = mask[c]; m; m &= m - 1) vals[pos++] = p[__builtin_ctzll(m)];
    }
}

} // namespace

void plan_write(const vp9_parsed_frame_t *pf, PrepLayout &L, uint8_t *dst)
{
    const int P = (int)L.chunks.size();
    std::vector<RangeBase> at((size_t)P);
    RangeBase acc = {0, 0, L.n_mc_not0, {0, 0, 0, 0}};
    for (int c = 0; c < P; c++) {
        at[c] = acc;
        const PlanChunk &k = L.chunks[c];
        acc.n += k.n_blocks;
        acc.nm_not0 += k.n_mc_not0;
        acc.nm_0 += k.n_mc_0;
        for (int t = 0; t < 4; t++) acc.tx[t] += k.ntx[t];
    }
    /* the parser's coefficients, verbatim, in P slices; or, compact, each
     * slice's chunk masks and then its values, which start where the
     * previous slice's end: a slice waits for the previous one's count
     * (tasks are taken in index order, and a serial run takes them in
     * that order too, so the slice it waits for is running or done), then
     * packs its values while its coefficients are still in cache */
    const size_t cbytes = (size_t)L.num_coeffs * 2;
    const int16_t *pool = (const int16_t *)pf->coeffs;
    uint64_t *cmask = (uint64_t *)(dst + L.off_cmask);
    std::vector<uint32_t> slice_end((size_t)P + 1, 0);
    std::atomic<int> counted{0}; /* slices whose end position is known */
    auto copy_slice = [&](int c) {
        if (L.compact) {
            const uint32_t a = (uint32_t)((uint64_t)L.n_chunks * c / P),
                           b = (uint32_t)((uint64_t)L.n_chunks * (c + 1) / P);
            const uint32_t cnt = chunk_masks(pool, L.num_coeffs, a, b, cmask);
            while (counted.load(std::memory_order_acquire) < c) __builtin_ia32_pause();
            const uint32_t base = slice_end[c];
            slice_end[c + 1] = base + cnt;
            counted.store(c + 1, std::memory_order_release);
            chunk_values(pool, a, b, cmask, base, (uint32_t *)(dst + L.off_cstart),
                         (int16_t *)(dst + L.off_cvals));
            return;
        }
        const size_t a = cbytes * c / P, b = cbytes * (c + 1) / P;
        if (b > a) memcpy(dst + L.off_coeffs + a, (const uint8_t *)pf->coeffs + a, b - a);
    };
    /* the intra order (the longest task, a serial pass over the intra
     * blocks) beside the tables and the coefficient slices */
    auto order = [&]() {
        thread_local std::vector<IntraGeom> g;
        g.clear();
        for (int c = 0; c < P; c++)
            for (IntraGeom x : L.chunks[c].intra) {
                x.b += at[c].n;
                g.push_back(x);
            }
        order_intra(g, (uint32_t *)(dst + L.off_intra), L.W, L.H, L.levels, L.level_small,
                    L.chains ? (uint32_t *)(dst + L.off_chain) : nullptr, L.n_chains, L.n_chained,
                    (uint32_t *)(dst + L.off_px), &L.n_px_small);
    };
    if (P == 1) {
        write_range(pf, L, dst, L.chunks[0], at[0]);
        copy_slice(0);
        order();
    } else {
        HostPool::get().run(2 * P + 1, [&](int t) {
            if (t == 0) order();
            else if (t <= P) write_range(pf, L, dst, L.chunks[t - 1], at[t - 1]);
            else copy_slice(t - 1 - P);
        });
    }
    if (L.compact) {
        L.nnz = slice_end[P];
        /* at least one value: the device reads entry 0 for zero entries */
        if (L.nnz == 0) *(int16_t *)(dst + L.off_cvals) = 0;
        L.upload = L.off_cvals + 2ull * (L.nnz ? L.nnz : 1);
    }
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
    j.px = (const uint32_t *)(dev + L.off_px);
    j.n_px_small = L.n_px_small;
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
    if (L.chains && L.n_intra) {
        j.chain = (const uint32_t *)(dev + L.off_chain);
        j.n_chains = L.n_chains;
    }
    if (L.compact && L.n_chunks) {
        j.cmask = (const uint64_t *)(dev + L.off_cmask);
        j.cstart = (const uint32_t *)(dev + L.off_cstart);
        j.cvals = (const int16_t *)(dev + L.off_cvals);
        j.n_chunks = L.n_chunks;
    }
    j.n_blocks = L.n_blocks;
    j.n_inter = L.n_inter;
    j.n_mc = L.n_mc;
    j.n_mc_a = L.n_mc; /* all early unless the engine splits them */
    j.n_intra = L.n_intra;
    j.ow = (L.W + 3) / 4;
    j.oh = (L.H + 3) / 4;
    j.epoch = 0; /* set on the device by the batch's first kernel */
    j.lf_thr = L.lf_thr;
    j.lf_wx = ((L.W + 3) / 8 + 1 + 63) / 64; /* waves of 64 deblocking units per band */
    j.lf_wy = (L.H + 1) / 8 + 1;             /* bands of 8 rows at 8k - 2 */
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
    return align256((uint64_t)njobs * sizeof(DJob)) + align256(15ull * (njobs + 1) * 4) + CVP9_BAR_BYTES +
           3ull * CVP9_MAX_HEAD_LEVELS * (njobs + 1) * 4;
}

void BatchBuilder::build(const DJob *jobs, const PrepLayout *const *layouts, uint32_t njobs,
                         uint8_t *host, uint8_t *dev, const Tune &tune, uint32_t spin_limit, bool fused)
{
    memcpy(host, jobs, (size_t)njobs * sizeof(DJob));
    const uint64_t pre_off = align256((uint64_t)njobs * sizeof(DJob));
    uint32_t *pre = (uint32_t *)(host + pre_off);
    const uint32_t stride = njobs + 1;
    enum { P_FILL, P_BLK, P_INTER, P_INTER2, P_G0, P_G1, P_G2, P_G3, P_INTRA, P_LF, P_NV12, P_EXP, P_PX, P_N };
    static_assert(P_N + 2 <= 15, "prefix rows, the ticket row and the start row fit the descriptor");
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
        c[P_EXP] = div_up(J.n_chunks, 4u);
        c[P_PX] = div_up(J.n_px_small, 4u) + (J.n_intra - J.n_px_small);
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
    b_.pre_exp = dpre + P_EXP * stride;
    b_.pre_px = dpre + P_PX * stride;
    b_.jticket = (uint32_t *)(dev + pre_off) + P_N * stride;
    b_.jstart = (const uint32_t *)(dev + pre_off) + (P_N + 1) * stride;
    const uint64_t tk_off = pre_off + align256(15ull * stride * 4);
    b_.bar = (uint32_t *)(dev + tk_off);
    /* the fused kernel's barrier words start at zero */
    const size_t bar_bytes = fused ? CVP9_BAR_BYTES : 256;
    memset(host + tk_off, 0, bar_bytes);
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
    n_.exp = pre[P_EXP * stride + njobs];
    n_.px = pre[P_PX * stride + njobs];
    n_.chains = njobs == 1 ? jobs[0].n_chains : 0;
    n_.chained = njobs == 1 && jobs[0].n_chains ? layouts[0]->n_chained : 0;
    n_.tune = tune;
    used_ = tk_off + bar_bytes;

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
    if (hipError_t e = cvp9_launch_stage(STAGE_EXPAND, b, n, s)) return e;
    for (int st = 0; st < STAGE_FUSED; st++) {
        hipError_t e = cvp9_launch_stage(st, b, n, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

} // namespace cvp9
