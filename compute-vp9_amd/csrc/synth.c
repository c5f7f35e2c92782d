/*
 * synth.c — deterministic generator of parsed VP9 frames (cvp9_hip/synth.h).
 *
 * Presets (the five benchmark workloads):
 *   0  1280x720 style: frame 0 KEY all intra, later frames all inter; about 6
 *      nonzero coefficients at zigzag 0..5, Laplacian scale 40; MV +-64 px;
 *      level 32, sharpness 0; qindex 60.
 *   1  denser: 25% of each transform nonzero at the lowest zigzag positions,
 *      Laplacian scale 120; 30% intra blocks on inter frames; qindex 20..200.
 *   2  as 1 with 32x32 transforms at p=0.5 where allowed, MV +-256 px and a
 *      loop-filter level drawn in 0..63 per frame.
 *   3  as 1 (the multi-stream mix uses one seed per stream).
 *   4  as 1 (serving streams).
 * Build with -ffp-contract=off: the Laplacian draw uses double arithmetic.
 */
#include "cvp9_hip/synth.h"

#include <stddef.h>
#include <stdlib.h>
#include <string.h>

/* ---- counter-based randomness ------------------------------------------ */

typedef struct { uint64_t s; } rng_t;

static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static inline uint64_t rng_next(rng_t *r)
{
    r->s += 0x9e3779b97f4a7c15ULL;
    return mix64(r->s);
}

/* uniform integer in [0, n) */
static inline uint32_t rng_below(rng_t *r, uint32_t n)
{
    return (uint32_t)(((rng_next(r) >> 32) * (uint64_t)n) >> 32);
}

static inline int rng_permille(rng_t *r, uint32_t pm)
{
    return rng_below(r, 1000) < pm;
}

/* Natural log of x in (0, 1], from the exponent bits plus the atanh series of
 * the mantissa; only IEEE double ops, so the result is the same everywhere. */
static double det_log(double x)
{
    union { double d; uint64_t u; } v = { x };
    int e = (int)((v.u >> 52) & 0x7ff) - 1023;
    v.u = (v.u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL; /* m in [1,2) */
    double m = v.d;
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    const double z = (m - 1.0) / (m + 1.0), z2 = z * z;
    double term = z, sum = 0.0;
    for (int k = 1; k <= 29; k += 2) {
        sum = sum + term / (double)k;
        term = term * z2;
    }
    return 2.0 * sum + (double)e * 0.6931471805599453;
}

/* Laplacian(0, scale) rounded to an integer and clipped to int16. */
static int16_t draw_laplace(rng_t *r, uint32_t scale)
{
    const uint64_t bits = rng_next(r);
    const double u = (double)((bits >> 11) + 1) * (1.0 / 9007199254740992.0); /* (0,1] */
    double mag = -(double)scale * det_log(u);
    if (mag > 32767.0) mag = 32767.0;
    const int m = (int)(mag + 0.5);
    return (int16_t)((bits & 1) ? -m : m);
}

/* ---- zigzag scan (vp9_entropy.c:18-38) --------------------------------- */

static uint16_t zz4[16], zz8[64], zz16[256], zz32[1024];
static int zz_ready;

static void build_zigzag(uint16_t *scan, int n)
{
    /* anti-diagonals d = r + c; odd d run top-right to bottom-left (row
     * ascending), even d run bottom-left to top-right (column ascending) */
    int i = 0;
    for (int d = 0; d <= 2 * (n - 1); d++) {
        for (int t = 0; t < n; t++) {
            const int r = (d & 1) ? t : d - t;
            const int c = (d & 1) ? d - t : t;
            if (r >= 0 && r < n && c >= 0 && c < n) scan[i++] = (uint16_t)(r * n + c);
        }
    }
}

static const uint16_t *zigzag(int tx)
{
    if (!zz_ready) {
        build_zigzag(zz4, 4);
        build_zigzag(zz8, 8);
        build_zigzag(zz16, 16);
        build_zigzag(zz32, 32);
        zz_ready = 1;
    }
    static const uint16_t *const t[4] = { zz4, zz8, zz16, zz32 };
    return t[tx];
}

/* ---- presets ----------------------------------------------------------- */

int cvp9_synth_preset(cvp9_synth_params_t *p, int preset,
                      uint32_t width, uint32_t height, uint64_t seed)
{
    memset(p, 0, sizeof(*p));
    p->width = width;
    p->height = height;
    p->seed = seed;
    p->split_permille = 500;
    p->filter_level = 32;
    p->sharpness = 0;
    switch (preset) {
    case 0:
        p->intra_permille = 0;
        p->coef_fixed = 6;
        p->lap_scale = 40;
        p->mv_range = 512;
        p->qindex_lo = p->qindex_hi = 60;
        return 0;
    case 1:
    case 3:
    case 4:
        p->intra_permille = 300;
        p->coef_permille = 250;
        p->lap_scale = 120;
        p->mv_range = 512;
        p->qindex_lo = 20;
        p->qindex_hi = 200;
        return 0;
    case 2:
        p->intra_permille = 300;
        p->coef_permille = 250;
        p->lap_scale = 120;
        p->tx32_permille = 500;
        p->mv_range = 2048;
        p->filter_level = -1;
        p->qindex_lo = 20;
        p->qindex_hi = 200;
        return 0;
    default:
        return -1;
    }
}

/* ---- frame generation --------------------------------------------------- */

typedef struct {
    vp9_parsed_frame_t *pf;
    const cvp9_synth_params_t *p;
    rng_t r;
    int key;
} gen_t;

static int push_block(gen_t *g, vp9_macroblock_info_t *b)
{
    vp9_parsed_frame_t *pf = g->pf;
    if (pf->num_blocks == pf->blocks_capacity) {
        uint32_t cap = pf->blocks_capacity ? pf->blocks_capacity * 2 : 1024;
        void *nb = realloc(pf->blocks, (size_t)cap * sizeof(*b));
        if (!nb) return -1;
        pf->blocks = (vp9_macroblock_info_t *)nb;
        pf->blocks_capacity = cap;
    }
    /* member stores leave padding unspecified (C11 6.2.6.1p6): copy, then
     * clear the padding bytes so frames are byte-identical run to run */
    unsigned char *d = (unsigned char *)&pf->blocks[pf->num_blocks++];
    memcpy(d, b, sizeof(*b));
    memset(d + offsetof(vp9_macroblock_info_t, skip) + 1, 0,
           offsetof(vp9_macroblock_info_t, coeff_offset) - offsetof(vp9_macroblock_info_t, skip) - 1);
    memset(d + offsetof(vp9_macroblock_info_t, eob) + 2, 0,
           sizeof(*b) - offsetof(vp9_macroblock_info_t, eob) - 2);
    return 0;
}

static int reserve_coeffs(gen_t *g, uint32_t n)
{
    vp9_parsed_frame_t *pf = g->pf;
    if ((uint64_t)pf->num_coeffs + n > pf->coeffs_capacity) {
        uint64_t cap = pf->coeffs_capacity ? pf->coeffs_capacity : 65536;
        while (cap < (uint64_t)pf->num_coeffs + n) cap *= 2;
        if (cap > 0xffffffffULL) return -1;
        void *nc = realloc(pf->coeffs, (size_t)cap * sizeof(int16_t));
        if (!nc) return -1;
        pf->coeffs = (int16_t *)nc;
        pf->coeffs_capacity = (uint32_t)cap;
    }
    return 0;
}

static int emit_block(gen_t *g, int x, int y, int size)
{
    const cvp9_synth_params_t *p = g->p;
    rng_t *r = &g->r;
    vp9_macroblock_info_t b;
    memset(&b, 0, sizeof(b));
    b.x = (uint16_t)x;
    b.y = (uint16_t)y;
    b.width = (uint8_t)size;
    b.height = (uint8_t)size;

    /* transform size among those that fit the (square) block */
    const int max_tx = size >= 32 ? 3 : (size == 16 ? 2 : (size == 8 ? 1 : 0));
    int tx;
    if (p->tx32_permille && max_tx == 3 && rng_permille(r, p->tx32_permille))
        tx = 3;
    else if (p->tx32_permille && max_tx == 3)
        tx = (int)rng_below(r, 3);
    else
        tx = (int)rng_below(r, (uint32_t)max_tx + 1);
    b.tx_size = (uint8_t)tx;

    static const uint8_t modes[4] = { 0, 1, 2, 9 }; /* DC, V, H, TM */
    b.is_intra = (uint8_t)(g->key || rng_permille(r, p->intra_permille));
    if (b.is_intra) {
        b.y_mode = modes[rng_below(r, 4)];
        b.uv_mode = modes[rng_below(r, 4)];
    } else {
        b.ref_frame[0] = (uint8_t)(1 + rng_below(r, 3));
        const uint32_t span = (uint32_t)(2 * p->mv_range + 1);
        b.mv[0][0] = (int16_t)((int)rng_below(r, span) - p->mv_range);
        b.mv[0][1] = (int16_t)((int)rng_below(r, span) - p->mv_range);
    }
    b.skip = (uint8_t)rng_permille(r, p->skip_permille);
    b.coeff_offset = g->pf->num_coeffs;

    if (!b.skip) {
        const int W = (int)p->width, H = (int)p->height;
        const int bw = x + size > W ? W - x : size;
        const int bh = y + size > H ? H - y : size;
        const int T = 4 << tx, area = T * T;
        const int nx = bw / T, ny = bh / T;
        const uint32_t total = (uint32_t)(nx * ny * area);
        if (reserve_coeffs(g, total)) return -1;
        int16_t *c = g->pf->coeffs + g->pf->num_coeffs;
        memset(c, 0, (size_t)total * sizeof(int16_t));
        const int nnz = p->coef_fixed ? (int)p->coef_fixed
                                      : (int)((area * p->coef_permille + 500) / 1000);
        const int cnt = nnz > area ? area : nnz;
        const uint16_t *scan = zigzag(tx);
        for (int t = 0; t < nx * ny; t++)
            for (int i = 0; i < cnt; i++) c[t * area + scan[i]] = draw_laplace(r, p->lap_scale);
        g->pf->num_coeffs += total;
        b.eob = (uint16_t)cnt;
    }
    return push_block(g, &b);
}

/* recursive quadtree split of one square node, children in Z order */
static int split_node(gen_t *g, int x, int y, int size)
{
    if (size > 8 && rng_permille(&g->r, g->p->split_permille)) {
        const int h = size / 2;
        for (int q = 0; q < 4; q++) {
            const int cx = x + (q & 1) * h, cy = y + (q >> 1) * h;
            if (cx < (int)g->p->width && cy < (int)g->p->height)
                if (split_node(g, cx, cy, h)) return -1;
        }
        return 0;
    }
    return emit_block(g, x, y, size);
}

vp9_parsed_frame_t *cvp9_synth_frame(const cvp9_synth_params_t *p, uint32_t index)
{
    if (!p || p->width == 0 || p->height == 0 || p->width > 65535 || p->height > 65535)
        return NULL;
    vp9_parsed_frame_t *pf = (vp9_parsed_frame_t *)calloc(1, sizeof(*pf));
    if (!pf) return NULL;
    gen_t g;
    g.pf = pf;
    g.p = p;
    g.r.s = mix64(p->seed ^ mix64(0x5eedULL + index));
    g.key = p->key_interval ? (index % p->key_interval == 0) : (index == 0);

    vp9_frame_header_t *h = &pf->hdr;
    h->frame_type = g.key ? VP9_FRAME_KEY : VP9_FRAME_NON_KEY;
    h->width = h->render_width = p->width;
    h->height = h->render_height = p->height;
    h->bit_depth = 8;
    h->subsampling_x = h->subsampling_y = 1;
    h->show_frame = true;
    h->filter_level = (uint8_t)(p->filter_level < 0 ? rng_below(&g.r, 64) : (uint32_t)p->filter_level);
    h->sharpness_level = (uint8_t)p->sharpness;
    h->base_qindex = (int16_t)(p->qindex_lo +
                               (int32_t)rng_below(&g.r, (uint32_t)(p->qindex_hi - p->qindex_lo + 1)));
    h->refresh_frame_flags = g.key ? 0xff : 0x01;
    h->ref_frame_idx[0] = 0;
    h->ref_frame_idx[1] = 1;
    h->ref_frame_idx[2] = 2;

    for (uint32_t sy = 0; sy < p->height; sy += 64)
        for (uint32_t sx = 0; sx < p->width; sx += 64)
            if (split_node(&g, (int)sx, (int)sy, 64)) {
                cvp9_synth_free(pf);
                return NULL;
            }
    if (!pf->coeffs) { /* keep a valid pointer even when all blocks skip */
        pf->coeffs = (int16_t *)calloc(1, sizeof(int16_t));
        pf->coeffs_capacity = 1;
    }
    return pf;
}

void cvp9_synth_free(vp9_parsed_frame_t *pf)
{
    if (!pf) return;
    free(pf->blocks);
    free(pf->coeffs);
    free(pf);
}

uint64_t cvp9_synth_algo_bytes(const vp9_parsed_frame_t *pf, int nv12)
{
    const int W = (int)pf->hdr.width, H = (int)pf->hdr.height;
    uint64_t b = 0;
    for (uint32_t i = 0; i < pf->num_blocks; i++) {
        const vp9_macroblock_info_t *k = &pf->blocks[i];
        const int bw = k->x + k->width > W ? W - k->x : k->width;
        const int bh = k->y + k->height > H ? H - k->y : k->height;
        b += 32;
        if (bw <= 0 || bh <= 0) continue;
        if (!k->skip) {
            const int T = 4 << k->tx_size;
            b += 2ull * (uint64_t)((bw / T) * (bh / T) * T * T);
        }
        if (!k->is_intra) {
            const int cbw = (bw + 1) / 2, cbh = (bh + 1) / 2;
            b += (uint64_t)((bw + 1) * (bh + 1) + 2 * (cbw + 1) * (cbh + 1));
        }
    }
    const uint64_t cw = (uint64_t)(W + 1) / 2, ch = (uint64_t)(H + 1) / 2;
    const uint64_t i420 = (uint64_t)W * H + 2 * cw * ch;
    return b + i420 * (nv12 ? 2 : 1);
}
