/*
 * cvp9_oracle.c — CPU restatement of the reference reconstruction.
 * TEST INFRASTRUCTURE ONLY (see cvp9_oracle.h). Build with -ffp-contract=off.
 */
#include "cvp9_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

/* ---- frame buffers: vp9_frame.c:8-63 ----------------------------------- */

int orc_frame_alloc(cvp9_frame_info_t *f, uint32_t width, uint32_t height)
{
    /* stride_y = W, stride_uv = (W+1)/2, every byte starts at 128
     * (vp9_frame.c:10-30) */
    f->width = width;
    f->height = height;
    f->stride_y = width;
    f->stride_uv = (width + 1) / 2;
    size_t ny = (size_t)f->stride_y * height;
    size_t nc = (size_t)f->stride_uv * ((height + 1) / 2);
    f->plane_y = (uint8_t *)malloc(ny ? ny : 1);
    f->plane_u = (uint8_t *)malloc(nc ? nc : 1);
    f->plane_v = (uint8_t *)malloc(nc ? nc : 1);
    if (!f->plane_y || !f->plane_u || !f->plane_v) {
        orc_frame_free(f);
        return 0;
    }
    memset(f->plane_y, 128, ny);
    memset(f->plane_u, 128, nc);
    memset(f->plane_v, 128, nc);
    f->pts = 0;
    return 1;
}

void orc_frame_free(cvp9_frame_info_t *f)
{
    if (!f) return;
    free(f->plane_y);
    free(f->plane_u);
    free(f->plane_v);
    f->plane_y = f->plane_u = f->plane_v = NULL;
}

int orc_frame_copy(cvp9_frame_info_t *dst, const cvp9_frame_info_t *src)
{
    /* vp9_frame.c:47-63: free, re-allocate at src size, copy planes + pts */
    if (!dst || !src) return 0;
    orc_frame_free(dst);
    if (!orc_frame_alloc(dst, src->width, src->height)) return 0;
    size_t ny = (size_t)src->stride_y * src->height;
    size_t nc = (size_t)src->stride_uv * ((src->height + 1) / 2);
    memcpy(dst->plane_y, src->plane_y, ny);
    memcpy(dst->plane_u, src->plane_u, nc);
    memcpy(dst->plane_v, src->plane_v, nc);
    dst->pts = src->pts;
    return 1;
}

/* ---- inverse transform: vp9_reconstruct.c:13-60 ------------------------ */

void orc_idct_matrix(int n, double *m)
{
    /* Row i scaled by sqrt(1/N) (i = 0) or sqrt(2/N); the cosine argument is
     * evaluated left to right as ((2j+1)*i*pi) / (2N)  (:19-28). */
    const double s0 = sqrt(1.0 / n);
    const double s1 = sqrt(2.0 / n);
    for (int i = 0; i < n; i++) {
        const double s = i == 0 ? s0 : s1;
        for (int j = 0; j < n; j++)
            m[i * n + j] = s * cos(((2.0 * j + 1.0) * i * 3.141592653589793) / (2.0 * n));
    }
}

void orc_idct_2d(const int16_t *in, int16_t *out, int n)
{
    double m[32 * 32];
    double t[32 * 32];
    orc_idct_matrix(n, m); /* rebuilt on every call, as the reference does */

    /* t = M^T * in: t[i][j] = sum_k M[k][i] * in[k][j], k ascending (:30-39) */
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double acc = 0.0;
            for (int k = 0; k < n; k++) acc += m[k * n + i] * in[k * n + j];
            t[i * n + j] = acc;
        }

    /* out = t * M: out[i][j] = sum_k t[i][k] * M[k][j], k ascending (:41-50),
     * then round half away from zero and saturate to int16 (:52-59) */
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double acc = 0.0;
            for (int k = 0; k < n; k++) acc += t[i * n + k] * m[k * n + j];
            int32_t r = (int32_t)(acc + (acc >= 0 ? 0.5 : -0.5));
            out[i * n + j] = (int16_t)clampi(r, -32768, 32767);
        }
}

/* ---- intra prediction: vp9_reconstruct.c:63-91 ------------------------- */

void orc_intra_predict(uint8_t *dst, int stride, int w, int h,
                       const uint8_t *above, const uint8_t *left,
                       uint8_t top_left, int mode)
{
    if (mode == 0) { /* DC: rounded mean over w above + h left samples */
        uint32_t sum = 0;
        for (int i = 0; i < w; i++) sum += above[i];
        for (int i = 0; i < h; i++) sum += left[i];
        uint8_t dc = (uint8_t)((sum + (uint32_t)((w + h) >> 1)) / (uint32_t)(w + h));
        for (int r = 0; r < h; r++) memset(dst + (size_t)r * stride, dc, (size_t)w);
    } else if (mode == 1) { /* V */
        for (int r = 0; r < h; r++) memcpy(dst + (size_t)r * stride, above, (size_t)w);
    } else if (mode == 2) { /* H */
        for (int r = 0; r < h; r++) memset(dst + (size_t)r * stride, left[r], (size_t)w);
    } else { /* every other mode value acts as TM (:83-90) */
        for (int r = 0; r < h; r++)
            for (int c = 0; c < w; c++)
                dst[(size_t)r * stride + c] =
                    (uint8_t)clampi(above[c] + left[r] - top_left, 0, 255);
    }
}

/* ---- motion compensation: vp9_reconstruct.c:94-131 --------------------- */

void orc_motion_compensate(uint8_t *dst, int dst_stride,
                           const uint8_t *ref, int ref_stride,
                           int w, int h, int block_x, int block_y,
                           int mv_x, int mv_y, int ref_w, int ref_h)
{
    /* position in 1/8 pel; integer part by arithmetic shift, phase = low 3 bits */
    const int px8 = block_x * 8 + mv_x;
    const int py8 = block_y * 8 + mv_y;
    const int ix = px8 >> 3, iy = py8 >> 3;
    const int fx = px8 & 7, fy = py8 & 7;
    for (int r = 0; r < h; r++) {
        const int y0 = clampi(iy + r, 0, ref_h - 1);
        const int y1 = clampi(iy + r + 1, 0, ref_h - 1);
        const uint8_t *row0 = ref + (size_t)y0 * ref_stride;
        const uint8_t *row1 = ref + (size_t)y1 * ref_stride;
        for (int c = 0; c < w; c++) {
            const int x0 = clampi(ix + c, 0, ref_w - 1);
            const int x1 = clampi(ix + c + 1, 0, ref_w - 1);
            /* horizontal taps per row, rounded, then vertical, rounded */
            uint32_t top = ((uint32_t)row0[x0] * (8 - fx) + (uint32_t)row0[x1] * fx + 4) >> 3;
            uint32_t bot = ((uint32_t)row1[x0] * (8 - fx) + (uint32_t)row1[x1] * fx + 4) >> 3;
            dst[(size_t)r * dst_stride + c] = (uint8_t)((top * (8 - fy) + bot * fy + 4) >> 3);
        }
    }
}

/* ---- loop filter: vp9_reconstruct.c:134-191 ----------------------------- */

int orc_lf_threshold(uint8_t level, uint8_t sharpness)
{
    /* :139-144 */
    uint32_t thr = level;
    if (sharpness > 0) {
        thr >>= (sharpness > 4) ? 2 : 1;
        if (thr > (uint32_t)(9 - sharpness)) thr = (uint32_t)(9 - sharpness);
    }
    if (thr == 0) thr = 1;
    return (int)thr;
}

/* One 4-tap edge across samples p1 p0 | q0 q1 (:154-165). */
static inline void lf_edge(uint8_t *p1, uint8_t *p0, uint8_t *q0, uint8_t *q1, int thr)
{
    const int a = *p1, b = *p0, c = *q0, d = *q1;
    if (abs(b - c) * 2 + abs(a - d) / 2 <= thr) {
        const int delta = clampi((3 * (c - b) + 4) >> 3, -thr, thr);
        *p0 = (uint8_t)clampi(b + delta, 0, 255);
        *q0 = (uint8_t)clampi(c - delta, 0, 255);
        *p1 = (uint8_t)clampi(a + delta / 2, 0, 255);
        *q1 = (uint8_t)clampi(d - delta / 2, 0, 255);
    }
}

void orc_loop_filter(cvp9_frame_info_t *f, uint8_t level, uint8_t sharpness)
{
    if (level == 0) return;
    const int thr = orc_lf_threshold(level, sharpness);
    uint8_t *y = f->plane_y;
    const size_t s = f->stride_y;
    /* horizontal edges y = 8, 16, ... < H, every column (:146-167) */
    for (uint32_t r = 8; r < f->height; r += 8)
        for (uint32_t c = 0; c < f->width; c++)
            lf_edge(&y[(r - 2) * s + c], &y[(r - 1) * s + c],
                    &y[r * s + c], &y[(r + 1) * s + c], thr);
    /* then vertical edges x = 8, 16, ... < W, every row (:169-190) */
    for (uint32_t c = 8; c < f->width; c += 8)
        for (uint32_t r = 0; r < f->height; r++) {
            uint8_t *row = &y[r * s];
            lf_edge(&row[c - 2], &row[c - 1], &row[c], &row[c + 1], thr);
        }
}

/* ---- frame reconstruction: vp9_reconstruct.c:194-370 ------------------- */

/* Gathers the above row, left column and top-left sample of a w x h region
 * at (x, y) of one plane; missing neighbours are 127 (:226-244). */
static uint8_t gather_neighbours(const uint8_t *plane, size_t stride, int x, int y,
                                 int w, int h, uint8_t *above, uint8_t *left)
{
    uint8_t tl = 127;
    if (y > 0) {
        memcpy(above, plane + (size_t)(y - 1) * stride + x, (size_t)w);
        if (x > 0) tl = plane[(size_t)(y - 1) * stride + x - 1];
    } else {
        memset(above, 127, (size_t)w);
    }
    if (x > 0) {
        for (int i = 0; i < h; i++) left[i] = plane[(size_t)(y + i) * stride + x - 1];
    } else {
        memset(left, 127, (size_t)h);
    }
    return tl;
}

static void fill_rect(uint8_t *p, size_t stride, int w, int h, uint8_t v)
{
    for (int r = 0; r < h; r++) memset(p + (size_t)r * stride, v, (size_t)w);
}

int orc_reconstruct_frame_ex(const vp9_parsed_frame_t *pf,
                             const cvp9_frame_info_t *refs[8],
                             cvp9_frame_info_t *out, int apply_lf)
{
    if (!orc_frame_alloc(out, pf->hdr.width, pf->hdr.height)) return 0;
    const int W = (int)pf->hdr.width, H = (int)pf->hdr.height;
    const size_t sy = out->stride_y, sc = out->stride_uv;
    uint8_t above[256], left[256];
    int16_t res[1024];

    for (uint32_t b = 0; b < pf->num_blocks; b++) {
        const vp9_macroblock_info_t *blk = &pf->blocks[b];
        const int bx = blk->x, by = blk->y;
        /* clamp to the frame, drop empty (:218-221) */
        const int bw = bx + blk->width > W ? W - bx : blk->width;
        const int bh = by + blk->height > H ? H - by : blk->height;
        if (bw <= 0 || bh <= 0) continue;

        /* reference for inter blocks: slot ref_frame[0]-1, none -> grey 127
         * (:250-251, :313-314) */
        const int ridx = blk->ref_frame[0] - 1;
        const cvp9_frame_info_t *ref = (!blk->is_intra && ridx >= 0 && ridx < 8) ? refs[ridx] : NULL;

        /* luma prediction (:224-264) */
        uint8_t *dy = out->plane_y + (size_t)by * sy + bx;
        if (blk->is_intra) {
            uint8_t tl = gather_neighbours(out->plane_y, sy, bx, by, bw, bh, above, left);
            orc_intra_predict(dy, (int)sy, bw, bh, above, left, tl, blk->y_mode);
        } else if (ref && ref->plane_y) {
            orc_motion_compensate(dy, (int)sy, ref->plane_y, (int)ref->stride_y, bw, bh,
                                  bx, by, blk->mv[0][0], blk->mv[0][1],
                                  (int)ref->width, (int)ref->height);
        } else {
            fill_rect(dy, sy, bw, bh, 127);
        }

        /* chroma prediction, 4:2:0 (:266-334) */
        const int cbx = bx / 2, cby = by / 2, cbw = (bw + 1) / 2, cbh = (bh + 1) / 2;
        uint8_t *du = out->plane_u + (size_t)cby * sc + cbx;
        uint8_t *dv = out->plane_v + (size_t)cby * sc + cbx;
        if (blk->is_intra) {
            uint8_t tl = gather_neighbours(out->plane_u, sc, cbx, cby, cbw, cbh, above, left);
            orc_intra_predict(du, (int)sc, cbw, cbh, above, left, tl, blk->uv_mode);
            tl = gather_neighbours(out->plane_v, sc, cbx, cby, cbw, cbh, above, left);
            orc_intra_predict(dv, (int)sc, cbw, cbh, above, left, tl, blk->uv_mode);
        } else if (ref && ref->plane_u && ref->plane_v) {
            /* chroma MV = luma MV / 2 truncated toward zero; bounds use half
             * the reference luma size (:321-327) */
            const int cmx = blk->mv[0][0] / 2, cmy = blk->mv[0][1] / 2;
            const int rw = (int)ref->width / 2, rh = (int)ref->height / 2;
            orc_motion_compensate(du, (int)sc, ref->plane_u, (int)ref->stride_uv, cbw, cbh,
                                  cbx, cby, cmx, cmy, rw, rh);
            orc_motion_compensate(dv, (int)sc, ref->plane_v, (int)ref->stride_uv, cbw, cbh,
                                  cbx, cby, cmx, cmy, rw, rh);
        } else {
            fill_rect(du, sc, cbw, cbh, 127);
            fill_rect(dv, sc, cbw, cbh, 127);
        }

        /* luma residual: floor(bw/T) x floor(bh/T) transforms in raster order,
         * T*T coefficients each from coeff_offset (:336-363) */
        if (!blk->skip) {
            const int T = 1 << (blk->tx_size + 2);
            const int nx = bw / T, ny = bh / T;
            for (int ty = 0; ty < ny; ty++)
                for (int tx = 0; tx < nx; tx++) {
                    const uint32_t off = blk->coeff_offset + (uint32_t)((ty * nx + tx) * T * T);
                    orc_idct_2d(&pf->coeffs[off], res, T);
                    for (int r = 0; r < T; r++) {
                        uint8_t *p = out->plane_y + (size_t)(by + ty * T + r) * sy + bx + tx * T;
                        for (int c = 0; c < T; c++)
                            p[c] = (uint8_t)clampi(p[c] + res[r * T + c], 0, 255);
                    }
                }
        }
    }

    if (apply_lf) orc_loop_filter(out, pf->hdr.filter_level, pf->hdr.sharpness_level);
    return 1;
}

int orc_reconstruct_frame(const vp9_parsed_frame_t *pf,
                          const cvp9_frame_info_t *refs[8],
                          cvp9_frame_info_t *out)
{
    return orc_reconstruct_frame_ex(pf, refs, out, 1);
}

/* ---- decoder-level reference pool: vp9_decoder.c:209-231 --------------- */

struct orc_decoder {
    cvp9_frame_info_t refs[8]; /* zeroed until the first frame: NULL planes */
};

orc_decoder_t *orc_decoder_new(void)
{
    return (orc_decoder_t *)calloc(1, sizeof(orc_decoder_t));
}

void orc_decoder_free(orc_decoder_t *d)
{
    if (!d) return;
    for (int i = 0; i < 8; i++) orc_frame_free(&d->refs[i]);
    free(d);
}

int orc_decoder_frame(orc_decoder_t *d, const vp9_parsed_frame_t *pf,
                      int64_t pts, cvp9_frame_info_t *out)
{
    const cvp9_frame_info_t *refs[8];
    for (int i = 0; i < 8; i++) refs[i] = &d->refs[i];
    memset(out, 0, sizeof(*out));
    if (!orc_reconstruct_frame(pf, refs, out)) return -1;
    out->pts = pts;
    /* KEY frames refresh all 8 slots, every other frame slot 0 only */
    if (pf->hdr.frame_type == VP9_FRAME_KEY) {
        for (int i = 0; i < 8; i++) orc_frame_copy(&d->refs[i], out);
    } else {
        orc_frame_copy(&d->refs[0], out);
    }
    return 0;
}

void orc_frame_pack_i420(const cvp9_frame_info_t *f, uint8_t *dst)
{
    const uint32_t cw = (f->width + 1) / 2, ch = (f->height + 1) / 2;
    for (uint32_t r = 0; r < f->height; r++)
        memcpy(dst + (size_t)r * f->width, f->plane_y + (size_t)r * f->stride_y, f->width);
    dst += (size_t)f->width * f->height;
    for (uint32_t r = 0; r < ch; r++)
        memcpy(dst + (size_t)r * cw, f->plane_u + (size_t)r * f->stride_uv, cw);
    dst += (size_t)cw * ch;
    for (uint32_t r = 0; r < ch; r++)
        memcpy(dst + (size_t)r * cw, f->plane_v + (size_t)r * f->stride_uv, cw);
}

int orc_decoder_run(orc_decoder_t *d, const vp9_parsed_frame_t *const *pfs,
                    int n, uint8_t *out)
{
    int done = 0;
    for (int i = 0; i < n; i++) {
        cvp9_frame_info_t f;
        if (orc_decoder_frame(d, pfs[i], i, &f) != 0) break;
        if (out) {
            orc_frame_pack_i420(&f, out);
            const size_t cw = (f.width + 1) / 2, ch = (f.height + 1) / 2;
            out += (size_t)f.width * f.height + 2 * cw * ch;
        }
        orc_frame_free(&f);
        done++;
    }
    return done;
}

/* ---- NV12: vaapi_entrypoint.c:131-153 ---------------------------------- */

void orc_nv12_pack(const cvp9_frame_info_t *f, uint8_t *dst)
{
    const uint32_t w = f->width, h = f->height;
    const uint32_t cw = (w + 1) / 2, ch = (h + 1) / 2;
    for (uint32_t r = 0; r < h; r++)
        memcpy(dst + (size_t)r * w, f->plane_y + (size_t)r * f->stride_y, w);
    uint8_t *uv = dst + (size_t)w * h;
    for (uint32_t r = 0; r < ch; r++) {
        const uint8_t *u = f->plane_u + (size_t)r * f->stride_uv;
        const uint8_t *v = f->plane_v + (size_t)r * f->stride_uv;
        uint8_t *o = uv + (size_t)r * 2 * cw;
        for (uint32_t c = 0; c < cw; c++) {
            o[2 * c] = u[c];
            o[2 * c + 1] = v[c];
        }
    }
}
