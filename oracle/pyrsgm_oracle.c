/*
 * pyrsgm_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see pyrsgm_oracle.h).
 *
 * A plain, single-threaded, loop-by-loop restatement of the semantics of the
 * reference kernels.  It is deliberately naive: every stage is a separate
 * pass over dense host buffers, in the order the reference host code runs
 * them.  Citations are "file:line" in gegeges/pyrsgm.
 *
 * Arithmetic is C integer arithmetic (arithmetic >> on negatives, truncating
 * /), exactly like the CUDA kernels.  The only floating point is the
 * confidence ratio, the disparity/16 conversion and the adaptive-P2 table.
 * The adaptive-P2 table uses glibc expf where the reference uses CUDA __expf
 * (sgm_kernels.cu:68); values where the two could round differently are not
 * pinned by anything in the reference.
 *
 * Build with -ffp-contract=off (oracle/Makefile) so float expressions are
 * evaluated operation by operation.
 */
#include "pyrsgm_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXS SHRT_MAX

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
/* sgm_kernels.cu:49-51 */
static inline int16_t sat16(int v) { return (int16_t)imin(SHRT_MAX, imax(SHRT_MIN, v)); }

/* ------------------------------------------------------------------------
 * pyrDown: pyrdown.cu:16-52 (kernel), :56-74 (argument checks).
 * ---------------------------------------------------------------------- */
static int reflect101(int i, int n)
{
    if (i < 0) i = -i;
    if (i >= n) i = 2 * n - 2 - i;
    return i < 0 ? 0 : i;
}

int orc_pyrdown_u8(const uint8_t* src, size_t sstep, int sw, int sh,
                   uint8_t* dst, size_t dstep, int dw, int dh)
{
    static const int k[5] = {1, 4, 6, 4, 1};
    if (!src || !dst || sw < 2 || sh < 2 || dw != (sw + 1) / 2 || dh != (sh + 1) / 2) return -1;
    for (int y = 0; y < dh; ++y) {
        for (int x = 0; x < dw; ++x) {
            int sum = 0;
            for (int ky = 0; ky < 5; ++ky) {
                const uint8_t* row = src + (size_t)reflect101(2 * y + ky - 2, sh) * sstep;
                int rs = 0;
                for (int kx = 0; kx < 5; ++kx) rs += k[kx] * (int)row[reflect101(2 * x + kx - 2, sw)];
                sum += k[ky] * rs;
            }
            dst[(size_t)y * dstep + x] = (uint8_t)((sum + 128) >> 8);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------
 * Census 5x5: sgm_kernels.cu:251-298.  Replicate border, bits MSB-first in
 * row-major window order, bit = neighbour < centre.
 * ---------------------------------------------------------------------- */
void orc_census5x5(const uint8_t* img, size_t step, int w, int h, uint32_t* out)
{
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const int c = img[(size_t)y * step + x];
            uint32_t cv = 0;
            for (int dy = -2; dy <= 2; ++dy) {
                const int yy = iclamp(y + dy, 0, h - 1);
                for (int dx = -2; dx <= 2; ++dx) {
                    const int xx = iclamp(x + dx, 0, w - 1);
                    cv = (cv << 1) | (uint32_t)(img[(size_t)yy * step + xx] < c);
                }
            }
            out[(size_t)y * w + x] = cv;
        }
    }
}

static inline int popc32(uint32_t v) { return __builtin_popcount(v); }

/* sgm_kernels.cu:303-327: x2 = x-(d+minD) out of range -> 24. */
void orc_census_cost(const uint32_t* cl, const uint32_t* cr, int w, int h,
                     int D, int minD, int16_t* cost)
{
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int d = 0; d < D; ++d) {
                const int x2 = x - (d + minD);
                int16_t c = 24;
                if (x2 >= 0 && x2 < w) c = (int16_t)popc32(cl[(size_t)y * w + x] ^ cr[(size_t)y * w + x2]);
                cost[((size_t)y * w + x) * D + d] = c;
            }
}

/* ------------------------------------------------------------------------
 * BT: gradient precompute sgm_kernels.cu:94-118; per-channel BT
 * :1002-1028 (identical to the tiled kernel_BTCost :139-239 because that
 * kernel clamps its smem halo loads to [0, W-1]); pixel cost :1030-1043.
 * ---------------------------------------------------------------------- */
void orc_bt_grad(const uint8_t* img, size_t step, int w, int h, int cap, uint8_t* out)
{
    const int ftzero = imax(cap, 15) | 1;
    for (int y = 0; y < h; ++y) {
        const int yp = imax(y - 1, 0), yn = imin(y + 1, h - 1);
        for (int x = 0; x < w; ++x) {
            const int xp = imax(x - 1, 0), xn = imin(x + 1, w - 1);
            const uint8_t* r = img + (size_t)y * step;
            const uint8_t* u = img + (size_t)yp * step;
            const uint8_t* dn = img + (size_t)yn * step;
            int g = ((int)r[xn] - (int)r[xp]) * 2;
            g += (int)u[xn] - (int)u[xp];
            g += (int)dn[xn] - (int)dn[xp];
            g = iclamp(g, -ftzero, ftzero);
            out[(size_t)y * w + x] = (uint8_t)(g + ftzero);
        }
    }
}

static int bt_channel(const uint8_t* l, const uint8_t* r, int w, int x1, int x2)
{
    if (x2 < 0 || x2 >= w) return 255;
    const int u = l[x1];
    const int ul = (u + l[imax(x1 - 1, 0)]) / 2;
    const int ur = (u + l[imin(x1 + 1, w - 1)]) / 2;
    const int u0 = imin(imin(ul, ur), u), u1 = imax(imax(ul, ur), u);
    const int v = r[x2];
    const int vl = (v + r[imax(x2 - 1, 0)]) / 2;
    const int vr = (v + r[imin(x2 + 1, w - 1)]) / 2;
    const int v0 = imin(imin(vl, vr), v), v1 = imax(imax(vl, vr), v);
    const int c0 = imax(imax(0, u - v1), v0 - u);
    const int c1 = imax(imax(0, v - u1), u0 - v);
    return imin(c0, c1);
}

/* grad rows are dense (stride w); image rows pitched. */
static int bt_pixel(const uint8_t* g1, const uint8_t* g2, const uint8_t* i1, size_t s1,
                    const uint8_t* i2, size_t s2, int w, int y, int x, int real)
{
    const int x2 = x - real;
    if (x2 < 0 || x2 >= w) return 255;
    const int cg = bt_channel(g1 + (size_t)y * w, g2 + (size_t)y * w, w, x, x2);
    const int ci = bt_channel(i1 + (size_t)y * s1, i2 + (size_t)y * s2, w, x, x2);
    return imin(255, cg + (ci >> 2));
}

void orc_bt_cost(const uint8_t* g1, const uint8_t* g2,
                 const uint8_t* i1, size_t s1, const uint8_t* i2, size_t s2,
                 int w, int h, int D, int minD, int16_t* cost)
{
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int d = 0; d < D; ++d)
                cost[((size_t)y * w + x) * D + d] =
                    (int16_t)bt_pixel(g1, g2, i1, s1, i2, s2, w, y, x, d + minD);
}

/* Box sums with clamped borders, saturated at SHRT_MAX after each pass:
 * sgm_kernels.cu:354-379 (H) and :390-416 (V).  Written as direct clamped
 * window sums (the reference's sliding form is equal to it, :348-351). */
void orc_box_h(const int16_t* src, int16_t* dst, int w, int h, int D, int block)
{
    const int half = block / 2;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int d = 0; d < D; ++d) {
                int s = 0;
                for (int k = -half; k <= half; ++k)
                    s += src[((size_t)y * w + iclamp(x + k, 0, w - 1)) * D + d];
                dst[((size_t)y * w + x) * D + d] = (int16_t)imin(SHRT_MAX, s);
            }
}

void orc_box_v(const int16_t* src, int16_t* dst, int w, int h, int D, int block)
{
    const int half = block / 2;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int d = 0; d < D; ++d) {
                int s = 0;
                for (int k = -half; k <= half; ++k)
                    s += src[((size_t)iclamp(y + k, 0, h - 1) * w + x) * D + d];
                dst[((size_t)y * w + x) * D + d] = (int16_t)imin(SHRT_MAX, s);
            }
}

/* ------------------------------------------------------------------------
 * Adaptive P2 (sgm_kernels.cu:63-74) tabulated over |dI| in 0..255 after the
 * host sanitisation of :1671-1675 / :1891-1895.
 * ---------------------------------------------------------------------- */
int orc_p2_lut(int P2, int adaptive, float alpha, float beta, float lower, float upper,
               int32_t lut[256])
{
    if (alpha < 0.0f) alpha = 0.0f;
    beta = fmaxf(beta, 0.5f);
    upper = fmaxf(upper, 1.0f);
    lower = fminf(fmaxf(lower, 0.0f), upper);
    if (alpha <= 0.0f && lower >= 1.0f) adaptive = 0;
    for (int i = 0; i < 256; ++i) {
        if (!adaptive) { lut[i] = P2; continue; }
        const float diff = (float)i;
        const float e = expf(-diff / beta);
        const float ratio = lower + (1.0f + alpha - lower) * e;
        float a = (float)P2 * ratio;
        const float lo = fmaxf(1.0f, (float)P2 * lower);
        const float hi = (float)P2 * upper;
        a = fminf(hi, fmaxf(lo, a));
        lut[i] = (int32_t)(a + 0.5f);
    }
    return adaptive;
}

static inline int p2_at(int P2, int adaptive, const int32_t* lut, int cur, int prev)
{
    return adaptive ? lut[abs(cur - prev)] : P2;
}

/* ------------------------------------------------------------------------
 * Single-scale recurrences.
 * L(p,d) = sat16(C + min(L[d], L[d-1]+P1, L[d+1]+P1, min(MAX, minPrev+P2')) - minPrev)
 * with L[-1] = L[D] = MAX (sgm_kernels.cu:560-565).
 * ---------------------------------------------------------------------- */
static void ss_step(const int16_t* Lp, int16_t* Lc, const int16_t* c, int D, int P1, int p2,
                    int minPrev)
{
    const int lj = imin(SHRT_MAX, minPrev + p2);
    for (int d = 0; d < D; ++d) {
        const int lm = d > 0 ? (int)Lp[d - 1] + P1 : SHRT_MAX;
        const int ls = Lp[d];
        const int lp = d < D - 1 ? (int)Lp[d + 1] + P1 : SHRT_MAX;
        const int best = imin(imin(lm, lp), imin(ls, lj));
        Lc[d] = sat16((int)c[d] + best - minPrev);
    }
}

static int16_t vmin16(const int16_t* v, int D)
{
    int m = v[0];
    for (int d = 1; d < D; ++d) m = imin(m, v[d]);
    return (int16_t)m;
}

/* sgm_kernels.cu:520-578: rows, x in [startX, endX) only; writes acc. */
void orc_path_lr(int16_t* acc, const int16_t* C, const uint8_t* img, size_t istep,
                 int w, int h, int D, int startX, int endX, int P1, int P2,
                 int adaptive, const int32_t* lut)
{
    int16_t* Lp = malloc(sizeof(int16_t) * D);
    int16_t* Lc = malloc(sizeof(int16_t) * D);
    if (startX < 0) startX = 0;
    if (endX > w) endX = w;
    for (int y = 0; y < h && startX < endX; ++y) {
        const size_t row = (size_t)y * w;
        memcpy(Lp, C + (row + startX) * D, sizeof(int16_t) * D);
        int minPrev = vmin16(Lp, D);
        memcpy(acc + (row + startX) * D, Lp, sizeof(int16_t) * D);
        const uint8_t* ir = img + (size_t)y * istep;
        for (int x = startX + 1; x < endX; ++x) {
            ss_step(Lp, Lc, C + (row + x) * D, D, P1, p2_at(P2, adaptive, lut, ir[x], ir[x - 1]), minPrev);
            minPrev = vmin16(Lc, D);
            memcpy(acc + (row + x) * D, Lc, sizeof(int16_t) * D);
            int16_t* t = Lp; Lp = Lc; Lc = t;
        }
    }
    free(Lp); free(Lc);
}

/* sgm_kernels.cu:587-638: columns, all y; acc = (short)min(MAX, acc + L). */
void orc_path_tb(int16_t* acc, const int16_t* C, const uint8_t* img, size_t istep,
                 int w, int h, int D, int P1, int P2, int adaptive, const int32_t* lut)
{
    int16_t* Lp = malloc(sizeof(int16_t) * D);
    int16_t* Lc = malloc(sizeof(int16_t) * D);
    for (int x = 0; x < w; ++x) {
        memcpy(Lp, C + (size_t)x * D, sizeof(int16_t) * D);
        int minPrev = vmin16(Lp, D);
        for (int d = 0; d < D; ++d) {
            int16_t* a = acc + (size_t)x * D + d;
            *a = (int16_t)imin(SHRT_MAX, (int)*a + Lp[d]);
        }
        for (int y = 1; y < h; ++y) {
            const size_t p = (size_t)y * w + x;
            ss_step(Lp, Lc, C + p * D, D, P1,
                    p2_at(P2, adaptive, lut, img[(size_t)y * istep + x], img[(size_t)(y - 1) * istep + x]),
                    minPrev);
            minPrev = vmin16(Lc, D);
            for (int d = 0; d < D; ++d) {
                int16_t* a = acc + p * D + d;
                *a = (int16_t)imin(SHRT_MAX, (int)*a + Lc[d]);
            }
            int16_t* t = Lp; Lp = Lc; Lc = t;
        }
    }
    free(Lp); free(Lc);
}

/* sgm_kernels.cu:665-819: R->L recurrence fused with WTA, uniqueness,
 * sub-pixel, confidence and the sequential disp2 projection. */
void orc_path_rl_wta(int16_t* acc, const int16_t* C, const uint8_t* img, size_t istep,
                     int16_t* disp, int16_t* disp2, int16_t* disp2_cost, float* conf,
                     int w, int h, int D, int minD, int P1, int P2, int uniq,
                     int accumulate, int adaptive, const int32_t* lut)
{
    int16_t* Lp = malloc(sizeof(int16_t) * D);
    int16_t* Lc = malloc(sizeof(int16_t) * D);
    int16_t* my = malloc(sizeof(int16_t) * D);
    const int16_t invalid = (int16_t)((minD - 1) * 16);
    const int minX1 = imax(minD + D, 0);
    const int maxX1 = w + imin(minD, 0);
    for (int y = 0; y < h; ++y) {
        const size_t row = (size_t)y * w;
        const int xs = w - 1;
        memcpy(Lp, C + (row + xs) * D, sizeof(int16_t) * D);
        int minPrev = vmin16(Lp, D);
        for (int d = 0; d < D; ++d) {
            int16_t* a = acc + (row + xs) * D + d;
            *a = (int16_t)imin(SHRT_MAX, (accumulate ? (int)*a : 0) + Lp[d]);
        }
        disp[row + xs] = invalid;
        conf[row + xs] = 0.0f;
        const uint8_t* ir = img + (size_t)y * istep;
        for (int x = w - 2; x >= 0; --x) {
            ss_step(Lp, Lc, C + (row + x) * D, D, P1, p2_at(P2, adaptive, lut, ir[x], ir[x + 1]), minPrev);
            minPrev = vmin16(Lc, D);
            uint32_t bestKey = 0xFFFFFFFFu;
            for (int d = 0; d < D; ++d) {
                int16_t* a = acc + (row + x) * D + d;
                my[d] = (int16_t)imin(SHRT_MAX, (accumulate ? (int)*a : 0) + Lc[d]);
                *a = my[d];
                const uint32_t key = ((uint32_t)(uint16_t)my[d] << 16) | (uint32_t)d;
                if (key < bestKey) bestKey = key;
            }
            { int16_t* t = Lp; Lp = Lc; Lc = t; }
            const int16_t minCost = (int16_t)(bestKey >> 16);
            const int16_t bestD = (int16_t)(bestKey & 0xFFFFu);
            int alt = SHRT_MAX;
            for (int d = 0; d < D; ++d)
                if (abs(d - (int)bestD) > 1) alt = imin(alt, my[d]);
            const int16_t altCost = (int16_t)alt;
            const size_t p = row + x;
            if (x < minX1 || x >= maxX1) {
                disp[p] = invalid;
                conf[p] = 0.0f;
                continue;
            }
            int unique = 1;
            if (uniq > 0 && minCost < SHRT_MAX / 2 &&
                (int)altCost * (100 - uniq) < (int)minCost * 100) unique = 0;
            if (unique && minCost < SHRT_MAX / 2) {
                int dint;
                if (bestD > 0 && bestD < D - 1) {
                    const int c0 = my[bestD - 1], c1 = my[bestD], c2 = my[bestD + 1];
                    const int denom2 = imax(c0 + c2 - 2 * c1, 1);
                    dint = bestD * 16 + ((c0 - c2) * 16 + denom2) / (denom2 * 2);
                } else {
                    dint = bestD * 16;
                }
                const int16_t dval = (int16_t)(dint + minD * 16);
                const int x2 = x - bestD - minD;
                if (x2 >= 0 && x2 < w && disp2_cost[row + x2] > minCost) {
                    disp2_cost[row + x2] = minCost;
                    disp2[row + x2] = (int16_t)(bestD + minD);
                }
                disp[p] = dval;
                const float denom = fmaxf((float)minCost, 1.0f);
                const float cf = ((float)altCost - (float)minCost) / denom;
                conf[p] = fminf(fmaxf(cf, 0.0f), 1.0f);
            } else {
                disp[p] = invalid;
                conf[p] = 0.0f;
            }
        }
    }
    free(Lp); free(Lc); free(my);
}

/* ------------------------------------------------------------------------
 * Post-processing: sgm_kernels.cu:825-851 (pseudo LR), :855-889 (strict LR),
 * :891-920 (median), :922-931 (confidence clear).
 * ---------------------------------------------------------------------- */
void orc_pseudo_lr_check(int16_t* disp, const int16_t* disp2, float* conf,
                         int w, int h, int minD, int maxDiff)
{
    const int16_t invalid = (int16_t)((minD - 1) * 16);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            const int d1 = disp[i];
            if (d1 == invalid) continue;
            const int d0 = d1 >> 4, du = (d1 + 15) >> 4;
            const int xa = x - d0, xb = x - du;
            const int16_t* r = disp2 + (size_t)y * w;
            if (0 <= xa && xa < w && r[xa] >= minD && abs(r[xa] - d0) > maxDiff &&
                0 <= xb && xb < w && r[xb] >= minD && abs(r[xb] - du) > maxDiff) {
                disp[i] = invalid;
                conf[i] = 0.0f;
            }
        }
}

void orc_strict_lr_check(int16_t* disp, const int16_t* disp2, float* conf,
                         int w, int h, int minD, int maxDiff)
{
    const int16_t invalid = (int16_t)((minD - 1) * 16);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            const int d16 = disp[i];
            if (d16 == invalid) continue;
            const int d0 = d16 >> 4, d1 = (d16 + 15) >> 4;
            const int16_t* r = disp2 + (size_t)y * w;
            int ok = 0;
            if (x - d0 >= 0 && x - d0 < w && r[x - d0] >= minD && abs((int)r[x - d0] - d0) <= maxDiff) ok = 1;
            if (x - d1 >= 0 && x - d1 < w && r[x - d1] >= minD && abs((int)r[x - d1] - d1) <= maxDiff) ok = 1;
            if (!ok) { disp[i] = invalid; conf[i] = 0.0f; }
        }
}

void orc_median3x3(const int16_t* src, int16_t* dst, int w, int h)
{
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int16_t v[9];
            int n = 0;
            for (int dy = -1; dy <= 1; ++dy) {
                const int yy = y + dy;
                if (yy < 0 || yy >= h) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    const int xx = x + dx;
                    if (xx < 0 || xx >= w) continue;
                    v[n++] = src[(size_t)yy * w + xx];
                }
            }
            for (int i = 1; i < n; ++i) {
                const int16_t key = v[i];
                int j = i - 1;
                while (j >= 0 && v[j] > key) { v[j + 1] = v[j]; --j; }
                v[j + 1] = key;
            }
            dst[(size_t)y * w + x] = v[n / 2];
        }
}

/* ------------------------------------------------------------------------
 * Pyramid label helpers: sgm_kernels.cu:943-953 (coarse), :955-1000
 * (refine), :1045-1070 (label geometry, valid label count), :1135-1143.
 * ---------------------------------------------------------------------- */
void orc_prepare_refine_labels(const int16_t* pdisp, const float* pconf, int pw, int ph,
                               size_t pdstep, size_t pcstep, int16_t* base, uint8_t* valid,
                               int w, int h, int minD, int maxD, int localD, float minConf)
{
    const int pdp = (int)(pdstep / sizeof(int16_t));
    const int pcp = (int)(pcstep / sizeof(float));
    const int16_t pinv = (int16_t)((minD - 1) * 16);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int px = imin(x >> 1, pw - 1), py = imin(y >> 1, ph - 1);
            const int16_t raw = pdisp[(size_t)py * pdp + px];
            const float cf = pconf[(size_t)py * pcp + px];
            const size_t o = (size_t)y * w + x;
            if (raw == pinv || cf < minConf) { base[o] = (int16_t)minD; valid[o] = 0; continue; }
            const int p16 = (int)raw * 2;
            const int center = (p16 + 8) >> 4;
            const int res16 = p16 - center * 16;
            int b = center + (res16 < 0 ? -(localD / 2) : -(localD / 2) + 1);
            const int physMax = imin(maxD, x);
            const int largest = physMax - localD + 1;
            if (largest >= minD) b = imin(imax(b, minD), largest);
            else b = minD;
            base[o] = (int16_t)b;
            valid[o] = 1;
        }
}

static int label_geom_valid(const uint8_t* valid, const int16_t* base, size_t p, int x, int d,
                            int D, int minD, int maxD, int w)
{
    if (!valid[p] || d < 0 || d >= D) return 0;
    const int real = (int)base[p] + d;
    const int x2 = x - real;
    return real >= minD && real <= maxD && x2 >= 0 && x2 < w;
}

static int valid_label_count(const uint8_t* valid, const int16_t* base, size_t p, int x, int D,
                             int minD, int maxD, int w)
{
    if (!valid[p] || D <= 0) return 0;
    const int b = base[p];
    const int firstReal = imax(minD, x - (w - 1));
    const int lastReal = imin(maxD, x);
    const int fl = imax(0, firstReal - b);
    const int ll = imin(D - 1, lastReal - b);
    return ll >= fl ? ll - fl + 1 : 0;
}

/* sgm_kernels.cu:1112-1133 */
void orc_census_cost_local(const uint32_t* cl, const uint32_t* cr, const int16_t* base,
                           const uint8_t* valid, int16_t* cost, int w, int h, int D,
                           int minD, int maxD)
{
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t p = (size_t)y * w + x;
            for (int d = 0; d < D; ++d) {
                const int real = base[p] + d;
                cost[p * D + d] = label_geom_valid(valid, base, p, x, d, D, minD, maxD, w)
                    ? (int16_t)popc32(cl[p] ^ cr[(size_t)y * w + (x - real)])
                    : (int16_t)SHRT_MAX;
            }
        }
}

/* sgm_kernels.cu:1075-1110: direct clamped block^2 window of BT costs at the
 * centre pixel's real disparity, capped at SHRT_MAX-1. */
void orc_bt_window_cost_local(const uint8_t* g1, const uint8_t* g2,
                              const uint8_t* i1, size_t s1, const uint8_t* i2, size_t s2,
                              const int16_t* base, const uint8_t* valid, int16_t* cost,
                              int w, int h, int D, int minD, int maxD, int block)
{
    const int half = block / 2;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t p = (size_t)y * w + x;
            for (int d = 0; d < D; ++d) {
                if (!label_geom_valid(valid, base, p, x, d, D, minD, maxD, w)) {
                    cost[p * D + d] = SHRT_MAX;
                    continue;
                }
                const int real = base[p] + d;
                int sum = 0;
                for (int dy = -half; dy <= half; ++dy) {
                    const int yy = iclamp(y + dy, 0, h - 1);
                    for (int dx = -half; dx <= half; ++dx) {
                        const int xx = iclamp(x + dx, 0, w - 1);
                        sum += bt_pixel(g1, g2, i1, s1, i2, s2, w, yy, xx, real);
                    }
                }
                cost[p * D + d] = (int16_t)imin(sum, SHRT_MAX - 1);
            }
        }
}

/* Pyramid recurrence in real-disparity space, one pixel step
 * (sgm_kernels.cu:1230-1252, identical in the three path kernels). */
static void pyr_step(const int16_t* Lp, int16_t* Lc, const int16_t* C, const uint8_t* valid,
                     const int16_t* base, size_t p, int x, int D, int minD, int maxD, int w,
                     int prevActive, int prevBase, int minPrev, int P1, int p2, uint8_t* lvalid)
{
    const int b = base[p];
    for (int d = 0; d < D; ++d) {
        const int lv = label_geom_valid(valid, base, p, x, d, D, minD, maxD, w) &&
                       C[p * D + d] < SHRT_MAX;
        lvalid[d] = (uint8_t)lv;
        const int c = lv ? (int)C[p * D + d] : SHRT_MAX;
        if (!prevActive) { Lc[d] = (int16_t)c; continue; }
        if (!lv) { Lc[d] = SHRT_MAX; continue; }
        const int same = d + b - prevBase;
        const int l0 = (same >= 0 && same < D) ? (int)Lp[same] : SHRT_MAX;
        const int mi = same - 1, pl = same + 1;
        const int lm = (mi >= 0 && mi < D && Lp[mi] < SHRT_MAX) ? (int)Lp[mi] + P1 : SHRT_MAX;
        const int lp = (pl >= 0 && pl < D && Lp[pl] < SHRT_MAX) ? (int)Lp[pl] + P1 : SHRT_MAX;
        const int lj = imin(SHRT_MAX, minPrev + p2);
        const int best = imin(imin(l0, lm), imin(lp, lj));
        Lc[d] = best < SHRT_MAX ? sat16(c + best - minPrev) : (int16_t)SHRT_MAX;
    }
}

/* sgm_kernels.cu:1191-1261 */
void orc_path_lr_pyr(int16_t* acc, const int16_t* C, const int16_t* base, const uint8_t* valid,
                     const uint8_t* img, size_t istep, int w, int h, int D, int minD, int maxD,
                     int P1, int P2, int adaptive, const int32_t* lut)
{
    int16_t* Lp = malloc(sizeof(int16_t) * D);
    int16_t* Lc = malloc(sizeof(int16_t) * D);
    uint8_t* lv = malloc((size_t)D);
    for (int y = 0; y < h; ++y) {
        int prevActive = 0, prevBase = minD, minPrev = SHRT_MAX;
        const uint8_t* ir = img + (size_t)y * istep;
        for (int x = 0; x < w; ++x) {
            const size_t p = (size_t)y * w + x;
            if (valid_label_count(valid, base, p, x, D, minD, maxD, w) == 0) {
                for (int d = 0; d < D; ++d) acc[p * D + d] = SHRT_MAX;
                prevActive = 0;
                continue;
            }
            const int p2 = (prevActive && adaptive) ? lut[abs((int)ir[x] - (int)ir[x - 1])] : P2;
            pyr_step(Lp, Lc, C, valid, base, p, x, D, minD, maxD, w, prevActive, prevBase,
                     minPrev, P1, p2, lv);
            const int minCur = vmin16(Lc, D);
            memcpy(acc + p * D, Lc, sizeof(int16_t) * D);
            int16_t* t = Lp; Lp = Lc; Lc = t;
            minPrev = minCur;
            prevBase = base[p];
            prevActive = minCur < SHRT_MAX;
        }
    }
    free(Lp); free(Lc); free(lv);
}

/* sgm_kernels.cu:1263-1334 */
void orc_path_tb_pyr(int16_t* acc, const int16_t* C, const int16_t* base, const uint8_t* valid,
                     const uint8_t* img, size_t istep, int w, int h, int D, int minD, int maxD,
                     int P1, int P2, int adaptive, const int32_t* lut)
{
    int16_t* Lp = malloc(sizeof(int16_t) * D);
    int16_t* Lc = malloc(sizeof(int16_t) * D);
    uint8_t* lv = malloc((size_t)D);
    for (int x = 0; x < w; ++x) {
        int prevActive = 0, prevBase = minD, minPrev = SHRT_MAX;
        for (int y = 0; y < h; ++y) {
            const size_t p = (size_t)y * w + x;
            if (valid_label_count(valid, base, p, x, D, minD, maxD, w) == 0) {
                for (int d = 0; d < D; ++d) acc[p * D + d] = SHRT_MAX;
                prevActive = 0;
                continue;
            }
            const int p2 = (prevActive && adaptive)
                ? lut[abs((int)img[(size_t)y * istep + x] - (int)img[(size_t)(y - 1) * istep + x])] : P2;
            pyr_step(Lp, Lc, C, valid, base, p, x, D, minD, maxD, w, prevActive, prevBase,
                     minPrev, P1, p2, lv);
            const int minCur = vmin16(Lc, D);
            for (int d = 0; d < D; ++d) {
                const int old = acc[p * D + d];
                acc[p * D + d] = (old < SHRT_MAX && Lc[d] < SHRT_MAX)
                    ? (int16_t)imin(SHRT_MAX - 1, old + (int)Lc[d]) : (int16_t)SHRT_MAX;
            }
            int16_t* t = Lp; Lp = Lc; Lc = t;
            minPrev = minCur;
            prevBase = base[p];
            prevActive = minCur < SHRT_MAX;
        }
    }
    free(Lp); free(Lc); free(lv);
}

/* sgm_kernels.cu:1336-1478 */
void orc_path_rl_pyr_wta(int16_t* acc, const int16_t* C, const int16_t* base,
                         const uint8_t* valid, const uint8_t* img, size_t istep,
                         int16_t* disp, int16_t* disp2, int16_t* disp2_cost, float* conf,
                         int w, int h, int D, int minD, int maxD, int P1, int P2, int uniq,
                         float minConf, int accumulate, int adaptive, const int32_t* lut)
{
    int16_t* Lp = malloc(sizeof(int16_t) * D);
    int16_t* Lc = malloc(sizeof(int16_t) * D);
    int16_t* my = malloc(sizeof(int16_t) * D);
    uint8_t* lv = malloc((size_t)D);
    const int16_t invalid = (int16_t)((minD - 1) * 16);
    for (int y = 0; y < h; ++y) {
        int prevActive = 0, prevBase = minD, minPrev = SHRT_MAX;
        const uint8_t* ir = img + (size_t)y * istep;
        for (int x = w - 1; x >= 0; --x) {
            const size_t p = (size_t)y * w + x;
            const int vc = valid_label_count(valid, base, p, x, D, minD, maxD, w);
            if (vc == 0) {
                for (int d = 0; d < D; ++d) acc[p * D + d] = SHRT_MAX;
                disp[p] = invalid;
                conf[p] = 0.0f;
                prevActive = 0;
                continue;
            }
            const int p2 = (prevActive && adaptive) ? lut[abs((int)ir[x] - (int)ir[x + 1])] : P2;
            pyr_step(Lp, Lc, C, valid, base, p, x, D, minD, maxD, w, prevActive, prevBase,
                     minPrev, P1, p2, lv);
            const int minCur = vmin16(Lc, D);
            uint32_t bestKey = 0xFFFFFFFFu;
            for (int d = 0; d < D; ++d) {
                const int old = accumulate ? (int)acc[p * D + d] : 0;
                my[d] = (lv[d] && Lc[d] < SHRT_MAX && (!accumulate || old < SHRT_MAX))
                    ? (int16_t)imin(SHRT_MAX - 1, old + (int)Lc[d]) : (int16_t)SHRT_MAX;
                acc[p * D + d] = my[d];
                const uint32_t key = ((uint32_t)(uint16_t)my[d] << 16) | (uint32_t)d;
                if (key < bestKey) bestKey = key;
            }
            const int16_t minCost = (int16_t)(bestKey >> 16);
            const int bestD = (int)(bestKey & 0xFFFFu);
            int alt = SHRT_MAX;
            for (int d = 0; d < D; ++d)
                if (lv[d] && abs(d - bestD) > 1) alt = imin(alt, my[d]);
            const int16_t altCost = (int16_t)alt;

            const int b = base[p];
            int ok = vc >= 3 && minCost < SHRT_MAX && altCost < SHRT_MAX;
            if (ok && uniq > 0 && (int)altCost * (100 - uniq) < (int)minCost * 100) ok = 0;
            const int realBest = b + bestD;
            const int physMax = imin(maxD, x);
            if ((bestD == 0 && b > minD) || (bestD == D - 1 && b + D - 1 < physMax)) ok = 0;
            float cf = 0.0f;
            if (ok) {
                const float denom = fmaxf((float)minCost, 1.0f);
                cf = fminf(fmaxf(((float)altCost - minCost) / denom, 0.0f), 1.0f);
                if (cf < minConf) ok = 0;
            }
            if (ok) {
                int d16 = realBest * 16;
                if (bestD > 0 && bestD + 1 < D) {
                    const int c0 = my[bestD - 1], c1 = my[bestD], c2 = my[bestD + 1];
                    if (c0 < SHRT_MAX && c1 < SHRT_MAX && c2 < SHRT_MAX) {
                        const int denom2 = imax(c0 + c2 - 2 * c1, 1);
                        d16 += ((c0 - c2) * 16 + denom2) / (denom2 * 2);
                    }
                }
                disp[p] = (int16_t)d16;
                conf[p] = cf;
                const int x2 = x - realBest;
                if (x2 >= 0 && x2 < w) {
                    const size_t rp = (size_t)y * w + x2;
                    if (disp2_cost[rp] > minCost) { disp2_cost[rp] = minCost; disp2[rp] = (int16_t)realBest; }
                }
            } else {
                disp[p] = invalid;
                conf[p] = 0.0f;
            }
            int16_t* t = Lp; Lp = Lc; Lc = t;
            minPrev = minCur;
            prevBase = b;
            prevActive = minCur < SHRT_MAX;
        }
    }
    free(Lp); free(Lc); free(my); free(lv);
}

/* ------------------------------------------------------------------------
 * Output stage shared by both entry points: median + confidence clear
 * (sgm_kernels.cu:1827-1836 / :2042-2051), disp16 copy, float conversion
 * (:1530-1551), pyramid conf-only raw copy (:1852-1857).
 * ---------------------------------------------------------------------- */
static void outputs(int16_t* draw, int16_t* dmed, float* craw, int w, int h, int minD,
                    int postFlags, int pyramid, int16_t* disp, float* dispF, float* confF,
                    size_t dstep, size_t dfstep, size_t cfstep)
{
    const int16_t invalid = (int16_t)((minD - 1) * 16);
    const int16_t* fin = draw;
    if (postFlags & ORC_POST_MEDIAN) {
        orc_median3x3(draw, dmed, w, h);
        for (size_t i = 0; i < (size_t)w * h; ++i)
            if (draw[i] != dmed[i]) craw[i] = 0.0f;
        fin = dmed;
    }
    if (disp)
        for (int y = 0; y < h; ++y)
            memcpy((char*)disp + (size_t)y * dstep, fin + (size_t)y * w, sizeof(int16_t) * w);
    if (dispF && confF) {
        for (int y = 0; y < h; ++y) {
            float* df = (float*)((char*)dispF + (size_t)y * dfstep);
            float* cf = (float*)((char*)confF + (size_t)y * cfstep);
            for (int x = 0; x < w; ++x) {
                const int16_t raw = fin[(size_t)y * w + x];
                const float d = raw / 16.0f;
                const int ok = raw != invalid && d > 0.0f;
                df[x] = ok ? d : 0.0f;
                cf[x] = ok ? craw[(size_t)y * w + x] : 0.0f;
            }
        }
    } else if (confF && pyramid) {
        for (int y = 0; y < h; ++y)
            memcpy((char*)confF + (size_t)y * cfstep, craw + (size_t)y * w, sizeof(float) * w);
    }
}

typedef struct {
    uint8_t *g1, *g2;
    uint32_t *c1, *c2;
    int16_t *C, *acc, *draw, *dmed, *d2, *d2c, *base;
    float* craw;
    uint8_t* valid;
} ws_t;

static int ws_alloc(ws_t* s, int w, int h, int D)
{
    const size_t P = (size_t)w * h, N = P * (size_t)D;
    memset(s, 0, sizeof(*s));
    s->g1 = calloc(P, 1); s->g2 = calloc(P, 1);
    s->c1 = calloc(P, 4); s->c2 = calloc(P, 4);
    s->C = calloc(N, 2); s->acc = calloc(N, 2);
    s->draw = calloc(P, 2); s->dmed = calloc(P, 2); s->d2 = calloc(P, 2); s->d2c = calloc(P, 2);
    s->base = calloc(P, 2); s->craw = calloc(P, 4); s->valid = calloc(P, 1);
    return (s->g1 && s->g2 && s->c1 && s->c2 && s->C && s->acc && s->draw && s->dmed &&
            s->d2 && s->d2c && s->base && s->craw && s->valid) ? 0 : -1;
}

static void ws_free(ws_t* s)
{
    free(s->g1); free(s->g2); free(s->c1); free(s->c2); free(s->C); free(s->acc);
    free(s->draw); free(s->dmed); free(s->d2); free(s->d2c); free(s->base); free(s->craw);
    free(s->valid);
}

static void init_maps(ws_t* s, size_t P, int minD)
{
    const int16_t invalid = (int16_t)((minD - 1) * 16);
    for (size_t i = 0; i < P; ++i) {
        s->draw[i] = invalid; s->d2[i] = invalid; s->d2c[i] = SHRT_MAX; s->craw[i] = 0.0f;
    }
}

/* sgm_kernels.cu:1864-2076 (pyrsgm_compute_single_scale).  Argument checks
 * follow the build's C-ABI (it rejects what the reference leaves undefined:
 * D not a multiple of 16, D > 1024, block < 1, empty images). */
int orc_compute_single_scale(
    const uint8_t* left, const uint8_t* right, int16_t* disp, float* dispF, float* confF,
    size_t lstep, size_t rstep, size_t dstep, size_t dfstep, size_t cfstep,
    int w, int h, int D, int minD, int block, int P1, int P2, int uniq, int disp12,
    int cap, int costMode, int postFlags, int numPaths, int adaptive,
    float alpha, float beta, float lower, float upper)
{
    (void)disp12;
    if (!left || !right || w <= 0 || h <= 0 || D <= 0 || D > 1024 || (D % 16) != 0 || block < 1)
        return -1;
    int32_t lut[256];
    adaptive = orc_p2_lut(P2, adaptive, alpha, beta, lower, upper, lut);
    const int minX1 = imax(minD + D, 0);
    const int maxX1 = w + imin(minD, 0);
    if (numPaths < 1) numPaths = 1;
    if (numPaths > 3) numPaths = 3;
    ws_t s;
    if (ws_alloc(&s, w, h, D) != 0) { ws_free(&s); return -1; }
    const size_t P = (size_t)w * h;
    init_maps(&s, P, minD);
    if (costMode == ORC_COST_CENSUS) {
        orc_census5x5(left, lstep, w, h, s.c1);
        orc_census5x5(right, rstep, w, h, s.c2);
        orc_census_cost(s.c1, s.c2, w, h, D, minD, s.C);
    } else {
        orc_bt_grad(left, lstep, w, h, cap, s.g1);
        orc_bt_grad(right, rstep, w, h, cap, s.g2);
        orc_bt_cost(s.g1, s.g2, left, lstep, right, rstep, w, h, D, minD, s.C);
        orc_box_h(s.C, s.acc, w, h, D, block);
        orc_box_v(s.acc, s.C, w, h, D, block);
    }
    /* cost_acc outside [minX1, maxX1) is never read by a valid output
     * (sgm_kernels.cu:761-763); start from zeros. */
    memset(s.acc, 0, P * (size_t)D * 2);
    if (numPaths >= 2) orc_path_lr(s.acc, s.C, left, lstep, w, h, D, minX1, maxX1, P1, P2, adaptive, lut);
    if (numPaths >= 3) orc_path_tb(s.acc, s.C, left, lstep, w, h, D, P1, P2, adaptive, lut);
    orc_path_rl_wta(s.acc, s.C, left, lstep, s.draw, s.d2, s.d2c, s.craw, w, h, D, minD, P1, P2,
                    uniq, numPaths >= 2, adaptive, lut);
    if (postFlags & ORC_POST_LRCHECK) orc_pseudo_lr_check(s.draw, s.d2, s.craw, w, h, minD, disp12);
    outputs(s.draw, s.dmed, s.craw, w, h, minD, postFlags, 0, disp, dispF, confF, dstep, dfstep, cfstep);
    ws_free(&s);
    return 0;
}

/* sgm_kernels.cu:1617-1862 (pyrsgm_compute_level). */
int orc_compute_level(
    const uint8_t* left, const uint8_t* right, const int16_t* pdisp, const float* pconf,
    int16_t* disp, float* dispF, float* confF,
    size_t lstep, size_t rstep, size_t pdstep, size_t pcstep,
    size_t dstep, size_t dfstep, size_t cfstep,
    int w, int h, int pw, int ph, int D, int minD, int maxD, int block, int P1, int P2,
    int uniq, int disp12, int cap, int costMode, int postFlags, int numPaths,
    int refine, float minConf, int adaptive, float alpha, float beta, float lower,
    float upper)
{
    if (!left || !right || w <= 0 || h <= 0 || D <= 0 || D > 1024 || (D % 16) != 0 ||
        (refine && D > 256) || block < 1)
        return -1;
    if (refine && (!pdisp || !pconf || pw <= 0 || ph <= 0)) return -1;
    maxD = imin(maxD, w - 1);
    maxD = imin(maxD, SHRT_MAX / 16);
    if (!refine) maxD = imin(maxD, minD + D - 1);
    if (maxD < minD) return -1;
    int32_t lut[256];
    adaptive = orc_p2_lut(P2, adaptive, alpha, beta, lower, upper, lut);
    minConf = fminf(fmaxf(minConf, 0.0f), 1.0f);
    if (numPaths < 1) numPaths = 1;
    if (numPaths > 3) numPaths = 3;
    ws_t s;
    if (ws_alloc(&s, w, h, D) != 0) { ws_free(&s); return -1; }
    const size_t P = (size_t)w * h;
    init_maps(&s, P, minD);
    if (refine) {
        orc_prepare_refine_labels(pdisp, pconf, pw, ph, pdstep, pcstep, s.base, s.valid, w, h,
                                  minD, maxD, D, minConf);
    } else {
        for (size_t i = 0; i < P; ++i) { s.base[i] = (int16_t)minD; s.valid[i] = 1; }
    }
    if (costMode == ORC_COST_CENSUS) {
        orc_census5x5(left, lstep, w, h, s.c1);
        orc_census5x5(right, rstep, w, h, s.c2);
        if (refine) orc_census_cost_local(s.c1, s.c2, s.base, s.valid, s.C, w, h, D, minD, maxD);
        else orc_census_cost(s.c1, s.c2, w, h, D, minD, s.C);
    } else {
        orc_bt_grad(left, lstep, w, h, cap, s.g1);
        orc_bt_grad(right, rstep, w, h, cap, s.g2);
        if (refine) {
            orc_bt_window_cost_local(s.g1, s.g2, left, lstep, right, rstep, s.base, s.valid, s.C,
                                     w, h, D, minD, maxD, block);
        } else {
            orc_bt_cost(s.g1, s.g2, left, lstep, right, rstep, w, h, D, minD, s.C);
            orc_box_h(s.C, s.acc, w, h, D, block);
            orc_box_v(s.acc, s.C, w, h, D, block);
        }
    }
    if (numPaths >= 2)
        orc_path_lr_pyr(s.acc, s.C, s.base, s.valid, left, lstep, w, h, D, minD, maxD, P1, P2, adaptive, lut);
    if (numPaths >= 3)
        orc_path_tb_pyr(s.acc, s.C, s.base, s.valid, left, lstep, w, h, D, minD, maxD, P1, P2, adaptive, lut);
    orc_path_rl_pyr_wta(s.acc, s.C, s.base, s.valid, left, lstep, s.draw, s.d2, s.d2c, s.craw,
                        w, h, D, minD, maxD, P1, P2, uniq, minConf, numPaths >= 2, adaptive, lut);
    if (postFlags & ORC_POST_LRCHECK) orc_strict_lr_check(s.draw, s.d2, s.craw, w, h, minD, disp12);
    outputs(s.draw, s.dmed, s.craw, w, h, minD, postFlags, 1, disp, dispF, confF, dstep, dfstep, cfstep);
    ws_free(&s);
    return 0;
}

/* ------------------------------------------------------------------------
 * StereoMatcher: level geometry stereo_matcher.cpp:169-193, post flags
 * :229-236, compute :238-338 (dense buffers here).
 * ---------------------------------------------------------------------- */
static int align16(int v) { return imax(16, ((v + 15) / 16) * 16); }

int orc_level_geometry(int w, int h, const orc_params* p, orc_level_info info[5])
{
    const int levels = p->pyramid_levels;
    int lw = w, lh = h;
    for (int l = 0; l <= levels; ++l) {
        if (l > 0) { lw = (lw + 1) / 2; lh = (lh + 1) / 2; }
        orc_level_info* li = &info[l];
        li->level = l; li->width = lw; li->height = lh;
        if (levels == 0) {
            li->refine = 0;
            li->num_labels = p->num_disparities;
            li->max_disparity = p->min_disparity + p->num_disparities - 1;
            li->block_size = p->block_size;
        } else {
            li->refine = l != levels;
            li->max_disparity = imin(p->pyramid_max_disparity >> l, lw - 1);
            li->num_labels = li->refine ? 16
                : align16(imin(p->pyramid_coarse_disparities, li->max_disparity + 1));
            li->block_size = l >= 2 ? 3 : p->block_size;
        }
    }
    return levels + 1;
}

int orc_matcher_compute(int w, int h, const orc_params* p, const uint8_t* left,
                        const uint8_t* right, int16_t* disp16, float* disp32, float* conf)
{
    if (!left || !right) return -1;
    if (!disp16 && !disp32 && !conf) return 0;
    orc_level_info info[5];
    orc_level_geometry(w, h, p, info);
    const int levels = p->pyramid_levels;
    float* scratch = NULL;
    float* oconf = conf;
    if (!oconf && disp32) { scratch = calloc((size_t)w * h, 4); oconf = scratch; }
    int ret = 0;
    if (levels == 0) {
        int flags = (p->lr_check ? ORC_POST_LRCHECK : 0) | (p->median_filter ? ORC_POST_MEDIAN : 0);
        ret = orc_compute_single_scale(left, right, disp16, disp32, oconf, w, w, 2 * (size_t)w,
                                       4 * (size_t)w, 4 * (size_t)w, w, h, p->num_disparities,
                                       p->min_disparity, p->block_size, p->p1, p->p2,
                                       p->uniqueness_ratio, p->disp12_max_diff, p->pre_filter_cap,
                                       p->cost_mode, flags, p->num_paths, p->adaptive_p2,
                                       p->p2_alpha, p->p2_beta, p->p2_lower_mult, p->p2_upper_mult);
        free(scratch);
        return ret;
    }
    uint8_t* pl[5] = {0};
    uint8_t* pr[5] = {0};
    int16_t* pd[5] = {0};
    float* pc[5] = {0};
    for (int l = 1; l <= levels; ++l) {
        const size_t P = (size_t)info[l].width * info[l].height;
        pl[l] = malloc(P); pr[l] = malloc(P); pd[l] = malloc(P * 2); pc[l] = malloc(P * 4);
        const uint8_t* sl = l == 1 ? left : pl[l - 1];
        const uint8_t* sr = l == 1 ? right : pr[l - 1];
        orc_pyrdown_u8(sl, info[l - 1].width, info[l - 1].width, info[l - 1].height,
                       pl[l], info[l].width, info[l].width, info[l].height);
        orc_pyrdown_u8(sr, info[l - 1].width, info[l - 1].width, info[l - 1].height,
                       pr[l], info[l].width, info[l].width, info[l].height);
    }
    for (int l = levels; l >= 0 && ret == 0; --l) {
        const orc_level_info* li = &info[l];
        const uint8_t* L = l == 0 ? left : pl[l];
        const uint8_t* R = l == 0 ? right : pr[l];
        const int flags = l != 0 ? ORC_POST_LRCHECK
            : ((p->lr_check ? ORC_POST_LRCHECK : 0) | (p->median_filter ? ORC_POST_MEDIAN : 0));
        int16_t* o16 = l == 0 ? disp16 : pd[l];
        float* o32 = l == 0 ? disp32 : NULL;
        float* oc = l == 0 ? oconf : pc[l];
        const int pw = li->refine ? info[l + 1].width : 0;
        const int ph = li->refine ? info[l + 1].height : 0;
        ret = orc_compute_level(L, R, li->refine ? pd[l + 1] : NULL, li->refine ? pc[l + 1] : NULL,
                                o16, o32, oc, li->width, li->width, 2 * (size_t)pw, 4 * (size_t)pw,
                                2 * (size_t)li->width, 4 * (size_t)li->width, 4 * (size_t)li->width,
                                li->width, li->height, pw, ph, li->num_labels, p->min_disparity,
                                li->max_disparity, li->block_size, p->p1, p->p2,
                                p->uniqueness_ratio, p->disp12_max_diff, p->pre_filter_cap,
                                p->cost_mode, flags, p->num_paths, li->refine,
                                p->pyramid_min_confidence, p->adaptive_p2, p->p2_alpha,
                                p->p2_beta, p->p2_lower_mult, p->p2_upper_mult);
    }
    for (int l = 1; l <= levels; ++l) { free(pl[l]); free(pr[l]); free(pd[l]); free(pc[l]); }
    free(scratch);
    return ret;
}
