/*
 * oracle_kernels.c — TEST INFRASTRUCTURE ONLY (CPU oracle, see oracle.h).
 *
 * Scalar restatement of the reference's per-block kernels:
 *   dequantisation / transform skip   src/transform.rs:594-910
 *   inverse DCT 4..32 (partial butterfly) src/transform.rs:406-531
 *   inverse DST4 (matrix form)        src/transform.rs:30, :535-590
 *   intra prediction                  src/intra.rs:31-374
 *   luma deblocking segment           src/deblock.rs:216-330
 *   chroma QP table                   src/decode.rs:7599
 *   weighted MC sample ops            src/mc.rs:631-733
 *   inverse ACT                       src/act.rs:43
 */
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include "oracle.h"

static int64_t clamp64(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
static int32_t clamp32(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
static int32_t wrap16(int32_t v) { return (int32_t)(int16_t)(uint16_t)(uint32_t)v; }
static int ilog2(int n) { int l = 0; while ((1 << l) < n) l++; return l; }

/* ------------------------------------------------------------------------ */
/* Dequantisation (src/transform.rs:686-910)                                 */
/* ------------------------------------------------------------------------ */
static const int64_t LEVEL_SCALE[6] = {40, 45, 51, 57, 64, 72}; /* src/transform.rs:37 */

int oracle_max_range(int bit_depth, int extended) {
    /* src/transform.rs:686 max_log2_transform_dynamic_range */
    if (extended) return bit_depth + 6 > 15 ? bit_depth + 6 : 15;
    return 15;
}

typedef struct { int64_t factor, add, clip_min, clip_max; int shift; } dq_params;

static dq_params dequant_params(int n, int qp_prime, int bd, int extended) {
    /* src/transform.rs:701 dequant_params */
    dq_params p;
    int64_t log2n = ilog2(n);
    int64_t max_range = oracle_max_range(bd, extended);
    int64_t bd_shift = bd + log2n + 10 - max_range;
    int max_qp = 51 + 6 * (bd - 8);
    int64_t qp = clamp32(qp_prime, 0, max_qp);
    p.factor = LEVEL_SCALE[qp % 6] * ((int64_t)1 << (qp / 6)) * 16;
    p.add = (int64_t)1 << (bd_shift - 1);
    p.shift = (int)bd_shift;
    p.clip_min = -((int64_t)1 << max_range);
    p.clip_max = ((int64_t)1 << max_range) - 1;
    return p;
}

/* ScalingMatrix::coeff (src/transform.rs:668): 4x4 reads the first 16
 * entries directly; larger sizes upsample the 8x8 list, DC override n>=16. */
static int64_t scaling_coeff(const uint8_t* sc, int n, int idx) {
    int log2 = ilog2(n);
    if (log2 == 2) return sc[idx];
    if (idx == 0 && log2 >= 4) return sc[64];
    int y = idx / n, x = idx - y * n;
    int down = log2 > 3 ? log2 - 3 : 0;
    return sc[(y >> down) * 8 + (x >> down)];
}

void oracle_dequant(const int32_t* levels, int n, int qp_prime, int bd, int extended,
                    const uint8_t* scaling, int32_t* out) {
    /* src/transform.rs:764 dequantize_into_scalar / :794 dequantize_scaled_into_scalar;
     * a flat-16 list uses the flat path (src/decode.rs:5930-5936). */
    dq_params p = dequant_params(n, qp_prime, bd, extended);
    int flat = scaling == NULL || scaling[65];
    for (int i = 0; i < n * n; i++) {
        int64_t f = flat ? p.factor : (p.factor / 16) * scaling_coeff(scaling, n, i);
        int64_t l = clamp64(levels[i], p.clip_min, p.clip_max);
        out[i] = (int32_t)clamp64((l * f + p.add) >> p.shift, p.clip_min, p.clip_max);
    }
}

void oracle_dequant_ts(const int32_t* levels, int n, int qp_prime, int bd, int extended,
                       const uint8_t* scaling, int32_t* out) {
    /* src/transform.rs:728 transform_skip_params + :756 apply_transform_skip_shift
     * + :779/:810 */
    int log2n = ilog2(n);
    int max_range = oracle_max_range(bd, extended);
    int tr_shift = max_range - bd - log2n;
    if (extended && tr_shift < 0) tr_shift = 0;
    int tr_add = tr_shift > 0 ? 1 << (tr_shift - 1) : 0;
    int32_t deq[32 * 32];
    oracle_dequant(levels, n, qp_prime, bd, extended, scaling, deq);
    for (int i = 0; i < n * n; i++) {
        if (tr_shift >= 0) out[i] = (int32_t)((uint32_t)deq[i] + (uint32_t)tr_add) >> tr_shift;
        else out[i] = (int32_t)((uint32_t)deq[i] << (-tr_shift));
    }
}

/* ------------------------------------------------------------------------ */
/* Inverse transforms (src/transform.rs:30, :406-590)                       */
/* ------------------------------------------------------------------------ */

/* HEVC transform matrix.  Row k of the 32-point matrix holds
 * c(k*(2m+1) mod 128) where c(j) is the normative integer approximation of
 * 64*sqrt(2)*cos(j*pi/64) (sign folded by quadrant); row 0 is 64.  The
 * N-point matrix is rows 0, 32/N, 2*32/N, ... of the 32-point one restricted
 * to its first N columns (H.265 8.6.4.2). */
static const int COS64[33] = {64 /*unused*/, 90, 90, 90, 89, 88, 87, 85, 83, 82, 80, 78, 75, 73, 70, 67,
                              64, 61, 57, 54, 50, 46, 43, 38, 36, 31, 25, 22, 18, 13, 9, 4, 0};

int oracle_dct_coef(int n, int k, int m) {
    int r = k * (32 / n);
    if (r == 0) return 64;
    int a = (r * (2 * m + 1)) % 128;
    if (a < 32) return COS64[a];
    if (a < 64) return -COS64[64 - a];
    if (a < 96) return -COS64[a - 64];
    return COS64[128 - a];
}

static const int DST4[4][4] = {{29, 55, 74, 84}, {74, 74, 0, -74}, {84, -29, -74, 55}, {55, -84, 74, -29}};

/* 1-D partial butterfly (src/transform.rs:406 idct_raw_4, :415-449
 * idct_raw_8/16/32 with idct_odd_8/16/32): even half by recursion on the
 * even-indexed inputs, odd half from the odd basis rows, then mirrored. */
static void idct_raw(int n, const int32_t* c, int32_t* out) {
    if (n == 4) {
        int32_t e0 = 64 * (c[0] + c[2]);
        int32_t e1 = 64 * (c[0] - c[2]);
        int32_t o0 = 83 * c[1] + 36 * c[3];
        int32_t o1 = 36 * c[1] - 83 * c[3];
        out[0] = e0 + o0; out[1] = e1 + o1; out[2] = e1 - o1; out[3] = e0 - o0;
        return;
    }
    int32_t even_in[16], even[16], odd[16];
    int h = n / 2;
    for (int j = 0; j < h; j++) even_in[j] = c[2 * j];
    idct_raw(h, even_in, even);
    for (int k = 0; k < h; k++) {
        int32_t acc = 0;
        for (int j = 1; j < n; j += 2) acc += oracle_dct_coef(n, j, k) * c[j];
        odd[k] = acc;
    }
    for (int k = 0; k < h; k++) {
        out[k] = even[k] + odd[k];
        out[n - 1 - k] = even[k] - odd[k];
    }
}

void oracle_inv_dct(const int32_t* coeff, int n, int bd, int max_range, int nx, int i16, int32_t* out) {
    /* src/transform.rs:489 inv_dct_n_into */
    int shift2 = max_range + 5 - bd;
    int32_t add2 = 1 << (shift2 - 1);
    int32_t cmin = -(1 << max_range), cmax = (1 << max_range) - 1;
    int32_t tmp[32 * 32];
    memset(tmp, 0, sizeof(tmp));
    if (nx > n) nx = n;
    for (int c = 0; c < nx; c++) {
        int32_t col[32], raw[32];
        for (int k = 0; k < n; k++) col[k] = coeff[k * n + c];
        idct_raw(n, col, raw);
        for (int m = 0; m < n; m++) tmp[m * n + c] = clamp32((raw[m] + 64) >> 7, cmin, cmax);
    }
    for (int r = 0; r < n; r++) {
        int32_t raw[32];
        idct_raw(n, &tmp[r * n], raw);
        for (int m = 0; m < n; m++) {
            int32_t v = (int32_t)((uint32_t)raw[m] + (uint32_t)add2) >> shift2;
            out[r * n + m] = i16 ? wrap16(v) : v;
        }
    }
}

void oracle_inv_dct_dense(const int32_t* coeff, int n, int bd, int max_range, int32_t* out) {
    int shift2 = max_range + 5 - bd;
    int32_t cmin = -(1 << max_range), cmax = (1 << max_range) - 1;
    int32_t tmp[32 * 32];
    for (int c = 0; c < n; c++)
        for (int m = 0; m < n; m++) {
            int64_t acc = 0;
            for (int k = 0; k < n; k++) acc += (int64_t)oracle_dct_coef(n, k, m) * coeff[k * n + c];
            tmp[m * n + c] = clamp32((int32_t)((acc + 64) >> 7), cmin, cmax);
        }
    for (int r = 0; r < n; r++)
        for (int m = 0; m < n; m++) {
            int64_t acc = 0;
            for (int k = 0; k < n; k++) acc += (int64_t)oracle_dct_coef(n, k, m) * tmp[r * n + k];
            out[r * n + m] = (int32_t)((acc + (1 << (shift2 - 1))) >> shift2);
        }
}

void oracle_inv_dst4(const int32_t* coeff, int bd, int max_range, int i16, int32_t* out) {
    /* src/transform.rs:535 inv_transform_n_into::<4> with DST4 */
    int shift2 = max_range + 5 - bd;
    int32_t add2 = 1 << (shift2 - 1);
    int32_t cmin = -(1 << max_range), cmax = (1 << max_range) - 1;
    int32_t tmp[16];
    for (int c = 0; c < 4; c++) {
        int32_t acc[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; k++) {
            int32_t ck = coeff[k * 4 + c];
            if (!ck) continue;
            for (int m = 0; m < 4; m++) acc[m] += DST4[k][m] * ck;
        }
        for (int m = 0; m < 4; m++) tmp[m * 4 + c] = clamp32((acc[m] + 64) >> 7, cmin, cmax);
    }
    for (int r = 0; r < 4; r++) {
        int32_t acc[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; k++) {
            int32_t rk = tmp[r * 4 + k];
            if (!rk) continue;
            for (int m = 0; m < 4; m++) acc[m] += DST4[k][m] * rk;
        }
        for (int m = 0; m < 4; m++) {
            int32_t v = (int32_t)((uint32_t)acc[m] + (uint32_t)add2) >> shift2;
            out[r * 4 + m] = i16 ? wrap16(v) : v;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Intra prediction (src/intra.rs)                                          */
/* ------------------------------------------------------------------------ */
static const int INTRA_PRED_ANGLE[33] = {32, 26, 21, 17, 13, 9, 5, 2, 0, -2, -5, -9, -13, -17, -21, -26, -32,
                                         -26, -21, -17, -13, -9, -5, -2, 0, 2, 5, 9, 13, 17, 21, 26, 32};
static const int INV_ANGLE[15] = {-4096, -1638, -910, -630, -482, -390, -315, -256,
                                  -315, -390, -482, -630, -910, -1638, -4096};

static void substitute_refs(int corner, const int32_t* raw_above, const int32_t* raw_left, int n, int neutral,
                            int32_t* above, int32_t* left) {
    /* src/intra.rs:84 substitute_refs_into */
    int total = 4 * n + 1;
    int32_t s[129];
    int avail[129];
    for (int i = 0; i < total; i++) { s[i] = 0; avail[i] = 0; }
    for (int i = 0; i < 2 * n; i++)
        if (raw_left[2 * n - 1 - i] >= 0) { s[i] = raw_left[2 * n - 1 - i]; avail[i] = 1; }
    if (corner >= 0) { s[2 * n] = corner; avail[2 * n] = 1; }
    for (int i = 0; i < 2 * n; i++)
        if (raw_above[i] >= 0) { s[2 * n + 1 + i] = raw_above[i]; avail[2 * n + 1 + i] = 1; }
    int first = -1;
    for (int i = 0; i < total; i++) if (avail[i]) { first = i; break; }
    if (first < 0) {
        for (int i = 0; i < total; i++) s[i] = neutral;
    } else {
        for (int i = 0; i < first; i++) s[i] = s[first];
        for (int i = 1; i < total; i++) if (!avail[i]) s[i] = s[i - 1];
    }
    above[0] = s[2 * n];
    left[0] = s[2 * n];
    for (int i = 0; i < 2 * n; i++) {
        above[1 + i] = s[2 * n + 1 + i];
        left[1 + i] = s[2 * n - 1 - i];
    }
}

static void filter_refs(const int32_t* above, const int32_t* left, int n, int mode, int is_luma, int enabled,
                        int strong, int bd, int32_t* fa, int32_t* fl) {
    /* src/intra.rs:151 filter_refs_into */
    int len = 2 * n + 1;
    for (int i = 0; i < len; i++) { fa[i] = above[i]; fl[i] = left[i]; }
    if (!enabled || !is_luma || n == 4 || mode == 1) return;
    int d26 = abs(mode - 26), d10 = abs(mode - 10);
    int min_dist = d26 < d10 ? d26 : d10;
    int thresh = n == 8 ? 7 : (n == 16 ? 1 : (n == 32 ? 0 : 8));
    if (!(mode == 0 || min_dist > thresh)) return;
    if (strong && n == 32) {
        int thr = 1 << (bd - 5);
        int32_t bl = left[2 * n], tl = above[0], tr = above[2 * n], cl = left[n], ct = above[n];
        if (abs(bl + tl - 2 * cl) < thr && abs(tr + tl - 2 * ct) < thr) {
            for (int i = 1; i < 2 * n; i++) {
                fa[i] = ((2 * n - i) * tl + i * tr + n) >> 6;
                fl[i] = ((2 * n - i) * tl + i * bl + n) >> 6;
            }
            fa[0] = above[0]; fl[0] = left[0];
            fa[2 * n] = above[2 * n]; fl[2 * n] = left[2 * n];
            return;
        }
    }
    int32_t corner = above[0];
    fa[0] = (left[1] + 2 * corner + above[1] + 2) >> 2;
    fl[0] = fa[0];
    for (int i = 1; i < 2 * n; i++) {
        fa[i] = (above[i - 1] + 2 * above[i] + above[i + 1] + 2) >> 2;
        fl[i] = (left[i - 1] + 2 * left[i] + left[i + 1] + 2) >> 2;
    }
    fa[2 * n] = above[2 * n];
    fl[2 * n] = left[2 * n];
}

static void predict(int mode, const int32_t* above, const int32_t* left, int n, int is_luma, int bd, uint16_t* out) {
    /* src/intra.rs:257 predict_into_scalar */
    int log2n = ilog2(n);
    if (mode == 0) {
        int32_t tr = above[n + 1], bl = left[n + 1];
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) {
                int32_t h = (n - 1 - x) * left[y + 1] + (x + 1) * tr;
                int32_t v = (n - 1 - y) * above[x + 1] + (y + 1) * bl;
                out[y * n + x] = (uint16_t)((h + v + n) >> (log2n + 1));
            }
        return;
    }
    if (mode == 1) {
        int32_t sum = 0;
        for (int i = 1; i <= n; i++) sum += above[i] + left[i];
        int32_t dc = (sum + n) >> (log2n + 1);
        for (int i = 0; i < n * n; i++) out[i] = (uint16_t)dc;
        if (is_luma && n < 32) {
            out[0] = (uint16_t)((left[1] + 2 * dc + above[1] + 2) >> 2);
            for (int x = 1; x < n; x++) out[x] = (uint16_t)((above[x + 1] + 3 * dc + 2) >> 2);
            for (int y = 1; y < n; y++) out[y * n] = (uint16_t)((left[y + 1] + 3 * dc + 2) >> 2);
        }
        return;
    }
    int angle = INTRA_PRED_ANGLE[mode - 2];
    int vertical = mode >= 18;
    const int32_t* mainr = vertical ? above : left;
    const int32_t* side = vertical ? left : above;
    int32_t refs[3 * 32 + 1];
    int base = n;
    for (int i = 0; i <= 2 * n; i++) refs[n + i] = mainr[i];
    if (angle < 0) {
        int inv = INV_ANGLE[mode - 11];
        int lim = (n * angle) >> 5;
        for (int k = -1; k >= lim; k--) {
            int idx = (k * inv + 128) >> 8;
            if (idx > 2 * n) idx = 2 * n;
            if (idx < 0) idx = 0;
            refs[k + base] = side[idx];
        }
    }
    int32_t max_v = (1 << bd) - 1;
    if (vertical) {
        for (int row = 0; row < n; row++) {
            int pos = (row + 1) * angle, i_idx = pos >> 5, frac = pos & 31;
            for (int col = 0; col < n; col++) {
                int32_t r0 = refs[col + i_idx + 1 + base];
                int32_t v = frac == 0 ? r0 : ((32 - frac) * r0 + frac * refs[col + i_idx + 2 + base] + 16) >> 5;
                out[row * n + col] = (uint16_t)clamp32(v, 0, max_v);
            }
        }
    } else {
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) {
                int pos = (x + 1) * angle, i_idx = pos >> 5, frac = pos & 31;
                int32_t r0 = refs[y + i_idx + 1 + base];
                int32_t v = frac == 0 ? r0 : ((32 - frac) * r0 + frac * refs[y + i_idx + 2 + base] + 16) >> 5;
                out[y * n + x] = (uint16_t)clamp32(v, 0, max_v);
            }
    }
    if (is_luma && n < 32) {
        if (mode == 26) {
            for (int y = 0; y < n; y++)
                out[y * n] = (uint16_t)clamp32(above[1] + ((left[y + 1] - above[0]) >> 1), 0, max_v);
        } else if (mode == 10) {
            for (int x = 0; x < n; x++)
                out[x] = (uint16_t)clamp32(left[1] + ((above[x + 1] - above[0]) >> 1), 0, max_v);
        }
    }
}

void oracle_intra_filter_refs(const int32_t* above, const int32_t* left, int n, int mode, int is_luma, int enabled,
                              int strong, int bd, int32_t* fa, int32_t* fl) {
    filter_refs(above, left, n, mode, is_luma, enabled, strong, bd, fa, fl);
}

void oracle_intra_predict(int corner, const int32_t* raw_above, const int32_t* raw_left, int n, int mode,
                          int filter_is_luma, int smoothing_enabled, int strong, int edge_filters, int bd,
                          uint16_t* out) {
    int32_t above[65], left[65], fa[65], fl[65];
    substitute_refs(corner, raw_above, raw_left, n, 1 << (bd - 1), above, left);
    filter_refs(above, left, n, mode, filter_is_luma, smoothing_enabled, strong, bd, fa, fl);
    predict(mode, fa, fl, n, edge_filters, bd, out);
}

/* ------------------------------------------------------------------------ */
/* Deblocking (src/deblock.rs:216-330)                                       */
/* ------------------------------------------------------------------------ */
/* lines[line][tap+4]: tap -4..-1 = p3..p0, 0..3 = q0..q3 */
#define G(line, tap) (lines[(line)][(tap) + 4])

int oracle_luma_decision(int32_t lines[4][8], int beta, int tc, int* do_p1, int* do_q1) {
    /* src/deblock.rs:216 luma_decision */
    int dp0 = abs(G(0, -3) - 2 * G(0, -2) + G(0, -1));
    int dp3 = abs(G(3, -3) - 2 * G(3, -2) + G(3, -1));
    int dq0 = abs(G(0, 2) - 2 * G(0, 1) + G(0, 0));
    int dq3 = abs(G(3, 2) - 2 * G(3, 1) + G(3, 0));
    *do_p1 = *do_q1 = 0;
    if (dp0 + dp3 + dq0 + dq3 >= beta) return 0;
    int ok0 = 2 * (dp0 + dq0) < (beta >> 2) && abs(G(0, -4) - G(0, -1)) + abs(G(0, 0) - G(0, 3)) < (beta >> 3) &&
              abs(G(0, -1) - G(0, 0)) < ((5 * tc + 1) >> 1);
    int ok3 = 2 * (dp3 + dq3) < (beta >> 2) && abs(G(3, -4) - G(3, -1)) + abs(G(3, 0) - G(3, 3)) < (beta >> 3) &&
              abs(G(3, -1) - G(3, 0)) < ((5 * tc + 1) >> 1);
    if (ok0 && ok3) return 1;
    int side_thr = (beta + (beta >> 1)) >> 3;
    *do_p1 = dp0 + dp3 < side_thr;
    *do_q1 = dq0 + dq3 < side_thr;
    return 2;
}

void oracle_deblock_luma_segment(int32_t lines[4][8], int beta, int tc, int max_val) {
    /* src/deblock.rs:241 deblock_luma_segment */
    int do_p1, do_q1;
    int dec = oracle_luma_decision(lines, beta, tc, &do_p1, &do_q1);
    if (dec == 0) return;
    int tc2 = tc >> 1;
    for (int l = 0; l < 4; l++) {
        int32_t p0 = G(l, -1), p1 = G(l, -2), p2 = G(l, -3), p3 = G(l, -4);
        int32_t q0 = G(l, 0), q1 = G(l, 1), q2 = G(l, 2), q3 = G(l, 3);
        if (dec == 1) {
#define SC(orig, v) clamp32(clamp32((v), (orig)-2 * tc, (orig) + 2 * tc), 0, max_val)
            G(l, -1) = SC(p0, (p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3);
            G(l, -2) = SC(p1, (p2 + p1 + p0 + q0 + 2) >> 2);
            G(l, -3) = SC(p2, (2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3);
            G(l, 0) = SC(q0, (p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3);
            G(l, 1) = SC(q1, (p0 + q0 + q1 + q2 + 2) >> 2);
            G(l, 2) = SC(q2, (p0 + q0 + q1 + 3 * q2 + 2 * q3 + 4) >> 3);
#undef SC
        } else {
            int32_t delta0 = (9 * (q0 - p0) - 3 * (q1 - p1) + 8) >> 4;
            if (abs(delta0) < tc * 10) {
                int32_t delta = clamp32(delta0, -tc, tc);
                G(l, -1) = clamp32(p0 + delta, 0, max_val);
                G(l, 0) = clamp32(q0 - delta, 0, max_val);
                if (do_p1) {
                    int32_t dp1 = clamp32((((p2 + p0 + 1) >> 1) - p1 + delta) >> 1, -tc2, tc2);
                    G(l, -2) = clamp32(p1 + dp1, 0, max_val);
                }
                if (do_q1) {
                    int32_t dq1 = clamp32((((q2 + q0 + 1) >> 1) - q1 - delta) >> 1, -tc2, tc2);
                    G(l, 1) = clamp32(q1 + dq1, 0, max_val);
                }
            }
        }
    }
}
#undef G

int oracle_qpc(int qpi, int chroma_idc, int bd) {
    /* src/decode.rs:7599 qpc */
    static const int T[14] = {29, 30, 31, 32, 33, 33, 34, 34, 35, 35, 36, 36, 37, 37};
    qpi = clamp32(qpi, -6 * (bd - 8), 57);
    if (chroma_idc != 1) return qpi < 51 ? qpi : 51;
    if (qpi < 30) return qpi;
    if (qpi > 43) return qpi - 6;
    return T[qpi - 30];
}

/* ------------------------------------------------------------------------ */
/* Weighted prediction sample ops (src/mc.rs:631 / :682) and ACT             */
/* ------------------------------------------------------------------------ */
int oracle_weighted_uni(int s, int w, int o, int log2_denom, int bd) {
    int log2wd = log2_denom + 14 - bd;
    int max = (1 << bd) - 1;
    int v = log2wd >= 1 ? ((s * w + (1 << (log2wd - 1))) >> log2wd) + o : s * w + o;
    return clamp32(v, 0, max);
}

int oracle_weighted_bi(int a, int b, int w0, int w1, int o0, int o1, int log2_denom, int bd) {
    int log2wd = log2_denom + 14 - bd;
    int max = (1 << bd) - 1;
    int64_t rnd = ((int64_t)o0 + o1 + 1) << log2wd;
    int32_t v = (int32_t)(((int64_t)a * w0 + (int64_t)b * w1 + rnd) >> (log2wd + 1));
    return clamp32(v, 0, max);
}

void oracle_act_inverse(int32_t* y, int32_t* cg, int32_t* co) {
    /* src/act.rs:43 inverse_sample */
    /* two's-complement wrap-around like the reference's release build */
    int32_t t = (int32_t)((uint32_t)*y - (uint32_t)(*cg >> 1));
    int32_t g = (int32_t)((uint32_t)*cg + (uint32_t)t);
    int32_t b = (int32_t)((uint32_t)t - (uint32_t)(*co >> 1));
    int32_t r = (int32_t)((uint32_t)b + (uint32_t)*co);
    *y = g; *cg = b; *co = r;
}
