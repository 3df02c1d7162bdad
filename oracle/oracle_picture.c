/*
 * oracle_picture.c — TEST INFRASTRUCTURE ONLY (CPU oracle, see oracle.h).
 *
 * Picture-level restatement of the reference's reconstruction order:
 *   per-TU reconstruction in decode order   src/decode.rs:3478-3611 (transform_unit),
 *       :4641-4870 (availability, gather, reconstruct_luma),
 *       :5007-5504 (do_chroma, reconstruct_chroma, cross-component prediction),
 *       :4356 rext_post_ops, :8418 apply_rext_residual_ops,
 *       src/reconstruct.rs:125,226 (narrow, add_residual_generic)
 *   motion compensation (before the TBs)   src/decode.rs:7000-7485 motion_compensate,
 *       src/mc.rs:310 RefPlane::sample, :351 / :442 luma / chroma interpolation,
 *       :538-733 uni / bi / weighted sample prediction
 *   serial deblocking                      src/decode.rs:1673-2143
 *   serial SAO                             src/decode.rs:2192-2493, src/sao.rs:179-470
 *   YUV -> RGB                             src/yuv.rs:187-668
 *   grid stitch / RGB8 narrowing           src/lib.rs:1610-1721, :1964-1977
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include "oracle.h"

static int32_t clamp32(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
static int64_t clamp64(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
static int32_t wrap16(int32_t v) { return (int32_t)(int16_t)(uint16_t)(uint32_t)v; }
static int32_t wadd32(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

typedef struct {
    const hpvcd_batch* b;
    const hpvcd_picture* P;
    int w, h, cw, ch, sub_w, sub_h, gw, gh;
    int bd, bd_c, ep;
    uint16_t *y, *cb, *cr;
    uint8_t* decoded;        /* per 4x4 luma cell: reconstructed so far */
    const uint32_t* cell;    /* cell info words */
    int32_t ccl[32 * 32];    /* cross_comp_luma (src/decode.rs:279) */
    uint16_t* const* refp;   /* final planes of earlier pictures: [3 * picture + plane], NULL if absent */
} pic_ctx;

static int sub_w_of(int idc) { return (idc == 1 || idc == 2) ? 2 : 1; }
static int sub_h_of(int idc) { return idc == 1 ? 2 : 1; }

/* ------------------------------------------------------------------------ */
/* availability (src/decode.rs:4641-4691)                                   */
/* ------------------------------------------------------------------------ */
static int cell_at(const pic_ctx* c, int x, int y) { return (y >> 2) * c->gw + (x >> 2); }

static int luma_avail(const pic_ctx* c, int x, int y, uint32_t cur_cell) {
    if (x < 0 || y < 0 || x >= c->w || y >= c->h) return 0;
    uint32_t info = c->cell[cell_at(c, x, y)];
    if ((c->P->flags & HPVCD_PF_TILES) && HPVCD_CELL_TILE(info) != HPVCD_CELL_TILE(cur_cell)) return 0;
    if (c->P->num_slices > 1 && HPVCD_CELL_SLICE(info) != HPVCD_CELL_SLICE(cur_cell)) return 0;
    if ((c->P->flags & HPVCD_PF_CONSTRAINED_INTRA) && !(info & HPVCD_CELL_INTRA)) return 0;
    return c->decoded[cell_at(c, x, y)];
}

static int chroma_avail(const pic_ctx* c, int cx, int cy, uint32_t cur_cell) {
    if (cx < 0 || cy < 0 || cx >= c->cw || cy >= c->ch) return 0;
    int lx = cx * c->sub_w, ly = cy * c->sub_h;
    if (lx >= c->w || ly >= c->h) return 0;
    return luma_avail(c, lx, ly, cur_cell);
}

static void predict_block(pic_ctx* c, int comp, int x0, int y0, int n, int mode, uint16_t* pred) {
    /* src/decode.rs:4956 predict_luma_block_into / :5202 predict_chroma_block_into */
    const hpvcd_picture* P = c->P;
    int lx = comp ? x0 * c->sub_w : x0, ly = comp ? y0 * c->sub_h : y0;
    uint32_t cur = c->cell[cell_at(c, lx < c->w ? lx : c->w - 1, ly < c->h ? ly : c->h - 1)];
    uint16_t* plane = comp == 0 ? c->y : (comp == 1 ? c->cb : c->cr);
    int stride = comp ? c->cw : c->w;
    int32_t raw_above[64], raw_left[64];
    int corner;
    if (comp == 0) {
        corner = luma_avail(c, x0 - 1, y0 - 1, cur) ? plane[(y0 - 1) * stride + x0 - 1] : -1;
        for (int i = 0; i < 2 * n; i++) {
            raw_above[i] = luma_avail(c, x0 + i, y0 - 1, cur) ? plane[(y0 - 1) * stride + x0 + i] : -1;
            raw_left[i] = luma_avail(c, x0 - 1, y0 + i, cur) ? plane[(y0 + i) * stride + x0 - 1] : -1;
        }
        int strong = (P->flags & HPVCD_PF_STRONG_INTRA_SMOOTHING) != 0; /* strong_smoothing is always set (src/decode.rs:583) */
        oracle_intra_predict(corner, raw_above, raw_left, n, mode, 1,
                             !(P->flags & HPVCD_PF_INTRA_SMOOTHING_DISABLED), strong,
                             !(P->flags & HPVCD_PF_INTRA_BOUNDARY_FILTER_OFF), c->bd, pred);
    } else {
        corner = chroma_avail(c, x0 - 1, y0 - 1, cur) ? plane[(y0 - 1) * stride + x0 - 1] : -1;
        for (int i = 0; i < 2 * n; i++) {
            raw_above[i] = chroma_avail(c, x0 + i, y0 - 1, cur) ? plane[(y0 - 1) * stride + x0 + i] : -1;
            raw_left[i] = chroma_avail(c, x0 - 1, y0 + i, cur) ? plane[(y0 + i) * stride + x0 - 1] : -1;
        }
        /* 4:4:4 chroma filters with the luma [1 2 1] rule, no strong smoothing;
         * other formats never filter; chroma never gets edge filters. */
        int is444 = P->chroma_idc == 3;
        oracle_intra_predict(corner, raw_above, raw_left, n, mode, is444,
                             is444 && !(P->flags & HPVCD_PF_INTRA_SMOOTHING_DISABLED), 0, 0, c->bd_c, pred);
    }
}

/* ------------------------------------------------------------------------ */
/* residual path                                                            */
/* ------------------------------------------------------------------------ */
static const uint8_t* scaling_for(const pic_ctx* c, int comp, int n, int inter, int ts) {
    /* src/decode.rs:5874 scaling_matrix_from_lists */
    if (ts && n != 4) return NULL;
    if (c->P->scaling_index < 0) return NULL;
    int size_id = (n == 4 ? 0 : n == 8 ? 1 : n == 16 ? 2 : 3);
    int matrix_id = (comp < 2 ? comp : 2) + (inter ? 3 : 0);
    const uint8_t* set = c->b->scaling + (size_t)c->P->scaling_index * HPVCD_SCALING_SET_BYTES;
    return set + (size_id * 6 + matrix_id) * HPVCD_SCALING_ENTRY_BYTES;
}

static void rext_ops(int32_t* res, int n, int rotate, int rdpcm, int i16) {
    /* src/decode.rs:8418 apply_rext_residual_ops (i16 storage wraps) */
    if (rotate) {
        for (int i = 0; i < n * n / 2; i++) {
            int32_t t = res[i]; res[i] = res[n * n - 1 - i]; res[n * n - 1 - i] = t;
        }
    }
    if (rdpcm == 1) {
        for (int y = 1; y < n; y++)
            for (int x = 0; x < n; x++) {
                int32_t v = wadd32(res[y * n + x], res[(y - 1) * n + x]);
                res[y * n + x] = i16 ? wrap16(v) : v;
            }
    } else if (rdpcm == 0) {
        for (int y = 0; y < n; y++)
            for (int x = 1; x < n; x++) {
                int32_t v = wadd32(res[y * n + x], res[y * n + x - 1]);
                res[y * n + x] = i16 ? wrap16(v) : v;
            }
    }
}

/* Residual of one TB (the dequant / transform half of reconstruct_luma and
 * reconstruct_chroma).  Returns 1 when the i16 storage path was used. */
static int tb_residual(pic_ctx* c, const hpvcd_tb* t, int32_t* res) {
    int n = 1 << t->log2n;
    int comp = t->comp;
    int bd = comp ? c->bd_c : c->bd;
    int inter = (t->flags & HPVCD_TB_INTER) != 0;
    int ts = (t->flags & HPVCD_TB_TS) != 0;
    int tqb = (t->flags & HPVCD_TB_TQB) != 0;
    int32_t levels[32 * 32];
    memset(levels, 0, sizeof(int32_t) * n * n);
    const uint32_t* co = c->b->coeffs + c->P->coeff_begin + t->coeff_off;
    for (uint32_t i = 0; i < t->coeff_cnt; i++) levels[HPVCD_COEFF_POS(co[i])] = HPVCD_COEFF_LEVEL(co[i]);

    /* src/decode.rs:4356 rext_post_ops */
    int rotate = 0, rdpcm = -1;
    if (ts || tqb) {
        rotate = ts && !tqb && (c->P->flags & HPVCD_PF_TS_ROTATION) && n == 4 && !inter;
        if (inter) {
            if (t->flags & HPVCD_TB_RDPCM_V) rdpcm = 1;
            else if (t->flags & HPVCD_TB_RDPCM_H) rdpcm = 0;
        } else if ((c->P->flags & HPVCD_PF_IMPLICIT_RDPCM) && (t->mode == 10 || t->mode == 26)) {
            rdpcm = t->mode == 26;
        }
    }
    if (t->flags & HPVCD_TB_PCM) {
        /* src/decode.rs:2788 pcm_sample << (BitDepth - PcmBitDepth) */
        for (int i = 0; i < n * n; i++) res[i] = (int32_t)((uint32_t)levels[i] << t->qp);
        return 0;
    }
    /* i16 storage: luma at bd<=8 without EP (src/decode.rs:4755); chroma
     * additionally needs res_scale == 0 (src/decode.rs:5315).  The ACT path
     * keeps all three residuals in i32 (src/decode.rs:3660). */
    int i16 = bd <= 8 && !c->ep && (comp == 0 || t->res_scale == 0) && !(t->flags & HPVCD_TB_ACT);
    int max_range = oracle_max_range(bd, c->ep);
    if (tqb) {
        for (int i = 0; i < n * n; i++)
            res[i] = i16 ? clamp32(levels[i], -32768, 32767) : levels[i]; /* src/reconstruct.rs:125 */
        rext_ops(res, n, rotate, rdpcm, i16);
    } else {
        const uint8_t* sc = scaling_for(c, comp, n, inter, ts);
        if (ts) {
            oracle_dequant_ts(levels, n, t->qp, bd, c->ep, sc, res);
            if (i16) for (int i = 0; i < n * n; i++) res[i] = wrap16(res[i]);
            rext_ops(res, n, rotate, rdpcm, i16);
        } else {
            int32_t deq[32 * 32];
            oracle_dequant(levels, n, t->qp, bd, c->ep, sc, deq);
            if (comp == 0 && n == 4 && !inter) oracle_inv_dst4(deq, bd, max_range, i16, res);
            else oracle_inv_dct(deq, n, bd, max_range, t->nx, i16, res);
        }
    }
    return i16;
}

static void add_residual(uint16_t* plane, int stride, int pw, int ph, int x0, int y0, int n, const uint16_t* pred,
                         const int32_t* res, int bd) {
    /* src/reconstruct.rs:226 add_residual_generic over the valid region */
    int vw = pw - x0 < n ? pw - x0 : n, vh = ph - y0 < n ? ph - y0 : n;
    int32_t max = (1 << bd) - 1;
    for (int y = 0; y < vh; y++)
        for (int x = 0; x < vw; x++)
            plane[(y0 + y) * stride + x0 + x] =
                (uint16_t)clamp32(wadd32(pred[y * n + x], res ? res[y * n + x] : 0), 0, max);
}

static void mark_decoded(pic_ctx* c, int x0, int y0, int n) {
    for (int y = y0; y < y0 + n && y < c->h; y += 4)
        for (int x = x0; x < x0 + n && x < c->w; x += 4) c->decoded[cell_at(c, x, y)] = 1;
}

/* Prediction of one TB: intra, or the MC result already in the plane for an
 * inter TB (load_plane_block_clipped, src/decode.rs:4744), or zero for I_PCM
 * (its samples are the whole reconstruction). */
static void tb_prediction(pic_ctx* c, const hpvcd_tb* t, uint16_t* pred) {
    int n = 1 << t->log2n;
    int comp = t->comp;
    uint16_t* plane = comp == 0 ? c->y : (comp == 1 ? c->cb : c->cr);
    int stride = comp ? c->cw : c->w, pw = comp ? c->cw : c->w, ph = comp ? c->ch : c->h;
    if (t->flags & HPVCD_TB_PCM) {
        memset(pred, 0, sizeof(uint16_t) * n * n);
    } else if (!(t->flags & HPVCD_TB_INTER)) {
        predict_block(c, comp, t->x, t->y, n, t->mode, pred);
    } else {
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) {
                int px = t->x + x, py = t->y + y;
                pred[y * n + x] = (px < pw && py < ph) ? plane[py * stride + px] : 0;
            }
    }
}

/* One ACT triple (src/decode.rs:3619 transform_unit_act): the three i32
 * residuals (zero where cbf is 0), cross-component prediction of Cb and Cr
 * from the luma residual, the inverse colour transform per sample, then
 * prediction and residual add in the order luma, Cb, Cr. */
static void recon_act(pic_ctx* c, const hpvcd_tb* t) {
    int n = 1 << t->log2n;
    static __thread int32_t res[3][32 * 32];
    uint16_t pred[32 * 32];
    for (int k = 0; k < 3; k++) {
        if (t[k].flags & HPVCD_TB_CBF) tb_residual(c, &t[k], res[k]);
        else memset(res[k], 0, sizeof(int32_t) * n * n);
    }
    for (int k = 1; k < 3; k++) {
        if (!t[k].res_scale) continue;
        for (int i = 0; i < n * n; i++) { /* src/decode.rs:5488 apply_cross_comp_residual */
            int64_t sy = ((int64_t)res[0][i] << c->bd_c) >> c->bd;
            int64_t delta = ((int64_t)t[k].res_scale * sy) >> 3;
            res[k][i] = (int32_t)clamp64((int64_t)res[k][i] + delta, INT32_MIN, INT32_MAX);
        }
    }
    for (int i = 0; i < n * n; i++) oracle_act_inverse(&res[0][i], &res[1][i], &res[2][i]);
    for (int k = 0; k < 3; k++) {
        int comp = t[k].comp;
        uint16_t* plane = comp == 0 ? c->y : (comp == 1 ? c->cb : c->cr);
        int stride = comp ? c->cw : c->w, ph = comp ? c->ch : c->h;
        tb_prediction(c, &t[k], pred);
        add_residual(plane, stride, stride, ph, t[k].x, t[k].y, n, pred, res[k], comp ? c->bd_c : c->bd);
        if (comp == 0) mark_decoded(c, t[k].x, t[k].y, n);
    }
}

static void recon_tb(pic_ctx* c, const hpvcd_tb* t) {
    int n = 1 << t->log2n;
    int comp = t->comp;
    int inter = (t->flags & HPVCD_TB_INTER) != 0;
    int cbf = (t->flags & (HPVCD_TB_CBF | HPVCD_TB_PCM)) != 0; /* PCM samples travel as levels */
    uint16_t* plane = comp == 0 ? c->y : (comp == 1 ? c->cb : c->cr);
    int stride = comp ? c->cw : c->w, pw = comp ? c->cw : c->w, ph = comp ? c->ch : c->h;
    int bd = comp ? c->bd_c : c->bd;
    uint16_t pred[32 * 32];
    int32_t res[32 * 32];

    tb_prediction(c, t, pred);
    if (comp == 0) {
        if (cbf) {
            tb_residual(c, t, res);
            memcpy(c->ccl, res, sizeof(int32_t) * n * n); /* capture for CCP (src/decode.rs:4808,4856) */
            add_residual(plane, stride, pw, ph, t->x, t->y, n, pred, res, bd);
        } else if (!inter) {
            add_residual(plane, stride, pw, ph, t->x, t->y, n, pred, NULL, bd); /* predict_only_luma */
        }
        mark_decoded(c, t->x, t->y, n);
        return;
    }
    /* chroma (src/decode.rs:5280 reconstruct_chroma, :5428, :5480) */
    if (cbf) {
        tb_residual(c, t, res);
    } else if (t->res_scale != 0) {
        memset(res, 0, sizeof(int32_t) * n * n);
    } else {
        if (!inter) add_residual(plane, stride, pw, ph, t->x, t->y, n, pred, NULL, bd);
        return;
    }
    if (t->res_scale != 0) {
        /* src/decode.rs:5488 apply_cross_comp_residual */
        for (int i = 0; i < n * n; i++) {
            int64_t sy = ((int64_t)c->ccl[i] << c->bd_c) >> c->bd;
            int64_t delta = ((int64_t)t->res_scale * sy) >> 3;
            res[i] = (int32_t)clamp64((int64_t)res[i] + delta, INT32_MIN, INT32_MAX);
        }
    }
    add_residual(plane, stride, pw, ph, t->x, t->y, n, pred, res, bd);
}


/* ------------------------------------------------------------------------ */
/* motion compensation (src/mc.rs, src/decode.rs:7000-7485)                  */
/* ------------------------------------------------------------------------ */
/* HEVC luma (quarter-sample) and chroma (eighth-sample) interpolation
 * filters, src/mc.rs:30-46 (Tables 8-11/8-12 of the standard). */
static const int LUMA_FILTER[4][8] = {{0, 0, 0, 64, 0, 0, 0, 0},
                                      {-1, 4, -10, 58, 17, -5, 1, 0},
                                      {-1, 4, -11, 40, 40, -11, 4, -1},
                                      {0, 1, -5, 17, 58, -10, 4, -1}};
static const int CHROMA_FILTER[8][4] = {{0, 64, 0, 0},    {-2, 58, 10, -2}, {-4, 54, 16, -2}, {-6, 46, 28, -4},
                                        {-4, 36, 36, -4}, {-4, 28, 46, -6}, {-2, 16, 54, -4}, {-2, 10, 58, -2}};

/* src/mc.rs:310 RefPlane::sample: reads clamp to the plane */
static int ref_sample(const uint16_t* d, int w, int h, int x, int y) {
    x = x < 0 ? 0 : (x >= w ? w - 1 : x);
    y = y < 0 ? 0 : (y >= h ? h - 1 : y);
    return d[(size_t)y * w + x];
}

/* src/mc.rs:351 luma_interp_scalar_scratch (taps 8, quarter fractions) and
 * :442 chroma_interp_scalar_scratch (taps 4, eighth fractions): the four
 * cases of the reference, 14-bit intermediates narrowed to i16. */
static void mc_interp(const uint16_t* d, int pw, int ph, int x0, int y0, int fx, int fy, int w, int h, int bd,
                      int taps, const int* hf, const int* vf, int16_t* dst) {
    int s1 = bd - 8, half = taps / 2 - 1;
    if (fx == 0 && fy == 0) {
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) dst[y * w + x] = (int16_t)((ref_sample(d, pw, ph, x0 + x, y0 + y) << 6) >> s1);
        return;
    }
    if (fy == 0) {
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int acc = 0;
                for (int t = 0; t < taps; t++) acc += hf[t] * ref_sample(d, pw, ph, x0 + x - half + t, y0 + y);
                dst[y * w + x] = (int16_t)(acc >> s1);
            }
        return;
    }
    if (fx == 0) {
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int acc = 0;
                for (int t = 0; t < taps; t++) acc += vf[t] * ref_sample(d, pw, ph, x0 + x, y0 + y - half + t);
                dst[y * w + x] = (int16_t)(acc >> s1);
            }
        return;
    }
    int th = h + taps - 1;
    int32_t* tmp = malloc(sizeof(int32_t) * (size_t)w * th);
    for (int ty = 0; ty < th; ty++)
        for (int x = 0; x < w; x++) {
            int acc = 0;
            for (int t = 0; t < taps; t++) acc += hf[t] * ref_sample(d, pw, ph, x0 + x - half + t, y0 + ty - half);
            tmp[ty * w + x] = acc >> s1;
        }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int acc = 0;
            for (int t = 0; t < taps; t++) acc += vf[t] * tmp[(y + t) * w + x];
            dst[y * w + x] = (int16_t)(acc >> 6);
        }
    free(tmp);
}

/* src/mc.rs:538-733: uni / bi / weighted uni / weighted bi into the valid
 * region of the plane; neither list usable gives mid-grey (src/decode.rs:7206) */
static void mc_write(uint16_t* plane, int stride, int vw, int vh, int w, const int16_t* p0, const int16_t* p1,
                     int use0, int use1, const hpvcd_ref* e0, const hpvcd_ref* e1, int comp, int bd) {
    int max = (1 << bd) - 1;
    int dn = comp ? 1 : 0;
    for (int y = 0; y < vh; y++)
        for (int x = 0; x < vw; x++) {
            int i = y * w + x, v;
            if (use0 && use1) {
                if (e0->weighted)
                    v = oracle_weighted_bi(p0[i], p1[i], e0->weight[comp], e1->weight[comp], e0->offset[comp],
                                           e1->offset[comp], e0->log2_denom[dn], bd);
                else {
                    int sh = 15 - bd;
                    v = clamp32((p0[i] + p1[i] + (1 << (sh - 1))) >> sh, 0, max);
                }
            } else if (use0 || use1) {
                const hpvcd_ref* e = use0 ? e0 : e1;
                int s = use0 ? p0[i] : p1[i];
                if (e->weighted)
                    v = oracle_weighted_uni(s, e->weight[comp], e->offset[comp], e->log2_denom[dn], bd);
                else {
                    int sh = 14 - bd;
                    v = clamp32((s + (sh > 0 ? 1 << (sh - 1) : 0)) >> sh, 0, max);
                }
            } else {
                v = 1 << (bd - 1);
            }
            plane[(size_t)y * stride + x] = (uint16_t)v;
        }
}

/* src/decode.rs:7000 motion_compensate of one PU: luma, then Cb and Cr with
 * the chroma vector mv * 2 / SubWidthC (eighth samples at 4:2:0). */
static void mc_pu(pic_ctx* c, const hpvcd_pu* pu) {
    const hpvcd_ref* e[2] = {NULL, NULL};
    int use[2] = {0, 0};
    for (int l = 0; l < 2; l++) {
        if (!((pu->flags >> l) & 1)) continue;
        e[l] = &c->b->refs[c->P->ref_begin + pu->ref[l]];
        use[l] = e[l]->picture != HPVCD_REF_MISSING && c->refp && c->refp[3 * e[l]->picture] != NULL;
    }
    int w = pu->w, h = pu->h;
    int16_t* pred[2] = {malloc(sizeof(int16_t) * w * h), malloc(sizeof(int16_t) * w * h)};
    for (int l = 0; l < 2; l++)
        if (use[l]) {
            int mx = pu->mv[l][0], my = pu->mv[l][1];
            mc_interp(c->refp[3 * e[l]->picture], c->w, c->h, pu->x + (mx >> 2), pu->y + (my >> 2), mx & 3, my & 3, w,
                      h, c->bd, 8, LUMA_FILTER[mx & 3], LUMA_FILTER[my & 3], pred[l]);
        }
    if (pu->x < c->w && pu->y < c->h) {
        int vw = c->w - pu->x < w ? c->w - pu->x : w, vh = c->h - pu->y < h ? c->h - pu->y : h;
        mc_write(c->y + (size_t)pu->y * c->w + pu->x, c->w, vw, vh, w, pred[0], pred[1], use[0], use[1], e[0], e[1],
                 0, c->bd);
    }
    if (c->P->chroma_idc) {
        int cw = w / c->sub_w, ch = h / c->sub_h, cx = pu->x / c->sub_w, cy = pu->y / c->sub_h;
        for (int pl = 1; pl <= 2; pl++) {
            for (int l = 0; l < 2; l++)
                if (use[l]) {
                    int mx = pu->mv[l][0] * 2 / c->sub_w, my = pu->mv[l][1] * 2 / c->sub_h;
                    mc_interp(c->refp[3 * e[l]->picture + pl], c->cw, c->ch, cx + (mx >> 3), cy + (my >> 3), mx & 7,
                              my & 7, cw, ch, c->bd_c, 4, CHROMA_FILTER[mx & 7], CHROMA_FILTER[my & 7], pred[l]);
                }
            if (cx >= c->cw || cy >= c->ch) continue;
            int vw = c->cw - cx < cw ? c->cw - cx : cw, vh = c->ch - cy < ch ? c->ch - cy : ch;
            uint16_t* plane = pl == 1 ? c->cb : c->cr;
            mc_write(plane + (size_t)cy * c->cw + cx, c->cw, vw, vh, cw, pred[0], pred[1], use[0], use[1], e[0], e[1],
                     pl, c->bd_c);
        }
    }
    free(pred[0]);
    free(pred[1]);
}

/* ------------------------------------------------------------------------ */
/* deblocking (src/decode.rs:1673-2143)                                      */
/* ------------------------------------------------------------------------ */
static const int BETA[52] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15,
                             16, 17, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56,
                             58, 60, 62, 64};
static const int TC[54] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                           2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 5, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 22, 24};

static uint32_t cell_info_at(const pic_ctx* c, int x, int y) { return c->cell[cell_at(c, x, y)]; }

static int slice_of(const pic_ctx* c, int x, int y) {
    if (x >= c->w || y >= c->h) return 0;
    return (int)HPVCD_CELL_SLICE(cell_info_at(c, x, y));
}

static hpvcd_slice_params slice_deblock_at(const pic_ctx* c, int x, int y) {
    int s = slice_of(c, x, y);
    if (s < c->P->num_slices) return c->P->slices[s];
    hpvcd_slice_params d = {(uint8_t)((c->P->flags & HPVCD_PF_DEBLOCK_DISABLED) != 0), 0, 0, 1};
    return d;
}

static int filter_across(const pic_ctx* c, int pxp, int pyp, int pxq, int pyq) {
    /* src/decode.rs:1712 filter_across_boundary */
    int sq = slice_of(c, pxq, pyq);
    int q_across = sq < c->P->num_slices ? c->P->slices[sq].lf_across
                                         : (c->P->flags & HPVCD_PF_LF_ACROSS_SLICES) != 0;
    if (!q_across && slice_of(c, pxp, pyp) != sq) return 0;
    if ((c->P->flags & HPVCD_PF_TILES) && !(c->P->flags & HPVCD_PF_LF_ACROSS_TILES)) {
        if (HPVCD_CELL_TILE(cell_info_at(c, pxp, pyp)) != HPVCD_CELL_TILE(cell_info_at(c, pxq, pyq))) return 0;
    }
    return 1;
}

static int deblock_bs_v(const pic_ctx* c, int px, int py) {
    /* src/decode.rs:1673 */
    if (px == 0 || px >= c->w || py >= c->h) return 0;
    uint32_t info = cell_info_at(c, px, py);
    if (!HPVCD_CELL_BSV(info)) return 0;
    if (!filter_across(c, px - 1, py, px, py)) return 0;
    return (int)HPVCD_CELL_BSV(info);
}

static int deblock_bs_h(const pic_ctx* c, int px, int py) {
    /* src/decode.rs:1692 */
    if (py == 0 || px >= c->w || py >= c->h) return 0;
    uint32_t info = cell_info_at(c, px, py);
    if (!HPVCD_CELL_BSH(info)) return 0;
    if (!filter_across(c, px, py - 1, px, py)) return 0;
    return (int)HPVCD_CELL_BSH(info);
}

static int qp_at(const pic_ctx* c, int px, int py) {
    int gx = px / 4, gy = py / 4;
    if (gx > c->gw - 1) gx = c->gw - 1;
    if (gy > c->gh - 1) gy = c->gh - 1;
    return HPVCD_CELL_QP(c->cell[gy * c->gw + gx]);
}

static int cell_suppressed(const pic_ctx* c, int g) {
    uint32_t info = c->cell[g];
    return (info & HPVCD_CELL_TQB) || ((c->P->flags & HPVCD_PF_PCM_LF_DISABLED) && (info & HPVCD_CELL_PCM));
}

static int any_suppressed(const pic_ctx* c) {
    for (int g = 0; g < c->gw * c->gh; g++) if (cell_suppressed(c, g)) return 1;
    return 0;
}

static void restore_luma(pic_ctx* c, const uint16_t* snap) {
    for (int gy = 0; gy < c->gh; gy++)
        for (int gx = 0; gx < c->gw; gx++) {
            if (!cell_suppressed(c, gy * c->gw + gx)) continue;
            for (int dy = 0; dy < 4 && gy * 4 + dy < c->h; dy++)
                for (int dx = 0; dx < 4 && gx * 4 + dx < c->w; dx++) {
                    int i = (gy * 4 + dy) * c->w + gx * 4 + dx;
                    c->y[i] = snap[i];
                }
        }
}

static void restore_chroma(pic_ctx* c, uint16_t* plane, const uint16_t* snap) {
    for (int cy = 0; cy < c->ch; cy++)
        for (int cx = 0; cx < c->cw; cx++) {
            int lx = cx * c->sub_w, ly = cy * c->sub_h;
            int g = (ly / 4) * c->gw + (lx / 4);
            if (g < c->gw * c->gh && cell_suppressed(c, g)) plane[cy * c->cw + cx] = snap[cy * c->cw + cx];
        }
}

static void luma_segment(pic_ctx* c, int vertical, int edge, int scan, int beta, int tc) {
    int32_t lines[4][8];
    int w = c->w;
    for (int l = 0; l < 4; l++)
        for (int t = -4; t < 4; t++)
            lines[l][t + 4] = vertical ? c->y[(scan + l) * w + edge + t] : c->y[(edge + t) * w + scan + l];
    oracle_deblock_luma_segment(lines, beta, tc, (1 << c->bd) - 1);
    for (int l = 0; l < 4; l++)
        for (int t = -4; t < 4; t++) {
            if (vertical) c->y[(scan + l) * w + edge + t] = (uint16_t)lines[l][t + 4];
            else c->y[(edge + t) * w + scan + l] = (uint16_t)lines[l][t + 4];
        }
}

static void apply_deblocking(pic_ctx* c) {
    /* src/decode.rs:1774 apply_deblocking (serial) */
    int w = c->w, h = c->h, cw = c->cw, ch = c->ch;
    int bd_shift = c->bd - 8, bd_shift_c = c->bd_c - 8;
    int supp = any_suppressed(c);
    uint16_t *snap_y = NULL, *snap_cb = NULL, *snap_cr = NULL;
    if (supp) {
        snap_y = malloc(sizeof(uint16_t) * w * h);
        memcpy(snap_y, c->y, sizeof(uint16_t) * w * h);
        if (cw && ch) {
            snap_cb = malloc(sizeof(uint16_t) * cw * ch);
            snap_cr = malloc(sizeof(uint16_t) * cw * ch);
            memcpy(snap_cb, c->cb, sizeof(uint16_t) * cw * ch);
            memcpy(snap_cr, c->cr, sizeof(uint16_t) * cw * ch);
        }
    }
    for (int pass = 0; pass < 2; pass++) {
        int edge_max = pass == 0 ? w : h, scan_max = pass == 0 ? h : w;
        int last = edge_max >= 4 ? edge_max - 4 : 0;
        for (int edge = 8; edge <= last; edge += 8) {
            for (int scan = 0; scan + 4 <= scan_max; scan += 4) {
                int mid = scan + 1;
                int bs = pass == 0 ? deblock_bs_v(c, edge, mid) : deblock_bs_h(c, mid, edge);
                if (!bs) continue;
                hpvcd_slice_params sd = pass == 0 ? slice_deblock_at(c, edge, mid) : slice_deblock_at(c, mid, edge);
                if (sd.disabled) continue;
                int qp_p = pass == 0 ? qp_at(c, edge - 1, mid) : qp_at(c, mid, edge - 1);
                int qp_q = pass == 0 ? qp_at(c, edge, mid) : qp_at(c, mid, edge);
                int avg = (qp_p + qp_q + 1) >> 1;
                int beta_i = clamp32(avg + sd.beta_offset_div2 * 2, 0, 51);
                int tc_i = clamp32(avg + 2 * (bs - 1) + sd.tc_offset_div2 * 2, 0, 53);
                int beta = BETA[beta_i] << bd_shift, tc = TC[tc_i] << bd_shift;
                if (!tc) continue;
                if (pass == 0) { if (scan + 4 <= h) luma_segment(c, 1, edge, scan, beta, tc); }
                else if (scan + 4 <= w && edge + 4 <= h) luma_segment(c, 0, edge, scan, beta, tc);
            }
        }
        if (snap_y) restore_luma(c, snap_y);
    }
    if (c->P->chroma_idc == 0) goto done;
    for (int pass = 0; pass < 2; pass++) {
        int scan_max = pass == 0 ? ch : cw;
        int chroma_edge_max = pass == 0 ? cw : ch;
        int last = chroma_edge_max >= 2 ? chroma_edge_max - 2 : 0;
        int maxv_c = (1 << c->bd_c) - 1;
        for (int edge = 8; edge <= last; edge += 8) {
            for (int scan = 0; scan + 4 <= scan_max; scan += 4) {
                int mid = scan + 1;
                int lex = pass == 0 ? edge * c->sub_w : mid * c->sub_w;
                int ley = pass == 0 ? mid * c->sub_h : edge * c->sub_h;
                int bs = pass == 0 ? deblock_bs_v(c, lex, ley) : deblock_bs_h(c, lex, ley);
                if (bs < 2) continue;
                int lqx = lex < w - 1 ? lex : w - 1, lqy = ley < h - 1 ? ley : h - 1;
                hpvcd_slice_params sd = slice_deblock_at(c, lqx, lqy);
                if (sd.disabled) continue;
                int tc_offset = sd.tc_offset_div2 * 2;
                int ppx, ppy;
                if (pass == 0) {
                    ppx = lex > 0 ? lex - 1 : 0; if (ppx > w - 1) ppx = w - 1;
                    ppy = ley < h - 1 ? ley : h - 1;
                } else {
                    ppx = lex < w - 1 ? lex : w - 1;
                    ppy = ley > 0 ? ley - 1 : 0; if (ppy > h - 1) ppy = h - 1;
                }
                int avg = (qp_at(c, ppx, ppy) + qp_at(c, lqx, lqy) + 1) >> 1;
                for (int pl = 0; pl < 2; pl++) {
                    int off = pl == 0 ? c->P->pps_cb_qp_offset : c->P->pps_cr_qp_offset;
                    int qpc = oracle_qpc(avg + off, c->P->chroma_idc, c->bd_c);
                    int tci = clamp32(qpc + 2 + tc_offset, 0, 53);
                    int tcc = TC[tci] << bd_shift_c;
                    if (!tcc) continue;
                    uint16_t* pix = pl == 0 ? c->cb : c->cr;
                    for (int s = scan; s < scan + 4; s++) {
                        if (s >= scan_max) continue;
                        int ip, iq, ip1, iq1;
                        if (pass == 0) { ip = s * cw + edge - 1; ip1 = s * cw + edge - 2; iq = s * cw + edge; iq1 = s * cw + edge + 1; }
                        else { ip = (edge - 1) * cw + s; ip1 = (edge - 2) * cw + s; iq = edge * cw + s; iq1 = (edge + 1) * cw + s; }
                        int32_t p0 = pix[ip], p1 = pix[ip1], q0 = pix[iq], q1 = pix[iq1];
                        int32_t delta = clamp32(((q0 - p0) * 4 + p1 - q1 + 4) >> 3, -tcc, tcc);
                        if (delta) {
                            pix[ip] = (uint16_t)clamp32(p0 + delta, 0, maxv_c);
                            pix[iq] = (uint16_t)clamp32(q0 - delta, 0, maxv_c);
                        }
                    }
                }
            }
        }
        if (snap_cb) { restore_chroma(c, c->cb, snap_cb); restore_chroma(c, c->cr, snap_cr); }
    }
done:
    free(snap_y); free(snap_cb); free(snap_cr);
}

/* ------------------------------------------------------------------------ */
/* SAO (src/decode.rs:2258 apply_sao, src/sao.rs:179-470)                    */
/* ------------------------------------------------------------------------ */
static void sao_band(uint16_t* dst, int w, int x0, int y0, int x1, int y1, const int16_t* off, int band_pos, int bd) {
    /* src/sao.rs:236 (in place); band index compared without wrap-around */
    int32_t max = (1 << bd) - 1;
    int shift = bd - 5;
    for (int y = y0; y < y1; y++)
        for (int x = x0; x < x1; x++) {
            int32_t s = dst[y * w + x];
            uint8_t rel = (uint8_t)((uint8_t)(s >> shift) - (uint8_t)band_pos);
            if (rel < 4) dst[y * w + x] = (uint16_t)clamp32(s + off[rel], 0, max);
        }
}

typedef struct {
    const pic_ctx* c;
    int chroma;
    int lf_across_slices;  /* picture-level: pps flag && every slice flag (src/decode.rs:2247) */
    int lf_across_tiles;
} sao_bnd;

static int sao_neighbor_ok(const sao_bnd* s, int cx, int cy, int nx, int ny) {
    /* src/sao.rs:378 luma_neighbor_ok / :417 chroma_neighbor_ok */
    const pic_ctx* c = s->c;
    if (s->chroma) { cx *= c->sub_w; cy *= c->sub_h; nx *= c->sub_w; ny *= c->sub_h; }
    int gc = (cy / 4) * c->gw + (cx / 4);
    int gn = (ny / 4) * c->gw + (nx / 4);
    uint32_t ic = c->cell[gc], in = c->cell[gn];
    if (ic & HPVCD_CELL_TQB) return 0;
    if ((c->P->flags & HPVCD_PF_PCM_LF_DISABLED) && (ic & HPVCD_CELL_PCM)) return 0;
    if (!s->lf_across_slices && HPVCD_CELL_SLICE(ic) != HPVCD_CELL_SLICE(in)) return 0;
    if (!s->lf_across_tiles && (c->P->flags & HPVCD_PF_TILES) && HPVCD_CELL_TILE(ic) != HPVCD_CELL_TILE(in)) return 0;
    return 1;
}

static void sao_edge(uint16_t* dst, const uint16_t* src, int w, int h, int x0, int y0, int x1, int y1,
                     const int16_t* off, int eo_class, int bd, const sao_bnd* gate) {
    /* src/sao.rs:266 apply_sao_edge_offset_scalar / :433 ..._gated */
    int32_t max = (1 << bd) - 1;
    int dx = eo_class == 1 ? 0 : 1;
    int dy = eo_class == 0 ? 0 : (eo_class == 3 ? -1 : 1);
    for (int y = y0; y < y1; y++)
        for (int x = x0; x < x1; x++) {
            int32_t s = src[y * w + x];
            int x1n = x + dx, y1n = y + dy, x2n = x - dx, y2n = y - dy;
            int ok1 = x1n >= 0 && y1n >= 0 && x1n < w && y1n < h;
            int ok2 = x2n >= 0 && y2n >= 0 && x2n < w && y2n < h;
            if (gate) {
                ok1 = ok1 && sao_neighbor_ok(gate, x, y, x1n, y1n);
                ok2 = ok2 && sao_neighbor_ok(gate, x, y, x2n, y2n);
            }
            if (!ok1 || !ok2) continue;
            int32_t n1 = src[y1n * w + x1n], n2 = src[y2n * w + x2n];
            int e = ((s > n1) - (s < n1)) + ((s > n2) - (s < n2)) + 2;
            int32_t o = e == 0 ? off[0] : e == 1 ? off[1] : e == 3 ? off[2] : e == 4 ? off[3] : 0;
            if (o) dst[y * w + x] = (uint16_t)clamp32(s + o, 0, max);
        }
}

void oracle_sao_ctb(uint16_t* dst, const uint16_t* src, int w, int h, int x0, int y0, int x1, int y1, int type,
                    const int16_t* offsets, int band_pos, int eo_class, int bd) {
    /* one CTB of one plane, ungated (src/sao.rs:179 apply_sao_plane_scalar) */
    if (type == 1) {
        for (int y = y0; y < y1; y++)
            for (int x = x0; x < x1; x++) dst[y * w + x] = src[y * w + x];
        sao_band(dst, w, x0, y0, x1, y1, offsets, band_pos, bd);
    } else if (type == 2) {
        sao_edge(dst, src, w, h, x0, y0, x1, y1, offsets, eo_class, bd, NULL);
    }
}

static void apply_sao(pic_ctx* c) {
    const hpvcd_picture* P = c->P;
    int sao_luma = (P->flags & HPVCD_PF_SAO_LUMA) != 0;
    int sao_chroma = (P->flags & HPVCD_PF_SAO_CHROMA) != 0 && c->cw && c->ch && P->chroma_idc;
    int ctb = 1 << P->log2_ctb;
    int ctb_cols = (c->w + ctb - 1) / ctb, ctb_rows = (c->h + ctb - 1) / ctb;
    const hpvcd_sao* sao = c->b->sao + P->ctb_begin;
    int active = 0, needs_src[3] = {0, 0, 0};
    for (int i = 0; i < ctb_cols * ctb_rows; i++) {
        if (sao_luma) { active |= sao[i].type_idx[0] != 0; needs_src[0] |= sao[i].type_idx[0] == 2; }
        if (sao_chroma) {
            active |= sao[i].type_idx[1] != 0 || sao[i].type_idx[2] != 0;
            needs_src[1] |= sao[i].type_idx[1] == 2; needs_src[2] |= sao[i].type_idx[2] == 2;
        }
    }
    if (!active) return;
    size_t ny = (size_t)c->w * c->h, nc = (size_t)c->cw * c->ch;
    uint16_t* orig[3] = {NULL, NULL, NULL};
    if (needs_src[0]) { orig[0] = malloc(ny * 2); memcpy(orig[0], c->y, ny * 2); }
    if (needs_src[1]) { orig[1] = malloc(nc * 2); memcpy(orig[1], c->cb, nc * 2); }
    if (needs_src[2]) { orig[2] = malloc(nc * 2); memcpy(orig[2], c->cr, nc * 2); }
    /* src/decode.rs:2210 sao_boundary_restricted */
    int all_slices_across = 1;
    for (int s = 0; s < P->num_slices; s++) all_slices_across &= P->slices[s].lf_across != 0;
    int supp = any_suppressed(c);
    int restricted = !all_slices_across || !(P->flags & HPVCD_PF_LF_ACROSS_SLICES) ||
                     ((P->flags & HPVCD_PF_TILES) && !(P->flags & HPVCD_PF_LF_ACROSS_TILES)) || supp;
    sao_bnd gl = {c, 0, (P->flags & HPVCD_PF_LF_ACROSS_SLICES) && all_slices_across,
                  !(P->flags & HPVCD_PF_TILES) || (P->flags & HPVCD_PF_LF_ACROSS_TILES)};
    sao_bnd gc = gl;
    gc.chroma = 1;
    uint16_t *snap_y = NULL, *snap_cb = NULL, *snap_cr = NULL;
    if (supp) {
        snap_y = malloc(ny * 2); memcpy(snap_y, c->y, ny * 2);
        if (nc) { snap_cb = malloc(nc * 2); memcpy(snap_cb, c->cb, nc * 2); snap_cr = malloc(nc * 2); memcpy(snap_cr, c->cr, nc * 2); }
    }
    for (int ry = 0; ry < ctb_rows; ry++)
        for (int rx = 0; rx < ctb_cols; rx++) {
            const hpvcd_sao* s = &sao[ry * ctb_cols + rx];
            int x0 = rx * ctb, y0 = ry * ctb;
            if (sao_luma && s->type_idx[0]) {
                int xe = x0 + ctb < c->w ? x0 + ctb : c->w, ye = y0 + ctb < c->h ? y0 + ctb : c->h;
                if (s->type_idx[0] == 1) sao_band(c->y, c->w, x0, y0, xe, ye, s->offsets[0], s->band_pos[0], c->bd);
                else sao_edge(c->y, orig[0], c->w, c->h, x0, y0, xe, ye, s->offsets[0], s->eo_class[0], c->bd,
                              restricted ? &gl : NULL);
            }
            if (sao_chroma) {
                int cx0 = x0 / c->sub_w, cy0 = y0 / c->sub_h;
                int cxe = (x0 + ctb) / c->sub_w, cye = (y0 + ctb) / c->sub_h;
                if (cxe > c->cw) cxe = c->cw;
                if (cye > c->ch) cye = c->ch;
                for (int pl = 0; pl < 2; pl++) {
                    int ci = 1 + pl;
                    uint16_t* buf = pl == 0 ? c->cb : c->cr;
                    if (s->type_idx[ci] == 1)
                        sao_band(buf, c->cw, cx0, cy0, cxe, cye, s->offsets[ci], s->band_pos[ci], c->bd_c);
                    else if (s->type_idx[ci] == 2)
                        sao_edge(buf, orig[ci], c->cw, c->ch, cx0, cy0, cxe, cye, s->offsets[ci], s->eo_class[ci],
                                 c->bd_c, restricted ? &gc : NULL);
                }
            }
        }
    if (supp) {
        restore_luma(c, snap_y);
        if (snap_cb) { restore_chroma(c, c->cb, snap_cb); restore_chroma(c, c->cr, snap_cr); }
    }
    free(orig[0]); free(orig[1]); free(orig[2]);
    free(snap_y); free(snap_cb); free(snap_cr);
}

/* ------------------------------------------------------------------------ */
/* picture driver                                                           */
/* ------------------------------------------------------------------------ */
/* Decode picture p with the final planes of its references in refp. */
static int decode_planes(const hpvcd_batch* b, uint32_t p, int stage, uint16_t* y, uint16_t* cb, uint16_t* cr,
                         uint16_t* const* refp) {
    if (p >= b->num_pictures) return -1;
    pic_ctx* c = calloc(1, sizeof(pic_ctx));
    const hpvcd_picture* P = &b->pictures[p];
    c->b = b; c->P = P;
    c->w = P->width; c->h = P->height;
    c->sub_w = sub_w_of(P->chroma_idc); c->sub_h = sub_h_of(P->chroma_idc);
    c->cw = P->chroma_idc ? (c->w + c->sub_w - 1) / c->sub_w : 0;
    c->ch = P->chroma_idc ? (c->h + c->sub_h - 1) / c->sub_h : 0;
    c->gw = (c->w + 3) / 4; c->gh = (c->h + 3) / 4;
    c->bd = P->bit_depth_luma; c->bd_c = P->bit_depth_chroma;
    c->ep = (P->flags & HPVCD_PF_EXTENDED_PRECISION) != 0;
    c->y = y; c->cb = cb; c->cr = cr;
    c->cell = b->cell_info + P->cell_begin;
    c->refp = refp;
    c->decoded = calloc((size_t)c->gw * c->gh, 1);
    memset(y, 0, sizeof(uint16_t) * c->w * c->h);
    if (c->cw) { memset(cb, 0, sizeof(uint16_t) * c->cw * c->ch); memset(cr, 0, sizeof(uint16_t) * c->cw * c->ch); }
    /* inter prediction reads only other pictures, so every PU is predicted
     * before the TBs that add residuals on top */
    for (uint32_t i = 0; i < P->pu_count; i++) mc_pu(c, &b->pus[P->pu_begin + i]);
    for (uint32_t i = 0; i < P->tb_count; i++) {
        const hpvcd_tb* t = &b->tbs[P->tb_begin + i];
        if ((t->flags & HPVCD_TB_ACT) && t->comp == 0 && i + 2 < P->tb_count) {
            recon_act(c, t);
            i += 2;
        } else {
            recon_tb(c, t);
        }
    }
    int deblock_any = 0;
    for (int s = 0; s < (P->num_slices ? P->num_slices : 1); s++)
        deblock_any |= P->num_slices ? !P->slices[s].disabled : !(P->flags & HPVCD_PF_DEBLOCK_DISABLED);
    if (stage >= ORACLE_STAGE_DEBLOCK && deblock_any) apply_deblocking(c);
    if (stage >= ORACLE_STAGE_SAO) apply_sao(c);
    free(c->decoded);
    free(c);
    return 0;
}

static size_t chroma_elems(const hpvcd_picture* P) {
    int sw = sub_w_of(P->chroma_idc), sh = sub_h_of(P->chroma_idc);
    return P->chroma_idc ? (size_t)((P->width + sw - 1) / sw) * ((P->height + sh - 1) / sh) : 0;
}

static void mark_refs(const hpvcd_batch* b, uint32_t p, uint8_t* need) {
    const hpvcd_picture* P = &b->pictures[p];
    for (uint32_t k = 0; k < P->ref_count; k++) {
        uint32_t r = b->refs[P->ref_begin + k].picture;
        if (r != HPVCD_REF_MISSING && r < p) need[r] = 1;
    }
}

int oracle_picture_planes(const hpvcd_batch* b, uint32_t p, int stage, uint16_t* y, uint16_t* cb, uint16_t* cr) {
    if (p >= b->num_pictures) return -1;
    /* the references' own references, transitively; decoded in batch order */
    uint8_t* need = calloc(p + 1, 1);
    mark_refs(b, p, need);
    for (uint32_t q = p; q-- > 0;)
        if (need[q]) mark_refs(b, q, need);
    uint16_t** refp = calloc(3 * (size_t)b->num_pictures + 3, sizeof(uint16_t*));
    int rc = 0;
    for (uint32_t q = 0; q < p && !rc; q++) {
        if (!need[q]) continue;
        const hpvcd_picture* Q = &b->pictures[q];
        size_t nc = chroma_elems(Q);
        refp[3 * q] = malloc(sizeof(uint16_t) * Q->width * Q->height);
        refp[3 * q + 1] = malloc(sizeof(uint16_t) * nc + 2);
        refp[3 * q + 2] = malloc(sizeof(uint16_t) * nc + 2);
        rc = decode_planes(b, q, ORACLE_STAGE_SAO, refp[3 * q], refp[3 * q + 1], refp[3 * q + 2], refp);
    }
    if (!rc) rc = decode_planes(b, p, stage, y, cb, cr, refp);
    for (size_t k = 0; k < 3 * (size_t)b->num_pictures; k++) free(refp[k]);
    free(refp);
    free(need);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* YUV -> RGB (src/yuv.rs:187-668)                                           */
/* ------------------------------------------------------------------------ */
int oracle_picture_rgb(const hpvcd_batch* b, uint32_t p, const uint16_t* Y, const uint16_t* Cb, const uint16_t* Cr,
                       void* out) {
    const hpvcd_picture* P = &b->pictures[p];
    const hpvcd_output* O = &b->outputs[P->out_index];
    int w = P->width, h = P->height;
    int sub_w = sub_w_of(P->chroma_idc), sub_h = sub_h_of(P->chroma_idc);
    int cw = P->chroma_idc ? (w + sub_w - 1) / sub_w : 0, ch = P->chroma_idc ? (h + sub_h - 1) / sub_h : 0;
    int dw = P->win_w, dh = P->win_h;
    if (dw <= 0 || dh <= 0 || w <= 0 || h <= 0) return 0;
    int x0 = P->crop_left < w - 1 ? P->crop_left : w - 1;
    int y0 = P->crop_top < h - 1 ? P->crop_top : h - 1;
    int bd = P->bit_depth_luma;
    int64_t scale = (int64_t)1 << (bd - 8);
    int full = (P->flags & HPVCD_PF_FULL_RANGE) != 0;
    /* src/yuv.rs:187 Cvt::new */
    int64_t k_y = full ? 8192 : 9539, k_c = full ? 8192 : 9326;
    int64_t m0, m1, m2, m3;
    switch (P->matrix) {
        case 5: case 6: m0 = 11485; m1 = -2819; m2 = -5851; m3 = 14516; break;
        case 9: m0 = 12080; m1 = -1348; m2 = -4681; m3 = 17546; break;
        case 0: m0 = m1 = m2 = m3 = 0; break;
        default: m0 = 12901; m1 = -1534; m2 = -3835; m3 = 15201; break;
    }
    int64_t y_black = full ? 0 : 16 * scale, c_black = y_black, neutral = 128 * scale;
    int64_t cr_r = (m0 * k_c + 4096) >> 13, cb_g = (m1 * k_c + 4096) >> 13;
    int64_t cr_g = (m2 * k_c + 4096) >> 13, cb_b = (m3 * k_c + 4096) >> 13;
    int kind = O->kind;
    int eight = bd == 8;
    int64_t cmax = eight ? 255 : (1 << bd) - 1;   /* RGB16 keeps bd-bit values */
    int narrow16 = kind == HPVCD_OUT_RGB8_FROM16; /* src/lib.rs:1969 */
    uint8_t* o8 = (uint8_t*)out;
    uint16_t* o16 = (uint16_t*)out;
    int mono = P->chroma_idc == 0;
    /* luma outputs: one channel; RGB outputs of a 4:0:0 picture repeat the
     * luma value (src/lib.rs:1978 decode_heic_rgb8_with Luma8/Luma16 arms) */
    int chn = (O->kind == HPVCD_OUT_LUMA8 || O->kind == HPVCD_OUT_LUMA16) ? 1 : 3;
    for (int yy = 0; yy < dh; yy++) {
        int src_y = y0 + yy;
        int y_row = src_y < h - 1 ? src_y : h - 1;
        int c_row = ch ? (src_y / sub_h < ch - 1 ? src_y / sub_h : ch - 1) : 0;
        for (int xx = 0; xx < dw; xx++) {
            int src_x = x0 + xx;
            int x_col = src_x < w - 1 ? src_x : w - 1;
            int c_col = cw ? (src_x / sub_w < cw - 1 ? src_x / sub_w : cw - 1) : 0;
            int64_t v[3];
            int64_t L = Y[y_row * w + x_col];
            if (mono) {
                /* src/yuv.rs:269 mono_rows (edge columns replicate the last in-range one, same as the clamp) */
                v[0] = v[1] = v[2] = (k_y * (L - y_black) + 4096) >> 13;
            } else {
                int64_t U = Cb[c_row * cw + c_col], V = Cr[c_row * cw + c_col];
                if (P->matrix == 0) {
                    /* src/yuv.rs:291 gbr_rows */
                    if (full) { v[0] = V; v[1] = L; v[2] = U; }
                    else {
                        v[0] = (k_c * (V - c_black) + 4096) >> 13;
                        v[1] = (k_y * (L - y_black) + 4096) >> 13;
                        v[2] = (k_c * (U - c_black) + 4096) >> 13;
                    }
                } else if (P->matrix == 8) {
                    /* src/yuv.rs:368 ycgco_rows */
                    int64_t cg = U - neutral, co = V - neutral, t = L - cg;
                    v[0] = t + co; v[1] = L + cg; v[2] = t - co;
                } else {
                    /* src/yuv.rs:232 pixel / :327 fast_rows */
                    int64_t yt = k_y * (L - y_black), cbc = U - neutral, crc = V - neutral;
                    v[0] = (yt + cr_r * crc + 4096) >> 13;
                    v[1] = (yt + cb_g * cbc + cr_g * crc + 4096) >> 13;
                    v[2] = (yt + cb_b * cbc + 4096) >> 13;
                }
            }
            for (int k = 0; k < chn; k++) {
                int64_t s = clamp64(v[k], 0, cmax);
                size_t idx = ((size_t)yy * dw + xx) * chn + k;
                if (eight) o8[idx] = (uint8_t)s;
                else if (narrow16) o8[idx] = (uint8_t)((uint16_t)s >> (bd - 8));
                else o16[idx] = (uint16_t)s;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* batch driver                                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
    const hpvcd_batch* b;
    void** outputs;
    int next;            /* shared work index */
    pthread_mutex_t mu;
    int njobs;
    int* job_first;      /* pictures per job: [job_first[j], job_first[j+1]) in job_pics order */
    uint32_t* job_pics;
    uint16_t** refp;     /* final planes kept for later pictures: [3 * picture + plane] */
    const uint8_t* keep; /* pictures referenced by a later picture */
    int err;
} batch_job;

/* Raw planes of one picture into its output (Decoder::decode_yuv,
 * src/lib.rs:762-1207).  Single item: plane_window (src/lib.rs:143) of the
 * coded planes at the conformance-window offset, dw x dh of the display size,
 * source rows and columns clamped to the plane (the padding repeats the last
 * sample of a row, the last row repeats).  Grid tile: stitch_native_yuv_band
 * (src/lib.rs:882): luma at (col * tile_w, row * tile_h), chroma at
 * (col * tile_cw, row * tile_ch) with tile_cw = ceil(tile_w / sub_w),
 * copies clipped to the output planes, crop offsets clamped to the plane.
 * In the records a single item is a picture whose window is the whole
 * output; tcw below is tile_cw for a tile and ceil(dw / sub_w) for it. */
static void oracle_picture_yuv(const hpvcd_batch* b, uint32_t p, const uint16_t* Y, const uint16_t* Cb,
                               const uint16_t* Cr, void* out) {
    const hpvcd_picture* P = &b->pictures[p];
    const hpvcd_output* O = &b->outputs[P->out_index];
    int wide = HPVCD_OUT_YUV_WIDE(O->kind);
    uint8_t* o8 = (uint8_t*)out;
    uint16_t* o16 = (uint16_t*)out;
    int w = P->width, h = P->height, W = O->width, H = O->height;
#define PUT(i, v) do { if (wide) o16[(i)] = (uint16_t)(v); else o8[(i)] = (uint8_t)(v); } while (0)
    int x0 = P->crop_left < w - 1 ? P->crop_left : w - 1;
    int y0 = P->crop_top < h - 1 ? P->crop_top : h - 1;
    for (int r = 0; r < P->copy_h; r++) {
        int sy = y0 + r < h - 1 ? y0 + r : h - 1;
        for (int x = 0; x < P->copy_w; x++) {
            int sx = x0 + x < w - 1 ? x0 + x : w - 1;
            PUT((size_t)(P->dst_y + r) * W + P->dst_x + x, Y[(size_t)sy * w + sx]);
        }
    }
    if (P->chroma_idc == 0 || P->win_w <= 0 || P->win_h <= 0) return;
    int sw = sub_w_of(P->chroma_idc), sh = sub_h_of(P->chroma_idc);
    int cw = (w + sw - 1) / sw, ch = (h + sh - 1) / sh;
    int cwo = (W + sw - 1) / sw, cho = (H + sh - 1) / sh;
    int tcw = ((P->win_w > P->copy_w ? P->win_w : P->copy_w) + sw - 1) / sw;
    int tch = ((P->win_h > P->copy_h ? P->win_h : P->copy_h) + sh - 1) / sh;
    int cdx = (P->dst_x / P->win_w) * tcw, cdy = (P->dst_y / P->win_h) * tch;
    int ccw = cwo - cdx < tcw ? cwo - cdx : tcw, cch = cho - cdy < tch ? cho - cdy : tch;
    if (P->copy_w <= 0 || P->copy_h <= 0 || ccw <= 0 || cch <= 0) return;
    int cl = P->crop_left / sw < cw - 1 ? P->crop_left / sw : cw - 1;
    int ct = P->crop_top / sh < ch - 1 ? P->crop_top / sh : ch - 1;
    for (int k = 0; k < 2; k++) {
        const uint16_t* C = k ? Cr : Cb;
        size_t base = (size_t)W * H + (size_t)k * cwo * cho;
        for (int r = 0; r < cch; r++) {
            int sy = ct + r < ch - 1 ? ct + r : ch - 1;
            for (int x = 0; x < ccw; x++) {
                int sx = cl + x < cw - 1 ? cl + x : cw - 1;
                PUT(base + (size_t)(cdy + r) * cwo + cdx + x, C[(size_t)sy * cw + sx]);
            }
        }
    }
#undef PUT
}

static int decode_one(const hpvcd_batch* b, uint32_t p, void** outputs, uint16_t** refp, int keep) {
    const hpvcd_picture* P = &b->pictures[p];
    const hpvcd_output* O = &b->outputs[P->out_index];
    size_t ny = (size_t)P->width * P->height;
    size_t nc = chroma_elems(P);
    uint16_t* y = malloc(ny * 2);
    uint16_t* cb = malloc(nc * 2 + 2);
    uint16_t* cr = malloc(nc * 2 + 2);
    int chn = (O->kind == HPVCD_OUT_LUMA8 || O->kind == HPVCD_OUT_LUMA16) ? 1 : 3;
    int bps = (O->kind == HPVCD_OUT_RGB16 || O->kind == HPVCD_OUT_LUMA16) ? 2 : 1;
    int raw = HPVCD_OUT_IS_YUV(O->kind);
    uint8_t* win = raw ? NULL : malloc((size_t)P->win_w * P->win_h * chn * bps + 1);
    int rc = decode_planes(b, p, ORACLE_STAGE_SAO, y, cb, cr, refp);
    if (!rc && raw && outputs[P->out_index]) oracle_picture_yuv(b, p, y, cb, cr, outputs[P->out_index]);
    if (!rc && !raw) rc = oracle_picture_rgb(b, p, y, cb, cr, win);
    if (!rc && !raw) {
        uint8_t* dst = (uint8_t*)outputs[P->out_index];
        size_t row = (size_t)P->copy_w * chn * bps;
        for (int r = 0; r < P->copy_h; r++)
            memcpy(dst + (((size_t)(P->dst_y + r) * O->width + P->dst_x) * chn) * bps,
                   win + (size_t)r * P->win_w * chn * bps, row);
    }
    if (keep && !rc) {
        refp[3 * p] = y; refp[3 * p + 1] = cb; refp[3 * p + 2] = cr;
    } else {
        free(y); free(cb); free(cr);
    }
    free(win);
    return rc;
}

static void* batch_worker(void* arg) {
    batch_job* J = (batch_job*)arg;
    for (;;) {
        pthread_mutex_lock(&J->mu);
        int j = J->next++;
        pthread_mutex_unlock(&J->mu);
        if (j >= J->njobs) break;
        for (int k = J->job_first[j]; k < J->job_first[j + 1]; k++) {
            uint32_t p = J->job_pics[k];
            if (decode_one(J->b, p, J->outputs, J->refp, J->keep[p])) J->err = 1;
        }
    }
    return NULL;
}

/* src/lib.rs:2017 apply_orientation with :2065 rotate_luma / :2116
 * rotate_buf on a whole output image: FlipH, Rotate180, FlipV, Rotate90 and
 * Rotate270 move pixels; Transpose and Transverse (5, 7) keep the buffer as
 * decoded (the reference's match arms fall through to the identity). */
static int orient_output(const hpvcd_output* O, void* buf) {
    int o = O->orientation;
    if (HPVCD_OUT_IS_YUV(O->kind)) return 0;  /* raw planes keep the coded orientation */
    if (o < HPVCD_ORIENT_FLIP_H || o == HPVCD_ORIENT_TRANSPOSE || o == HPVCD_ORIENT_TRANSVERSE) return 0;
    size_t px = (size_t)O->width * O->height;
    size_t chn = (O->kind == HPVCD_OUT_LUMA8 || O->kind == HPVCD_OUT_LUMA16) ? 1 : 3;
    size_t bps = (O->kind == HPVCD_OUT_RGB16 || O->kind == HPVCD_OUT_LUMA16) ? 2 : 1;
    size_t pb = chn * bps;
    uint8_t* src = malloc(px * pb + 1);
    if (!src) return -1;
    memcpy(src, buf, px * pb);
    uint8_t* dst = (uint8_t*)buf;
    size_t W = (size_t)O->width, H = (size_t)O->height;
    for (size_t y = 0; y < H; y++)
        for (size_t x = 0; x < W; x++) {
            size_t idx;
            if (o == HPVCD_ORIENT_FLIP_H) idx = y * W + (W - 1 - x);
            else if (o == HPVCD_ORIENT_ROTATE_180) idx = (H - 1 - y) * W + (W - 1 - x);
            else if (o == HPVCD_ORIENT_FLIP_V) idx = (H - 1 - y) * W + x;
            else if (o == HPVCD_ORIENT_ROTATE_90) idx = x * H + (H - 1 - y);
            else idx = (W - 1 - x) * H + y;
            memcpy(dst + idx * pb, src + (y * W + x) * pb, pb);
        }
    free(src);
    return 0;
}

int oracle_decode_batch(const hpvcd_batch* b, void** outputs, int threads, int rows_parallel) {
    /* src/lib.rs:1680 fill_grid_bands: one job per grid row (pictures sharing
     * dst_y of one output), tiles of a row serial; or one job per picture.
     * Inter pictures wait for their references: the batch runs in waves of
     * equal reference depth, each wave's pictures in parallel. */
    batch_job J;
    memset(&J, 0, sizeof(J));
    J.b = b; J.outputs = outputs;
    pthread_mutex_init(&J.mu, NULL);
    uint32_t n = b->num_pictures;
    J.job_pics = malloc(sizeof(uint32_t) * (n + 1));
    J.job_first = malloc(sizeof(int) * (n + 2));
    J.refp = calloc(3 * (size_t)n + 3, sizeof(uint16_t*));
    uint8_t* keep = calloc(n + 1, 1);
    uint32_t* depth = calloc(n + 1, sizeof(uint32_t));
    uint32_t max_depth = 0;
    for (uint32_t p = 0; p < n; p++) {
        const hpvcd_picture* P = &b->pictures[p];
        for (uint32_t k = 0; k < P->ref_count; k++) {
            uint32_t r = b->refs[P->ref_begin + k].picture;
            if (r == HPVCD_REF_MISSING || r >= p) continue;
            keep[r] = 1;
            if (depth[r] + 1 > depth[p]) depth[p] = depth[r] + 1;
        }
        if (depth[p] > max_depth) max_depth = depth[p];
    }
    J.keep = keep;
    if (threads < 1) threads = 1;
    pthread_t* th = malloc(sizeof(pthread_t) * threads);
    uint8_t* used = calloc(n + 1, 1);
    for (uint32_t d = 0; d <= max_depth; d++) {
        int nj = 0, k = 0;
        for (uint32_t p = 0; p < n; p++) {
            if (used[p] || depth[p] != d) continue;
            J.job_first[nj++] = k;
            J.job_pics[k++] = p; used[p] = 1;
            if (rows_parallel)
                for (uint32_t q = p + 1; q < n; q++)
                    if (!used[q] && depth[q] == d && b->pictures[q].out_index == b->pictures[p].out_index &&
                        b->pictures[q].dst_y == b->pictures[p].dst_y) { J.job_pics[k++] = q; used[q] = 1; }
        }
        J.job_first[nj] = k;
        J.njobs = nj;
        J.next = 0;
        for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, batch_worker, &J);
        for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    }
    for (size_t q = 0; q < 3 * (size_t)n; q++) free(J.refp[q]);
    for (uint32_t o = 0; o < b->num_outputs && !J.err; o++)
        if (outputs[o] && orient_output(&b->outputs[o], outputs[o])) J.err = 1;
    free(th); free(used); free(J.job_pics); free(J.job_first); free(J.refp); free(keep); free(depth);
    pthread_mutex_destroy(&J.mu);
    return J.err ? -1 : 0;
}
