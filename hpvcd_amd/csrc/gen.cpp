// gen.cpp — seeded syntax-level generator of reconstruction records.
//
// It draws the coding tree, prediction modes, CBFs and coefficient levels of
// synthetic intra pictures from configurable distributions and emits the same
// records the host syntax parse produces, in decode order.  The parse-side
// derivations follow the reference decoder:
//   coding quadtree / CU / transform tree / TU  src/decode.rs:2495, :2585, :3315, :3478
//   chroma mode derivation (Table 8-2)         src/decode.rs:3248 decode_chroma_mode
//   deblocking edge marking (bS = 2, intra)    src/decode.rs:3018 mark_block_edges
//   QP map / chroma QP                         src/decode.rs:3126 set_qp, :5045-5055, :7599 qpc
//   SAO offset signs and scaling               src/decode.rs:1507-1614 parse_sao
//   CCP ResScaleVal                            src/decode.rs:4257 (log2_res_scale_abs_plus1, sign)
//   grid placement / crop window               src/lib.rs:1610-1653, :1754-1817
//   inter CU, part_mode, PU edges, skip        src/decode.rs:6380 decode_inter_cu, :6462
//   inter bS (intra side, coefficients, motion) src/decode.rs:2955 finalize_coeff_bs, :620
//   pred_weight_table rows                     src/inter.rs:141 PredWeightTable
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "../../include/hpvcd.h"
#include "internal.h"

namespace hpvcd {

static inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t tile_seed(uint64_t base, uint32_t config_id, uint32_t tile) {
    uint64_t s = base ^ ((uint64_t)config_id << 32) ^ (uint64_t)tile;
    return splitmix64(s);
}

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() { return splitmix64(s); }
    bool bern(float p) { return (double)(next() >> 11) * (1.0 / 9007199254740992.0) < (double)p; }
    uint32_t uni(uint32_t n) { return (uint32_t)(next() % n); }
    int sint(int lo, int hi) { return lo + (int)uni((uint32_t)(hi - lo + 1)); }
};

constexpr uint8_t MODE_DC = 1;

struct PicGen {
    const hpvcd_gen_config& cfg;
    Rng rng;
    int w, h, gw, gh, sub_w, sub_h;
    int log2_min_cb, log2_min_tb, log2_max_tb;
    int qp_y, bd;
    std::vector<hpvcd_tb>& tbs;
    std::vector<uint32_t>& coeffs;
    std::vector<hpvcd_pu>& pus;
    std::vector<hpvcd_ref>& refs;
    // inter pictures: reference entries (list 0 first), per-cell motion and
    // the transform / prediction edges for the boundary strengths
    bool inter_pic = false;
    int nref[2] = {0, 0};
    std::vector<uint32_t> ref_pic;
    struct Motion {
        uint8_t intra = 1, flags = 0;
        int32_t pic[2] = {-1, -1};
        int16_t mv[2][2] = {{0, 0}, {0, 0}};
    };
    std::vector<Motion> motion;
    std::vector<uint8_t> tu_v, tu_h, pu_v, pu_h, nz;
    bool cu_inter = false;
    std::vector<uint32_t> cell;      // info word without QP (QP packed at the end)
    std::vector<int8_t> cell_qp;
    std::vector<uint8_t> mode_y;     // luma intra mode per 4x4 (src/decode.rs mode_y)
    uint32_t coeff_base;
    // current CU
    uint8_t cu_chroma_modes[4];
    bool cu_chroma_dm[4];
    int cu_x0, cu_y0, cu_pu;
    bool cu_tqb;
    // quantisation groups and the QpY predictor (src/decode.rs:2509 QG reset,
    // :2590 QG entry, :3084 predict_qp); qp_y is the running QpY
    int log2_qg = 6, log2_ctb = 6;
    int slice_qp = 0, qp_prev = 0, qp_bd_off = 0;
    bool qp_delta_coded = false;
    // per raster CTB: slice index and tile id
    std::vector<uint8_t> ctb_slice, ctb_tile;
    int ctb_cols = 1;
    size_t tb_first = 0;

    PicGen(const hpvcd_gen_config& c, uint64_t seed, std::vector<hpvcd_tb>& t, std::vector<uint32_t>& co,
           std::vector<hpvcd_pu>& pu_out, std::vector<hpvcd_ref>& ref_out, uint32_t pic_index)
        : cfg(c), rng{seed}, tbs(t), coeffs(co), pus(pu_out), refs(ref_out) {
        w = c.tile_width;
        h = c.tile_height;
        gw = (w + 3) / 4;
        gh = (h + 3) / 4;
        sub_w = (c.chroma_idc == 1 || c.chroma_idc == 2) ? 2 : 1;
        sub_h = c.chroma_idc == 1 ? 2 : 1;
        log2_min_cb = c.log2_min_cb;
        log2_min_tb = c.log2_min_tb;
        log2_max_tb = c.log2_max_tb;
        bd = c.bit_depth;
        qp_bd_off = 6 * (bd - 8);
        qp_y = std::clamp<int>(c.qp, -qp_bd_off, 51);
        slice_qp = qp_prev = qp_y;
        log2_ctb = c.log2_ctb;
        log2_qg = std::max<int>(c.log2_min_cb, c.log2_ctb - (c.cu_qp_delta ? c.diff_cu_qp_delta_depth : 0));
        ctb_cols = (w + (1 << log2_ctb) - 1) >> log2_ctb;
        tb_first = t.size();
        cell.assign((size_t)gw * gh, HPVCD_CELL_INTRA);
        cell_qp.assign((size_t)gw * gh, (int8_t)qp_y);
        mode_y.assign((size_t)gw * gh, MODE_DC);
        coeff_base = (uint32_t)co.size();
        inter_pic = c.inter && pic_index > 0;
        if (inter_pic) setup_refs(pic_index);
    }

    // Low-delay reference lists: list 0 holds the nearest earlier pictures,
    // list 1 (B pictures) the same pictures farthest first.  Each entry is a
    // pred_weight_table row (src/inter.rs:141): weights around 1 << denom,
    // offsets in 8-bit units scaled to the bit depth.
    void setup_refs(uint32_t pic_index) {
        int n = std::max(1, std::min<int>(cfg.num_refs ? cfg.num_refs : 1, (int)pic_index));
        nref[0] = n;
        nref[1] = cfg.p_bi > 0.f ? n : 0;
        for (int l = 0; l < 2; l++)
            for (int i = 0; i < nref[l]; i++) {
                hpvcd_ref e;
                std::memset(&e, 0, sizeof(e));
                e.picture = pic_index - 1 - (uint32_t)(l == 0 ? i : n - 1 - i);
                e.weighted = cfg.weighted ? 1 : 0;
                e.log2_denom[0] = 6;
                e.log2_denom[1] = 6;
                for (int k = 0; k < 3; k++) {
                    bool signalled = cfg.weighted && rng.bern(0.8f);
                    e.weight[k] = (int16_t)(signalled ? rng.sint(32, 96) : 64);
                    e.offset[k] = (int16_t)(signalled ? rng.sint(-16, 16) * (1 << (bd - 8)) : 0);
                }
                refs.push_back(e);
                ref_pic.push_back(e.picture);
            }
        size_t ncell = (size_t)gw * gh;
        motion.assign(ncell, Motion());
        tu_v.assign(ncell, 0); tu_h.assign(ncell, 0); pu_v.assign(ncell, 0); pu_h.assign(ncell, 0); nz.assign(ncell, 0);
    }

    // transform-block boundary of an inter CU or TU (edge flags only; the
    // strength is derived once the picture is complete)
    void mark_tu_edges(int x0, int y0, int size) {
        if (x0 > 0 && x0 < w)
            for (int yy = y0; yy < std::min(y0 + size, h); yy += 4) tu_v[(yy / 4) * gw + x0 / 4] = 1;
        if (y0 > 0 && y0 < h)
            for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) tu_h[(y0 / 4) * gw + xx / 4] = 1;
    }

    static bool motion_differs(const Motion& p, const Motion& q) {
        // src/decode.rs:620 inter_motion_differs (reference pictures compared by identity)
        int32_t pr[2] = {(p.flags & 1) ? p.pic[0] : -1, (p.flags & 2) ? p.pic[1] : -1};
        int32_t qr[2] = {(q.flags & 1) ? q.pic[0] : -1, (q.flags & 2) ? q.pic[1] : -1};
        bool direct = pr[0] == qr[0] && pr[1] == qr[1];
        bool crossed = pr[0] == qr[1] && pr[1] == qr[0];
        if (!direct && !crossed) return true;
        auto mv = [](const Motion& m, int l, int c) -> int { return (m.flags >> l) & 1 ? m.mv[l][c] : 0; };
        auto far = [&](int pl, int ql) {
            return std::abs(mv(p, pl, 0) - mv(q, ql, 0)) >= 4 || std::abs(mv(p, pl, 1) - mv(q, ql, 1)) >= 4;
        };
        if (pr[0] != pr[1]) return direct ? (far(0, 0) || far(1, 1)) : (far(0, 1) || far(1, 0));
        return (far(0, 0) || far(1, 1)) && (far(0, 1) || far(1, 0));
    }

    // src/decode.rs:2955 finalize_coeff_bs over the whole inter picture
    void finalize_bs() {
        for (int gy = 0; gy < gh; gy++)
            for (int gx = 0; gx < gw; gx++) {
                size_t g = (size_t)gy * gw + gx;
                uint32_t& c = cell[g];
                if (gx > 0 && (tu_v[g] || pu_v[g])) {
                    uint32_t bs = c & 3u;
                    if (bs < 2 && (motion[g].intra || motion[g - 1].intra)) bs = 2;
                    else if (bs < 1 && tu_v[g] && (nz[g] || nz[g - 1])) bs = 1;
                    else if (bs < 1 && motion_differs(motion[g], motion[g - 1])) bs = 1;
                    c = (c & ~3u) | bs;
                }
                if (gy > 0 && (tu_h[g] || pu_h[g])) {
                    uint32_t bs = (c >> 2) & 3u;
                    if (bs < 2 && (motion[g].intra || motion[g - gw].intra)) bs = 2;
                    else if (bs < 1 && tu_h[g] && (nz[g] || nz[g - gw])) bs = 1;
                    else if (bs < 1 && motion_differs(motion[g], motion[g - gw])) bs = 1;
                    c = (c & ~0xcu) | (bs << 2);
                }
            }
    }

    double gauss() {
        double u1 = ((double)(rng.next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        double u2 = (double)(rng.next() >> 11) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }

    // One PU: prediction lists (no bi for 8x4 / 4x8), reference entries and
    // quarter-sample motion vectors ~ N(0, mv_sigma) keeping the block within
    // 64 samples of the picture.
    void prediction_unit(int px, int py, int pw, int ph) {
        hpvcd_pu pu;
        std::memset(&pu, 0, sizeof(pu));
        pu.x = (uint16_t)px; pu.y = (uint16_t)py; pu.w = (uint8_t)pw; pu.h = (uint8_t)ph;
        if (nref[1] > 0) {
            if (pw + ph != 12 && rng.bern(cfg.p_bi)) pu.flags = HPVCD_PU_L0 | HPVCD_PU_L1;
            else pu.flags = (rng.next() & 1) ? HPVCD_PU_L0 : HPVCD_PU_L1;
        } else {
            pu.flags = HPVCD_PU_L0;
        }
        Motion m;
        m.intra = 0;
        m.flags = pu.flags;
        for (int l = 0; l < 2; l++) {
            if (!((pu.flags >> l) & 1)) continue;
            int idx = (int)rng.uni((uint32_t)nref[l]);
            pu.ref[l] = (uint16_t)(l == 0 ? idx : nref[0] + idx);
            double sig = cfg.mv_sigma * 4.0;
            int mvx = (int)std::lround(gauss() * sig), mvy = (int)std::lround(gauss() * sig);
            mvx = std::clamp(mvx, (-64 - px - pw) * 4, (w + 64 - px) * 4);
            mvy = std::clamp(mvy, (-64 - py - ph) * 4, (h + 64 - py) * 4);
            pu.mv[l][0] = (int16_t)mvx;
            pu.mv[l][1] = (int16_t)mvy;
            m.pic[l] = (int32_t)ref_pic[pu.ref[l]];
            m.mv[l][0] = pu.mv[l][0];
            m.mv[l][1] = pu.mv[l][1];
        }
        pus.push_back(pu);
        for (int yy = py; yy < std::min(py + ph, h); yy += 4)
            for (int xx = px; xx < std::min(px + pw, w); xx += 4) motion[(yy / 4) * gw + xx / 4] = m;
    }

    // TBs without residual tiling an inter CU that codes none (skip, or
    // rqt_root_cbf = 0): the MC prediction is the reconstruction.
    void inter_tiles(int x0, int y0, int log2) {
        if (log2 > log2_max_tb) {
            int half = 1 << (log2 - 1);
            for (int k = 0; k < 4; k++) inter_tiles(x0 + (k & 1) * half, y0 + (k >> 1) * half, log2 - 1);
            return;
        }
        tbs.push_back(make_tb(x0, y0, log2, 0, MODE_DC, false));
        if (!cfg.chroma_idc) return;
        int cl = cfg.chroma_idc == 3 ? log2 : log2 - 1;
        for (int comp = 1; comp <= 2; comp++) tbs.push_back(make_tb(x0 / sub_w, y0 / sub_h, cl, comp, MODE_DC, false));
    }

    // src/decode.rs:6380 decode_inter_cu: part_mode, PUs in coding order,
    // then the residual tree or none
    void inter_cu(int x0, int y0, int log2) {
        int cb = 1 << log2;
        cu_inter = true;
        for (int yy = y0; yy < std::min(y0 + cb, h); yy += 4)
            for (int xx = x0; xx < std::min(x0 + cb, w); xx += 4) cell[(yy / 4) * gw + xx / 4] &= ~HPVCD_CELL_INTRA;
        mark_tu_edges(x0, y0, cb);
        bool skip = rng.bern(cfg.p_skip);
        int part = 0;  // 2Nx2N, 2NxN, Nx2N, NxN, 2NxnU, 2NxnD, nLx2N, nRx2N
        if (!skip) {
            int opts[8], n = 0;
            opts[n++] = 0; opts[n++] = 1; opts[n++] = 2;
            if (log2 == log2_min_cb && log2 > 3) opts[n++] = 3;
            if (cfg.amp && log2 > log2_min_cb) for (int k = 4; k < 8; k++) opts[n++] = k;
            part = opts[rng.uni((uint32_t)n)];
        }
        int q = cb / 4, hf = cb / 2;
        int rects[4][4], nr = 0;
        auto add = [&](int x, int y, int rw, int rh) { rects[nr][0] = x; rects[nr][1] = y; rects[nr][2] = rw; rects[nr][3] = rh; nr++; };
        switch (part) {
            case 0: add(x0, y0, cb, cb); break;
            case 1: add(x0, y0, cb, hf); add(x0, y0 + hf, cb, hf); break;
            case 2: add(x0, y0, hf, cb); add(x0 + hf, y0, hf, cb); break;
            case 3: add(x0, y0, hf, hf); add(x0 + hf, y0, hf, hf); add(x0, y0 + hf, hf, hf); add(x0 + hf, y0 + hf, hf, hf); break;
            case 4: add(x0, y0, cb, q); add(x0, y0 + q, cb, cb - q); break;
            case 5: add(x0, y0, cb, cb - q); add(x0, y0 + cb - q, cb, q); break;
            case 6: add(x0, y0, q, cb); add(x0 + q, y0, cb - q, cb); break;
            default: add(x0, y0, cb - q, cb); add(x0 + cb - q, y0, q, cb); break;
        }
        for (int i = 0; i < nr; i++) {
            int px = rects[i][0], py = rects[i][1], pw = rects[i][2], ph = rects[i][3];
            if (px > x0)
                for (int yy = py; yy < std::min(py + ph, h); yy += 4) pu_v[(yy / 4) * gw + px / 4] = 1;
            if (py > y0)
                for (int xx = px; xx < std::min(px + pw, w); xx += 4) pu_h[(py / 4) * gw + xx / 4] = 1;
            prediction_unit(px, py, pw, ph);
        }
        bool root_cbf = !skip && rng.bern(0.85f);
        if (root_cbf) transform_tree(x0, y0, x0, y0, log2, 0, 0, false, cfg.max_transform_depth_intra, false, false);
        else inter_tiles(x0, y0, log2);
        for (int yy = y0; yy < std::min(y0 + cb, h); yy += 4)
            for (int xx = x0; xx < std::min(x0 + cb, w); xx += 4) cell_qp[(yy / 4) * gw + xx / 4] = (int8_t)qp_y;
        qp_prev = qp_y;
        cu_inter = false;
    }

    // src/decode.rs:3018 mark_block_edges (bS kept as the max of the marks)
    void mark_block_edges(int x0, int y0, int size, int bs) {
        if (x0 > 0 && x0 < w) {
            int gx = x0 / 4;
            for (int yy = y0; yy < std::min(y0 + size, h); yy += 4) {
                uint32_t& c = cell[(yy / 4) * gw + gx];
                uint32_t cur = c & 3u;
                if ((uint32_t)bs > cur) c = (c & ~3u) | (uint32_t)bs;
            }
        }
        if (y0 > 0 && y0 < h) {
            int gy = y0 / 4;
            for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) {
                uint32_t& c = cell[gy * gw + xx / 4];
                uint32_t cur = (c >> 2) & 3u;
                if ((uint32_t)bs > cur) c = (c & ~0xcu) | ((uint32_t)bs << 2);
            }
        }
    }

    void fill_cells(int x0, int y0, int size, uint32_t bit) {
        for (int yy = y0; yy < std::min(y0 + size, h); yy += 4)
            for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) cell[(yy / 4) * gw + xx / 4] |= bit;
    }

    void set_mode(int x0, int y0, int size, uint8_t mode) {
        for (int yy = y0; yy < std::min(y0 + size, h); yy += 4)
            for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) mode_y[(yy / 4) * gw + xx / 4] = mode;
    }

    // Table 8-2 chroma mode derivation (src/decode.rs:3248)
    static uint8_t derive_chroma_mode(uint32_t syntax, uint8_t luma_mode) {
        if (syntax == 4) return luma_mode;
        static const uint8_t cand[4] = {0, 26, 10, 1};
        uint8_t m = cand[syntax];
        return m == luma_mode ? 34 : m;
    }

    // Coefficient levels of one coded TB: Bernoulli significance, geometric
    // magnitude >= 1, random sign; at least one nonzero (the coded last
    // position exists whenever cbf is set).
    void emit_levels(hpvcd_tb& tb) {
        int n = 1 << tb.log2n;
        uint32_t start = (uint32_t)coeffs.size();
        int max_x = 0;
        int lmax = cfg.extended_precision ? HPVCD_COEFF_LEVEL_MAX : 32767;
        for (int pos = 0; pos < n * n; pos++) {
            if (!rng.bern(cfg.p_sig)) continue;
            int level = 1;
            while (level < lmax && !rng.bern(cfg.p_level_geom)) level++;
            if (rng.next() & 1) level = -level;
            coeffs.push_back(HPVCD_COEFF_PACK(pos, level));
            max_x = std::max(max_x, pos & (n - 1));
        }
        if (coeffs.size() == start) {
            int pos = (int)rng.uni((uint32_t)(n * n));
            int level = (rng.next() & 1) ? 1 : -1;
            coeffs.push_back(HPVCD_COEFF_PACK(pos, level));
            max_x = pos & (n - 1);
        }
        tb.coeff_off = start - coeff_base;
        tb.coeff_cnt = (uint32_t)coeffs.size() - start;
        tb.nx = (uint8_t)(max_x + 1);
    }

    int qp_prime(int qp, int bitd) {
        int off = 6 * (bitd - 8);
        return std::clamp(qp + off, 0, 51 + off);
    }
    static int qpc_table(int qpi, int chroma_idc, int bitd) {
        static const int T[14] = {29, 30, 31, 32, 33, 33, 34, 34, 35, 35, 36, 36, 37, 37};
        qpi = std::clamp(qpi, -6 * (bitd - 8), 57);
        if (chroma_idc != 1) return std::min(qpi, 51);
        if (qpi < 30) return qpi;
        if (qpi > 43) return qpi - 6;
        return T[qpi - 30];
    }

    // qPY_PRED of the quantisation group at (xqg, yqg): the left and above
    // QpY inside the same CTB, else the QpY of the previous CU in decode
    // order (src/decode.rs:3084 predict_qp; no WPP)
    int predict_qp(int xqg, int yqg) {
        const int ctb = 1 << log2_ctb;
        const int cx = (xqg / ctb) * ctb, cy = (yqg / ctb) * ctb;
        const int a = (xqg >= 1 && xqg - 1 >= cx) ? cell_qp[(size_t)(yqg / 4) * gw + (xqg - 1) / 4] : qp_prev;
        const int b = (yqg >= 1 && yqg - 1 >= cy) ? cell_qp[(size_t)((yqg - 1) / 4) * gw + xqg / 4] : qp_prev;
        return std::clamp((a + b + 1) >> 1, -qp_bd_off, 51);
    }
    // src/decode.rs:5768 derive_qpy_from_delta
    int qpy_from_delta(int pred, int delta) {
        const int m = 52 + qp_bd_off;
        int q = (pred + delta + 52 + 2 * qp_bd_off) % m;
        if (q < 0) q += m;
        return std::clamp(q - qp_bd_off, -qp_bd_off, 51);
    }
    // cu_qp_delta of the first TU with a coded residual in a quantisation
    // group (src/decode.rs:3528)
    void maybe_qp_delta(bool residual) {
        if (!cfg.cu_qp_delta || !residual || qp_delta_coded) return;
        qp_delta_coded = true;
        const int d = rng.sint(-(int)cfg.qp_delta_max, (int)cfg.qp_delta_max);
        qp_y = qpy_from_delta(qp_y, d);
    }

    // Qp' of a TB.  Under ACT the component offsets -5/-5/-3 apply to QpY
    // before the chroma mapping and no PPS chroma offsets (src/decode.rs:3788);
    // otherwise chroma adds the PPS and slice offsets (src/decode.rs:5045).
    int tb_qp(int comp, bool act) {
        if (!act) {
            if (comp == 0) return qp_prime(qp_y, bd);
            const int off = comp == 1 ? cfg.pps_cb_qp_offset + cfg.slice_cb_qp_offset
                                      : cfg.pps_cr_qp_offset + cfg.slice_cr_qp_offset;
            return qp_prime(qpc_table(qp_y + off, cfg.chroma_idc, bd), bd);
        }
        int qpi = qp_y + (comp == 2 ? -3 : -5);
        if (comp == 0) return qp_prime(std::clamp(qpi, -6 * (bd - 8), 51), bd);
        return qp_prime(qpc_table(qpi, cfg.chroma_idc, bd), bd);
    }

    hpvcd_tb make_tb(int x, int y, int log2n, int comp, uint8_t mode, bool cbf, bool act = false) {
        hpvcd_tb tb;
        std::memset(&tb, 0, sizeof(tb));
        tb.x = (uint16_t)x;
        tb.y = (uint16_t)y;
        tb.log2n = (uint8_t)log2n;
        tb.comp = (uint8_t)comp;
        tb.mode = mode;
        tb.qp = (uint8_t)tb_qp(comp, act);
        if (act) tb.flags |= HPVCD_TB_ACT;
        if (cu_tqb) tb.flags |= HPVCD_TB_TQB;
        if (cu_inter) tb.flags |= HPVCD_TB_INTER;
        if (cbf) {
            tb.flags |= HPVCD_TB_CBF;
            if (cfg.transform_skip && !cu_tqb && log2n <= cfg.log2_max_ts && rng.bern(cfg.p_ts)) tb.flags |= HPVCD_TB_TS;
            // explicit_rdpcm_flag / explicit_rdpcm_dir_flag of an inter TB with
            // transform skip or bypass (RExt, src/decode.rs:4356)
            if (cu_inter && cfg.implicit_rdpcm && (tb.flags & (HPVCD_TB_TS | HPVCD_TB_TQB))) {
                uint32_t r = rng.uni(3);
                if (r == 1) tb.flags |= HPVCD_TB_RDPCM_H;
                else if (r == 2) tb.flags |= HPVCD_TB_RDPCM_V;
            }
            emit_levels(tb);
        }
        return tb;
    }

    bool chroma_dm_at(int x0, int y0) {
        int col = (x0 - cu_x0) >= cu_pu, row = (y0 - cu_y0) >= cu_pu;
        return cu_chroma_dm[row * 2 + col];
    }
    uint8_t chroma_mode_at(int x0, int y0) {
        int col = (x0 - cu_x0) >= cu_pu, row = (y0 - cu_y0) >= cu_pu;
        return cu_chroma_modes[row * 2 + col];
    }

    int draw_res_scale() {
        int m = cfg.log2_res_scale_max;
        int v = rng.sint(-m, m);
        if (v == 0) return 0;
        int mag = 1 << (std::abs(v) - 1);
        return v < 0 ? -mag : mag;
    }

    void do_chroma(int lx, int ly, int luma_log2, bool cbf_cb, bool cbf_cr, bool cross_comp) {
        uint8_t mode = cu_inter ? MODE_DC : chroma_mode_at(lx, ly);
        int clog2 = cfg.chroma_idc == 3 ? luma_log2 : luma_log2 - 1;
        int cx0 = lx / sub_w, cy0 = ly / sub_h;
        for (int comp = 1; comp <= 2; comp++) {
            int scale = cross_comp ? draw_res_scale() : 0;
            bool cbf = comp == 1 ? cbf_cb : cbf_cr;
            hpvcd_tb tb = make_tb(cx0, cy0, clog2, comp, mode, cbf);
            tb.res_scale = (int8_t)scale;
            tbs.push_back(tb);
        }
    }

    // A co-located luma/Cb/Cr TB triple whose residuals go through the
    // inverse colour transform (src/decode.rs:3619 transform_unit_act).
    void act_unit(int x0, int y0, int log2, uint8_t luma_mode, bool cbf_luma, bool cbf_cb, bool cbf_cr) {
        bool cross_comp = cfg.cross_component && cbf_luma;
        tbs.push_back(make_tb(x0, y0, log2, 0, luma_mode, cbf_luma, true));
        uint8_t mode = cu_inter ? MODE_DC : chroma_mode_at(x0, y0);
        for (int comp = 1; comp <= 2; comp++) {
            int scale = cross_comp ? draw_res_scale() : 0;
            hpvcd_tb tb = make_tb(x0, y0, log2, comp, mode, comp == 1 ? cbf_cb : cbf_cr, true);
            tb.res_scale = (int8_t)scale;
            tbs.push_back(tb);
        }
    }

    void transform_unit(int x0, int y0, int xbase, int ybase, int log2, int blk, bool cbf_cb, bool cbf_cr) {
        bool cbf_luma = rng.bern(cfg.p_cbf);
        int size = 1 << log2;
        if (cu_inter) {
            mark_tu_edges(x0, y0, size);
            if (cbf_luma)
                for (int yy = y0; yy < std::min(y0 + size, h); yy += 4)
                    for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) nz[(yy / 4) * gw + xx / 4] = 1;
        } else {
            mark_block_edges(x0, y0, size, 2);
        }
        uint8_t luma_mode = cu_inter ? MODE_DC : mode_y[(y0 / 4) * gw + x0 / 4];
        maybe_qp_delta(cbf_luma || cbf_cb || cbf_cr);
        // tu_residual_act_flag is present for a TU with residual of an inter
        // CU, or of an intra CU whose chroma modes are all the derived mode
        // (src/decode.rs:4229)
        bool dm_all = cu_chroma_dm[0] && cu_chroma_dm[1] && cu_chroma_dm[2] && cu_chroma_dm[3];
        if (cfg.act && cfg.chroma_idc == 3 && (cbf_luma || cbf_cb || cbf_cr) && (cu_inter || dm_all) &&
            rng.bern(cfg.p_act)) {
            act_unit(x0, y0, log2, luma_mode, cbf_luma, cbf_cb, cbf_cr);
            return;
        }
        bool cross_comp = cfg.cross_component && cbf_luma && (cu_inter || chroma_dm_at(x0, y0));
        tbs.push_back(make_tb(x0, y0, log2, 0, luma_mode, cbf_luma));
        if (cfg.chroma_idc == 0) return;
        if (log2 > 2 || cfg.chroma_idc == 3) do_chroma(x0, y0, log2, cbf_cb, cbf_cr, cross_comp);
        else if (blk == 3) do_chroma(xbase, ybase, 3, cbf_cb, cbf_cr, false);
    }

    void transform_tree(int x0, int y0, int xbase, int ybase, int log2, int depth, int blk, bool intra_split,
                        int max_depth, bool pcb, bool pcr) {
        bool allowed = log2 <= log2_max_tb && log2 > log2_min_tb && depth < max_depth && !(intra_split && depth == 0);
        bool split = allowed ? rng.bern(cfg.p_split_tu) : (log2 > log2_max_tb || (intra_split && depth == 0));
        bool cbf_cb = pcb, cbf_cr = pcr;
        if (cfg.chroma_idc && (log2 > 2 || cfg.chroma_idc == 3)) {
            if (depth == 0 || pcb) cbf_cb = rng.bern(cfg.p_cbf);
            if (depth == 0 || pcr) cbf_cr = rng.bern(cfg.p_cbf);
        }
        if (split) {
            int half = 1 << (log2 - 1);
            transform_tree(x0, y0, x0, y0, log2 - 1, depth + 1, 0, intra_split, max_depth, cbf_cb, cbf_cr);
            transform_tree(x0 + half, y0, x0, y0, log2 - 1, depth + 1, 1, intra_split, max_depth, cbf_cb, cbf_cr);
            transform_tree(x0, y0 + half, x0, y0, log2 - 1, depth + 1, 2, intra_split, max_depth, cbf_cb, cbf_cr);
            transform_tree(x0 + half, y0 + half, x0, y0, log2 - 1, depth + 1, 3, intra_split, max_depth, cbf_cb, cbf_cr);
        } else {
            transform_unit(x0, y0, xbase, ybase, log2, blk, cbf_cb, cbf_cr);
        }
    }

    // I_PCM CU (src/decode.rs:2768 decode_pcm_cu): luma then Cb then Cr,
    // every sample a raw PcmBitDepth-bit value, scaled by the TB's shift.
    void pcm_unit(int x0, int y0, int log2) {
        int size = 1 << log2;
        int shift = bd - cfg.pcm_bit_depth;
        uint32_t vmax = 1u << cfg.pcm_bit_depth;
        fill_cells(x0, y0, size, HPVCD_CELL_PCM);
        set_mode(x0, y0, size, MODE_DC);
        mark_block_edges(x0, y0, size, 2);
        int ncomp = cfg.chroma_idc ? 3 : 1;
        for (int comp = 0; comp < ncomp; comp++) {
            int l2 = comp == 0 || cfg.chroma_idc == 3 ? log2 : log2 - 1;
            int n = 1 << l2;
            hpvcd_tb tb;
            std::memset(&tb, 0, sizeof(tb));
            tb.x = (uint16_t)(comp ? x0 / sub_w : x0);
            tb.y = (uint16_t)(comp ? y0 / sub_h : y0);
            tb.log2n = (uint8_t)l2;
            tb.comp = (uint8_t)comp;
            tb.mode = MODE_DC;
            tb.qp = (uint8_t)shift;
            tb.flags = HPVCD_TB_PCM | HPVCD_TB_CBF;
            tb.coeff_off = (uint32_t)coeffs.size() - coeff_base;
            for (int pos = 0; pos < n * n; pos++) coeffs.push_back(HPVCD_COEFF_PACK(pos, (int)rng.uni(vmax)));
            tb.coeff_cnt = (uint32_t)(n * n);
            tb.nx = (uint8_t)n;
            tbs.push_back(tb);
        }
        for (int yy = y0; yy < std::min(y0 + size, h); yy += 4)
            for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) cell_qp[(yy / 4) * gw + xx / 4] = (int8_t)qp_y;
        qp_prev = qp_y;
    }

    void coding_unit(int x0, int y0, int log2) {
        int size = 1 << log2;
        const int qgm = ~((1 << log2_qg) - 1);
        if ((x0 & qgm) == x0 && (y0 & qgm) == y0) qp_y = predict_qp(x0, y0);
        cu_tqb = cfg.p_tqb > 0.f && rng.bern(cfg.p_tqb);
        if (cu_tqb) fill_cells(x0, y0, size, HPVCD_CELL_TQB);
        // cu_skip_flag / pred_mode_flag of an inter picture
        if (inter_pic && !rng.bern(cfg.p_intra_cu)) {
            inter_cu(x0, y0, log2);
            return;
        }
        bool nxn = log2 == log2_min_cb && log2 > 2 && rng.bern(cfg.p_nxn);
        // pcm_flag follows part_mode for 2Nx2N CUs of 8x8..32x32 (src/decode.rs:2647)
        if (cfg.pcm && !nxn && log2 >= 3 && log2 <= 5 && rng.bern(cfg.p_pcm)) {
            pcm_unit(x0, y0, log2);
            return;
        }
        int npu = nxn ? 2 : 1, pu = size / npu;
        uint8_t luma_modes[4] = {MODE_DC, MODE_DC, MODE_DC, MODE_DC};
        for (int i = 0; i < npu * npu; i++) {
            luma_modes[i] = (uint8_t)rng.uni(35);
            set_mode(x0 + (i % npu) * pu, y0 + (i / npu) * pu, pu, luma_modes[i]);
        }
        int n_chroma = (cfg.chroma_idc == 3 && nxn) ? 4 : 1;
        for (int i = 0; i < 4; i++) { cu_chroma_modes[i] = MODE_DC; cu_chroma_dm[i] = false; }
        if (cfg.chroma_idc) {
            for (int i = 0; i < n_chroma; i++) {
                uint32_t syn = rng.uni(5);
                cu_chroma_modes[i] = derive_chroma_mode(syn, luma_modes[i]);
                cu_chroma_dm[i] = syn == 4;
            }
            if (n_chroma == 1)
                for (int i = 1; i < 4; i++) { cu_chroma_modes[i] = cu_chroma_modes[0]; cu_chroma_dm[i] = cu_chroma_dm[0]; }
        }
        cu_x0 = x0; cu_y0 = y0; cu_pu = std::max(pu, 1);
        int max_depth = cfg.max_transform_depth_intra + (nxn ? 1 : 0);
        transform_tree(x0, y0, x0, y0, log2, 0, 0, nxn, max_depth, false, false);
        if (nxn) {
            int half = size / 2;
            mark_block_edges(x0 + half, y0, half, 2);
            mark_block_edges(x0, y0 + half, half, 2);
        }
        for (int yy = y0; yy < std::min(y0 + size, h); yy += 4)
            for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) cell_qp[(yy / 4) * gw + xx / 4] = (int8_t)qp_y;
        qp_prev = qp_y;
    }

    void coding_quadtree(int x0, int y0, int log2) {
        int size = 1 << log2;
        if (x0 >= w || y0 >= h) return;
        bool split;
        if (x0 + size > w || y0 + size > h) split = log2 > log2_min_cb;
        else split = log2 > log2_min_cb && rng.bern(cfg.p_split_cu);
        if (cfg.cu_qp_delta && log2 >= log2_qg) qp_delta_coded = false;
        if (split) {
            int half = size / 2;
            coding_quadtree(x0, y0, log2 - 1);
            coding_quadtree(x0 + half, y0, log2 - 1);
            coding_quadtree(x0, y0 + half, log2 - 1);
            coding_quadtree(x0 + half, y0 + half, log2 - 1);
        } else {
            coding_unit(x0, y0, log2);
        }
    }

    // SAO parameters of CTB (rx, ry): merged from the left or above CTB
    // where that CTB is in the same slice and tile (src/decode.rs:1507
    // parse_sao), else drawn
    void sao_params(std::vector<hpvcd_sao>& all, int rx, int ry, int cols) {
        hpvcd_sao& s = all[(size_t)ry * cols + rx];
        std::memset(&s, 0, sizeof(s));
        if (!cfg.sao) return;
        const size_t me = (size_t)ry * cols + rx;
        auto avail = [&](size_t n) { return ctb_slice[n] == ctb_slice[me] && ctb_tile[n] == ctb_tile[me]; };
        if (cfg.p_sao_merge > 0.f) {
            if (rx > 0 && avail(me - 1) && rng.bern(cfg.p_sao_merge)) { s = all[me - 1]; return; }
            if (ry > 0 && avail(me - cols) && rng.bern(cfg.p_sao_merge)) { s = all[me - cols]; return; }
        }
        int ncomp = cfg.chroma_idc ? 3 : 1;
        for (int c = 0; c < ncomp; c++) {
            int type = c == 2 ? s.type_idx[1] : (int)rng.uni(3);
            s.type_idx[c] = (uint8_t)type;
            if (!type) continue;
            int v[4];
            for (int k = 0; k < 4; k++) v[k] = (int)rng.uni(8); /* |offset| in [0,7] */
            if (type == 1) {
                for (int k = 0; k < 4; k++) if (v[k] && (rng.next() & 1)) v[k] = -v[k];
                s.band_pos[c] = (uint8_t)rng.uni(32);
            } else {
                v[2] = -v[2];
                v[3] = -v[3];
                s.eo_class[c] = c == 2 ? s.eo_class[1] : (uint8_t)rng.uni(4);
            }
            for (int k = 0; k < 4; k++) s.offsets[c][k] = (int16_t)v[k];
        }
    }
};

struct TileOut {
    std::vector<hpvcd_tb> tbs;
    std::vector<uint32_t> coeffs, cells;
    std::vector<hpvcd_sao> sao;
    std::vector<hpvcd_pu> pus;
    std::vector<hpvcd_ref> refs;
    int num_slices = 1, num_tiles = 1;
};

// Uniformly spaced tile boundaries in CTBs (uniform_spacing_flag, H.265
// eq. 6-3/6-4) and the tile-scan order of the CTBs (CtbAddrTsToRs).
struct TileLayout {
    std::vector<int> col_bd, row_bd;
    std::vector<uint32_t> ts_to_rs;
    std::vector<uint8_t> tile_of;   // per raster CTB
    TileLayout(int cols, int rows, int ntc, int ntr) {
        ntc = std::max(1, std::min(ntc, cols));
        ntr = std::max(1, std::min(ntr, rows));
        for (int i = 0; i <= ntc; i++) col_bd.push_back(i * cols / ntc);
        for (int j = 0; j <= ntr; j++) row_bd.push_back(j * rows / ntr);
        tile_of.assign((size_t)cols * rows, 0);
        for (int tj = 0; tj < ntr; tj++)
            for (int ti = 0; ti < ntc; ti++)
                for (int ry = row_bd[tj]; ry < row_bd[tj + 1]; ry++)
                    for (int rx = col_bd[ti]; rx < col_bd[ti + 1]; rx++) {
                        ts_to_rs.push_back((uint32_t)(ry * cols + rx));
                        tile_of[(size_t)ry * cols + rx] = (uint8_t)(tj * ntc + ti);
                    }
    }
    int num_tiles() const { return (int)(col_bd.size() - 1) * (int)(row_bd.size() - 1); }
};

void gen_tile(const hpvcd_gen_config& cfg, uint32_t tile, TileOut& out) {
    PicGen g(cfg, tile_seed(cfg.base_seed, cfg.config_id, tile), out.tbs, out.coeffs, out.pus, out.refs, tile);
    int ctb = 1 << cfg.log2_ctb;
    int cols = (g.w + ctb - 1) / ctb, rows = (g.h + ctb - 1) / ctb;
    const int nctb = cols * rows;
    out.sao.resize((size_t)nctb);
    // CTBs in tile-scan order; slices are equal runs of that order
    TileLayout T(cols, rows, std::max<int>(1, cfg.tile_cols), std::max<int>(1, cfg.tile_rows));
    const int nslices = std::max(1, std::min<int>(std::max<int>(1, cfg.num_slices), std::min(nctb, HPVCD_MAX_SLICES)));
    g.ctb_tile = T.tile_of;
    g.ctb_slice.assign((size_t)nctb, 0);
    for (int ts = 0; ts < nctb; ts++) g.ctb_slice[T.ts_to_rs[ts]] = (uint8_t)((int64_t)ts * nslices / nctb);
    out.num_slices = nslices;
    out.num_tiles = T.num_tiles();
    for (int ts = 0; ts < nctb; ts++) {
        const uint32_t rs = T.ts_to_rs[ts];
        const int rx = (int)(rs % cols), ry = (int)(rs / cols);
        const bool slice_start = ts == 0 || g.ctb_slice[rs] != g.ctb_slice[T.ts_to_rs[ts - 1]];
        const bool tile_start = ts == 0 || g.ctb_tile[rs] != g.ctb_tile[T.ts_to_rs[ts - 1]];
        if (slice_start || tile_start) g.qp_prev = g.slice_qp;  // src/decode.rs:776, :1262
        if (slice_start) g.qp_y = g.slice_qp;
        g.sao_params(out.sao, rx, ry, cols);
        g.coding_quadtree(rx * ctb, ry * ctb, cfg.log2_ctb);
    }
    if (g.inter_pic) g.finalize_bs();
    out.cells.resize(g.cell.size());
    for (size_t i = 0; i < g.cell.size(); i++) {
        const int gx = (int)(i % (size_t)g.gw), gy = (int)(i / (size_t)g.gw);
        const size_t c = (size_t)((gy * 4) >> cfg.log2_ctb) * cols + ((gx * 4) >> cfg.log2_ctb);
        out.cells[i] = (g.cell[i] & 0xffu) | ((uint32_t)(uint8_t)g.cell_qp[i] << 8) |
                       ((uint32_t)g.ctb_slice[c] << 16) | ((uint32_t)g.ctb_tile[c] << 24);
    }
    // records list TBs in CTB raster order (decode order inside each CTB)
    if (T.num_tiles() > 1) {
        const int sw = g.sub_w, sh = g.sub_h;
        auto ctb_of = [&](const hpvcd_tb& t) {
            const int lx = t.comp ? t.x * sw : t.x, ly = t.comp ? t.y * sh : t.y;
            return (ly >> cfg.log2_ctb) * cols + (lx >> cfg.log2_ctb);
        };
        std::stable_sort(out.tbs.begin(), out.tbs.end(),
                         [&](const hpvcd_tb& a, const hpvcd_tb& b) { return ctb_of(a) < ctb_of(b); });
    }
}

// Default lists (H.265 Tables 7-5 and 7-6, raster order as
// src/config.rs:358-370 stores them).
const uint8_t kDefaultIntra8x8[64] = {
    16, 16, 16, 16, 17, 18, 21, 24, 16, 16, 16, 16, 17, 19, 22, 25, 16, 16, 17, 18, 20, 22, 25, 29,
    16, 16, 18, 21, 24, 27, 31, 36, 17, 17, 20, 24, 30, 35, 41, 47, 18, 19, 22, 27, 35, 44, 54, 65,
    21, 22, 25, 31, 41, 54, 70, 88, 24, 25, 29, 36, 47, 65, 88, 115};
const uint8_t kDefaultInter8x8[64] = {
    16, 16, 16, 16, 17, 18, 20, 24, 16, 16, 16, 17, 18, 20, 24, 25, 16, 16, 17, 18, 20, 24, 25, 28,
    16, 17, 18, 20, 24, 25, 28, 33, 17, 18, 20, 24, 25, 28, 33, 41, 18, 20, 24, 25, 28, 33, 41, 54,
    20, 24, 25, 28, 33, 41, 54, 71, 24, 25, 28, 33, 41, 54, 71, 91};

// One record scaling set (sizeId 0..3 x matrixId 0..5 entries of 64 raster
// coefficients, DC, flat flag): the default lists, or random lists of the
// kind scaling_list_data codes explicitly.  sizeId 3 signals matrixId 0 and
// 3 only; its chroma entries repeat sizeId 2's (src/config.rs:558-563).  The
// flat flag follows ScalingList::refresh_stats (src/config.rs:409).
void make_scaling_set(const hpvcd_gen_config& c, uint8_t* set) {
    Rng rng{tile_seed(c.base_seed ^ 0x5ca1eull, c.config_id, 0)};
    auto entry = [&](int size_id, int m) { return set + (size_t)(size_id * 6 + m) * HPVCD_SCALING_ENTRY_BYTES; };
    for (int size_id = 0; size_id < 4; size_id++)
        for (int m = 0; m < 6; m++) {
            uint8_t* e = entry(size_id, m);
            std::memset(e, 16, 64);
            e[64] = 16;
            if (size_id == 3 && m % 3) {
                std::memcpy(e, entry(2, m), HPVCD_SCALING_ENTRY_BYTES);
                continue;
            }
            const int used = size_id == 0 ? 16 : 64;
            if (c.scaling_list == 1) {
                if (size_id > 0) std::memcpy(e, m < 3 ? kDefaultIntra8x8 : kDefaultInter8x8, 64);
            } else {
                for (int k = 0; k < used; k++) e[k] = (uint8_t)rng.sint(4, 80);
                if (size_id >= 2) e[64] = (uint8_t)rng.sint(4, 80);
            }
            bool flat = size_id < 2 || e[64] == 16;
            for (int k = 0; k < used; k++) flat &= e[k] == 16;
            e[65] = flat ? 1 : 0;
        }
}

}  // namespace
}  // namespace hpvcd

using namespace hpvcd;

static_assert(sizeof(hpvcd_gen_config) == 168, "hpvcd_gen_config layout");
static_assert(sizeof(hpvcd_picture) == 140, "hpvcd_picture layout");
static_assert(sizeof(hpvcd_tb) == 24 && sizeof(hpvcd_sao) == 36 && sizeof(hpvcd_output) == 12, "record layout");
static_assert(sizeof(hpvcd_pu) == 20 && sizeof(hpvcd_ref) == 20, "inter record layout");

extern "C" void hpvcd_gen_default_config(hpvcd_gen_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->tile_width = 512;
    c->tile_height = 512;
    c->grid_cols = 1;
    c->grid_rows = 1;
    c->out_width = 512;
    c->out_height = 512;
    c->chroma_idc = 1;
    c->bit_depth = 8;
    c->log2_ctb = 6;
    c->log2_min_cb = 3;
    c->log2_min_tb = 2;
    c->log2_max_tb = 5;
    c->max_transform_depth_intra = 4;
    c->qp = 32;
    c->log2_max_ts = 2;
    c->strong_intra_smoothing = 1;
    c->sao = 1;
    c->deblock = 1;
    c->matrix = 1;
    c->full_range = 0;
    c->out_kind = HPVCD_OUT_RGB8;
    c->num_outputs = 1;
    c->log2_res_scale_max = 3;
    c->p_split_cu = 0.5f;
    c->p_nxn = 0.5f;
    c->p_split_tu = 0.4f;
    c->p_cbf = 0.6f;
    c->p_sig = 0.3f;
    c->p_level_geom = 0.5f;
    c->p_ts = 0.2f;
    c->p_tqb = 0.0f;
    c->num_refs = 4;
    c->amp = 1;
    c->p_intra_cu = 0.1f;
    c->p_skip = 0.3f;
    c->p_bi = 0.5f;
    c->mv_sigma = 16.0f;
    c->config_id = 0;
    c->base_seed = 0x48505643ull;
    c->qp_delta_max = 4;
    c->lf_across_slices = 1;
    c->lf_across_tiles = 1;
    c->num_slices = 1;
    c->tile_cols = 1;
    c->tile_rows = 1;
}

extern "C" hpvcd_status hpvcd_gen_records(const hpvcd_gen_config* cfg, hpvcd_records** out) {
    if (!cfg || !out) return set_error(HPVCD_E_INVALID, "null argument");
    *out = nullptr;
    const hpvcd_gen_config& c = *cfg;
    if (c.tile_width <= 0 || c.tile_height <= 0 || c.grid_cols <= 0 || c.grid_rows <= 0 ||
        c.tile_width % (1 << c.log2_min_cb) || c.tile_height % (1 << c.log2_min_cb) || c.tile_width > 8192 ||
        c.tile_height > 8192)
        return set_error(HPVCD_E_BAD_DIMENSIONS, "generator: tile size must be a positive multiple of the min CU");
    if (c.chroma_idc == 2) return set_error(HPVCD_E_UNSUPPORTED_CHROMA, "generator: 4:2:2 not generated");
    if (c.bit_depth < 8 || c.bit_depth > 12) return set_error(HPVCD_E_UNSUPPORTED_BITDEPTH, "generator: bit depth");
    if (c.pcm && (c.pcm_bit_depth < 1 || c.pcm_bit_depth > c.bit_depth))
        return set_error(HPVCD_E_PARAMSET, "generator: PCM bit depth must be 1..bit_depth");
    if (c.log2_ctb < 4 || c.log2_ctb > 6 || c.log2_min_cb < 3 || c.log2_min_cb > c.log2_ctb || c.log2_min_tb != 2 ||
        c.log2_max_tb > 5 || c.log2_max_tb < 3)
        return set_error(HPVCD_E_PARAMSET, "generator: unsupported block size limits");
    if (c.conf_win[0] + c.conf_win[1] >= c.tile_width || c.conf_win[2] + c.conf_win[3] >= c.tile_height)
        return set_error(HPVCD_E_BAD_DIMENSIONS, "generator: conformance window leaves no picture");
    {
        const int cw = (c.chroma_idc == 1 || c.chroma_idc == 2) ? 2 : 1, ch = c.chroma_idc == 1 ? 2 : 1;
        if (c.conf_win[0] % cw || c.conf_win[1] % cw || c.conf_win[2] % ch || c.conf_win[3] % ch)
            return set_error(HPVCD_E_PARAMSET, "generator: conformance window offsets in chroma units");
    }
    if (c.orientation > HPVCD_ORIENT_ROTATE_270) return set_error(HPVCD_E_INVALID, "generator: orientation");
    if (c.pps_cb_qp_offset < -12 || c.pps_cb_qp_offset > 12 || c.pps_cr_qp_offset < -12 || c.pps_cr_qp_offset > 12 ||
        c.pps_cb_qp_offset + c.slice_cb_qp_offset < -12 || c.pps_cb_qp_offset + c.slice_cb_qp_offset > 12 ||
        c.pps_cr_qp_offset + c.slice_cr_qp_offset < -12 || c.pps_cr_qp_offset + c.slice_cr_qp_offset > 12)
        return set_error(HPVCD_E_PARAMSET, "generator: chroma QP offsets outside [-12, 12]");
    if (c.beta_offset_div2 < -6 || c.beta_offset_div2 > 6 || c.tc_offset_div2 < -6 || c.tc_offset_div2 > 6)
        return set_error(HPVCD_E_PARAMSET, "generator: deblocking offsets outside [-6, 6]");
    if (c.cu_qp_delta && (c.diff_cu_qp_delta_depth > c.log2_ctb - c.log2_min_cb || c.qp_delta_max > 26))
        return set_error(HPVCD_E_PARAMSET, "generator: cu_qp_delta depth or range");
    uint32_t ntiles = (uint32_t)(c.grid_cols * c.grid_rows);
    std::vector<TileOut> tiles(ntiles);
    unsigned nth = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; t++)
        th.emplace_back([&, t]() {
            for (uint32_t i = t; i < ntiles; i += nth) gen_tile(c, i, tiles[i]);
        });
    for (auto& x : th) x.join();

    hpvcd_records* r = new (std::nothrow) hpvcd_records();
    if (!r) return set_error(HPVCD_E_ALLOC, "generator: out of memory");
    if (c.scaling_list) {
        r->scaling.resize(HPVCD_SCALING_SET_BYTES);
        make_scaling_set(c, r->scaling.data());
    }
    const int win_w = c.tile_width - c.conf_win[0] - c.conf_win[1];
    const int win_h = c.tile_height - c.conf_win[2] - c.conf_win[3];
    int nout = std::max<int>(1, c.num_outputs);
    for (int o = 0; o < nout; o++) {
        hpvcd_output ho;
        std::memset(&ho, 0, sizeof(ho));
        ho.width = c.out_width;
        ho.height = c.out_height;
        ho.kind = c.out_kind;
        ho.channels = (c.out_kind == HPVCD_OUT_LUMA8 || c.out_kind == HPVCD_OUT_LUMA16) ? 1 : 3;
        ho.bytes_per_sample = (c.out_kind == HPVCD_OUT_RGB16 || c.out_kind == HPVCD_OUT_LUMA16) ? 2 : 1;
        ho.orientation = c.orientation;
        r->outputs.push_back(ho);
    }
    for (uint32_t i = 0; i < ntiles; i++) {
        TileOut& t = tiles[i];
        hpvcd_picture p;
        std::memset(&p, 0, sizeof(p));
        p.width = c.tile_width;
        p.height = c.tile_height;
        p.chroma_idc = c.chroma_idc;
        p.bit_depth_luma = c.bit_depth;
        p.bit_depth_chroma = c.bit_depth;
        p.log2_ctb = c.log2_ctb;
        uint32_t f = c.lf_across_slices ? HPVCD_PF_LF_ACROSS_SLICES : 0u;
        if (c.strong_intra_smoothing) f |= HPVCD_PF_STRONG_INTRA_SMOOTHING;
        if (c.extended_precision) f |= HPVCD_PF_EXTENDED_PRECISION;
        if (c.implicit_rdpcm) f |= HPVCD_PF_IMPLICIT_RDPCM;
        if (c.ts_rotation) f |= HPVCD_PF_TS_ROTATION;
        if (c.sao) f |= HPVCD_PF_SAO_LUMA | (c.chroma_idc ? HPVCD_PF_SAO_CHROMA : 0u);
        if (c.full_range) f |= HPVCD_PF_FULL_RANGE;
        if (!c.deblock) f |= HPVCD_PF_DEBLOCK_DISABLED;
        if (c.pcm && c.pcm_lf_disabled) f |= HPVCD_PF_PCM_LF_DISABLED;
        if (t.num_tiles > 1) f |= HPVCD_PF_TILES | (c.lf_across_tiles ? HPVCD_PF_LF_ACROSS_TILES : 0u);
        p.flags = f;
        p.matrix = c.matrix;
        p.pps_cb_qp_offset = c.pps_cb_qp_offset;
        p.pps_cr_qp_offset = c.pps_cr_qp_offset;
        // slice deblocking parameters: the configured offsets, or per slice a
        // draw of disable, offsets and the across-slices flag
        p.num_slices = (uint8_t)t.num_slices;
        Rng prng{tile_seed(c.base_seed ^ 0x5b1cull, c.config_id, i)};
        for (int sl = 0; sl < t.num_slices; sl++) {
            hpvcd_slice_params& S = p.slices[sl];
            S.disabled = c.deblock ? 0 : 1;
            S.beta_offset_div2 = c.beta_offset_div2;
            S.tc_offset_div2 = c.tc_offset_div2;
            S.lf_across = c.lf_across_slices ? 1 : 0;
            if (c.random_slice_params) {
                if (c.deblock) S.disabled = prng.bern(0.2f) ? 1 : 0;
                S.beta_offset_div2 = (int8_t)prng.sint(-6, 6);
                S.tc_offset_div2 = (int8_t)prng.sint(-6, 6);
                if (c.lf_across_slices) S.lf_across = prng.bern(0.5f) ? 1 : 0;
            }
        }
        p.slice_qp = c.qp;
        p.out_kind = c.out_kind;
        uint32_t col = i % (uint32_t)c.grid_cols, row = i / (uint32_t)c.grid_cols;
        // conformance window: the visible tile, placed at (col, row) * window
        // size in a grid (src/lib.rs:143 plane_window, :1754-1793)
        p.crop_left = c.conf_win[0];
        p.crop_top = c.conf_win[2];
        p.win_w = win_w;
        p.win_h = win_h;
        // one output per picture (frames, independent images) or one grid
        p.dst_x = nout > 1 ? 0 : (int32_t)col * win_w;
        p.dst_y = nout > 1 ? 0 : (int32_t)row * win_h;
        p.copy_w = std::max(0, std::min(win_w, c.out_width - p.dst_x));
        p.copy_h = std::max(0, std::min(win_h, c.out_height - p.dst_y));
        p.out_index = nout > 1 ? i % (uint32_t)nout : 0u;
        p.scaling_index = c.scaling_list ? 0 : -1;
        p.tb_begin = (uint32_t)r->tbs.size();
        p.tb_count = (uint32_t)t.tbs.size();
        p.coeff_begin = (uint32_t)r->coeffs.size();
        p.coeff_count = (uint32_t)t.coeffs.size();
        p.ctb_begin = (uint32_t)r->sao.size();
        p.cell_begin = (uint32_t)r->cells.size();
        p.pu_begin = (uint32_t)r->pus.size();
        p.pu_count = (uint32_t)t.pus.size();
        p.ref_begin = (uint32_t)r->refs.size();
        p.ref_count = (uint32_t)t.refs.size();
        r->pictures.push_back(p);
        r->pus.insert(r->pus.end(), t.pus.begin(), t.pus.end());
        r->refs.insert(r->refs.end(), t.refs.begin(), t.refs.end());
        r->tbs.insert(r->tbs.end(), t.tbs.begin(), t.tbs.end());
        r->coeffs.insert(r->coeffs.end(), t.coeffs.begin(), t.coeffs.end());
        r->sao.insert(r->sao.end(), t.sao.begin(), t.sao.end());
        r->cells.insert(r->cells.end(), t.cells.begin(), t.cells.end());
        std::vector<hpvcd_tb>().swap(t.tbs);
        std::vector<uint32_t>().swap(t.coeffs);
    }
    *out = r;
    return HPVCD_OK;
}
