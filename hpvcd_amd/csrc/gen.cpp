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
#include <algorithm>
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
    std::vector<uint32_t> cell;      // info word without QP (QP packed at the end)
    std::vector<int8_t> cell_qp;
    std::vector<uint8_t> mode_y;     // luma intra mode per 4x4 (src/decode.rs mode_y)
    uint32_t coeff_base;
    // current CU
    uint8_t cu_chroma_modes[4];
    bool cu_chroma_dm[4];
    int cu_x0, cu_y0, cu_pu;
    bool cu_tqb;

    PicGen(const hpvcd_gen_config& c, uint64_t seed, std::vector<hpvcd_tb>& t, std::vector<uint32_t>& co)
        : cfg(c), rng{seed}, tbs(t), coeffs(co) {
        w = c.tile_width;
        h = c.tile_height;
        gw = (w + 3) / 4;
        gh = (h + 3) / 4;
        sub_w = (c.chroma_idc == 1 || c.chroma_idc == 2) ? 2 : 1;
        sub_h = c.chroma_idc == 1 ? 2 : 1;
        log2_min_cb = c.log2_min_cb;
        log2_min_tb = c.log2_min_tb;
        log2_max_tb = c.log2_max_tb;
        qp_y = c.qp;
        bd = c.bit_depth;
        cell.assign((size_t)gw * gh, HPVCD_CELL_INTRA);
        cell_qp.assign((size_t)gw * gh, (int8_t)qp_y);
        mode_y.assign((size_t)gw * gh, MODE_DC);
        coeff_base = (uint32_t)co.size();
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

    // Qp' of a TB.  Under ACT the component offsets -5/-5/-3 apply to QpY
    // before the chroma mapping and no PPS chroma offsets (src/decode.rs:3788).
    int tb_qp(int comp, bool act) {
        if (!act) return comp == 0 ? qp_prime(qp_y, bd) : qp_prime(qpc_table(qp_y, cfg.chroma_idc, bd), bd);
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
        if (cbf) {
            tb.flags |= HPVCD_TB_CBF;
            if (cfg.transform_skip && !cu_tqb && log2n <= cfg.log2_max_ts && rng.bern(cfg.p_ts)) tb.flags |= HPVCD_TB_TS;
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
        uint8_t mode = chroma_mode_at(lx, ly);
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
        uint8_t mode = chroma_mode_at(x0, y0);
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
        mark_block_edges(x0, y0, size, 2);
        uint8_t luma_mode = mode_y[(y0 / 4) * gw + x0 / 4];
        // tu_residual_act_flag is present for a TU with residual when every
        // chroma mode of the CU is the derived mode (src/decode.rs:4229)
        if (cfg.act && cfg.chroma_idc == 3 && (cbf_luma || cbf_cb || cbf_cr) && cu_chroma_dm[0] && cu_chroma_dm[1] &&
            cu_chroma_dm[2] && cu_chroma_dm[3] && rng.bern(cfg.p_act)) {
            act_unit(x0, y0, log2, luma_mode, cbf_luma, cbf_cb, cbf_cr);
            return;
        }
        bool cross_comp = cfg.cross_component && cbf_luma && chroma_dm_at(x0, y0);
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
    }

    void coding_unit(int x0, int y0, int log2) {
        int size = 1 << log2;
        cu_tqb = cfg.p_tqb > 0.f && rng.bern(cfg.p_tqb);
        if (cu_tqb) fill_cells(x0, y0, size, HPVCD_CELL_TQB);
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
    }

    void coding_quadtree(int x0, int y0, int log2) {
        int size = 1 << log2;
        if (x0 >= w || y0 >= h) return;
        bool split;
        if (x0 + size > w || y0 + size > h) split = log2 > log2_min_cb;
        else split = log2 > log2_min_cb && rng.bern(cfg.p_split_cu);
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

    void sao_params(hpvcd_sao& s) {
        std::memset(&s, 0, sizeof(s));
        if (!cfg.sao) return;
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
};

void gen_tile(const hpvcd_gen_config& cfg, uint32_t tile, TileOut& out) {
    PicGen g(cfg, tile_seed(cfg.base_seed, cfg.config_id, tile), out.tbs, out.coeffs);
    int ctb = 1 << cfg.log2_ctb;
    int cols = (g.w + ctb - 1) / ctb, rows = (g.h + ctb - 1) / ctb;
    out.sao.resize((size_t)cols * rows);
    for (int ry = 0; ry < rows; ry++)
        for (int rx = 0; rx < cols; rx++) {
            g.sao_params(out.sao[(size_t)ry * cols + rx]);
            g.coding_quadtree(rx * ctb, ry * ctb, cfg.log2_ctb);
        }
    out.cells.resize(g.cell.size());
    for (size_t i = 0; i < g.cell.size(); i++)
        out.cells[i] = (g.cell[i] & 0xffu) | ((uint32_t)(uint8_t)g.cell_qp[i] << 8);
}

}  // namespace
}  // namespace hpvcd

using namespace hpvcd;

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
    c->config_id = 0;
    c->base_seed = 0x48505643ull;
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
    int sub_w = (c.chroma_idc == 1 || c.chroma_idc == 2) ? 2 : 1;
    (void)sub_w;
    int nout = std::max<int>(1, c.num_outputs);
    for (int o = 0; o < nout; o++) {
        hpvcd_output ho;
        std::memset(&ho, 0, sizeof(ho));
        ho.width = c.out_width;
        ho.height = c.out_height;
        ho.kind = c.out_kind;
        ho.channels = c.chroma_idc ? 3 : 1;
        ho.bytes_per_sample = (c.out_kind == HPVCD_OUT_RGB16 || c.out_kind == HPVCD_OUT_LUMA16) ? 2 : 1;
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
        uint32_t f = HPVCD_PF_LF_ACROSS_SLICES;
        if (c.strong_intra_smoothing) f |= HPVCD_PF_STRONG_INTRA_SMOOTHING;
        if (c.extended_precision) f |= HPVCD_PF_EXTENDED_PRECISION;
        if (c.implicit_rdpcm) f |= HPVCD_PF_IMPLICIT_RDPCM;
        if (c.ts_rotation) f |= HPVCD_PF_TS_ROTATION;
        if (c.sao) f |= HPVCD_PF_SAO_LUMA | (c.chroma_idc ? HPVCD_PF_SAO_CHROMA : 0u);
        if (c.full_range) f |= HPVCD_PF_FULL_RANGE;
        if (!c.deblock) f |= HPVCD_PF_DEBLOCK_DISABLED;
        if (c.pcm && c.pcm_lf_disabled) f |= HPVCD_PF_PCM_LF_DISABLED;
        p.flags = f;
        p.matrix = c.matrix;
        p.num_slices = 1;
        p.slices[0].disabled = c.deblock ? 0 : 1;
        p.slices[0].lf_across = 1;
        p.slice_qp = c.qp;
        p.out_kind = c.out_kind;
        uint32_t col = i % (uint32_t)c.grid_cols, row = i / (uint32_t)c.grid_cols;
        p.crop_left = 0;
        p.crop_top = 0;
        p.win_w = c.tile_width;
        p.win_h = c.tile_height;
        p.dst_x = (int32_t)col * c.tile_width;
        p.dst_y = (int32_t)row * c.tile_height;
        p.copy_w = std::max(0, std::min(c.tile_width, c.out_width - p.dst_x));
        p.copy_h = std::max(0, std::min(c.tile_height, c.out_height - p.dst_y));
        p.out_index = (uint32_t)(i % (uint32_t)nout == 0 ? 0 : 0);
        if (nout > 1) {  // one output per tile (serving-style independent images)
            p.out_index = i % (uint32_t)nout;
        }
        p.scaling_index = -1;
        p.tb_begin = (uint32_t)r->tbs.size();
        p.tb_count = (uint32_t)t.tbs.size();
        p.coeff_begin = (uint32_t)r->coeffs.size();
        p.coeff_count = (uint32_t)t.coeffs.size();
        p.ctb_begin = (uint32_t)r->sao.size();
        p.cell_begin = (uint32_t)r->cells.size();
        r->pictures.push_back(p);
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
