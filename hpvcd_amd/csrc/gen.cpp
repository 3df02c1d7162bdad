// gen.cpp — seeded syntax-level generator of reconstruction records.
//
// It draws the coding tree, prediction modes, CBFs and coefficient levels of
// synthetic pictures from configurable distributions; the coding-tree walk
// that turns those decisions into records is the parser's (walker.cpp), so
// the records are the ones a parse of the same syntax yields.  Picture-level
// fields follow the reference decoder:
//   SAO offset signs and scaling               src/decode.rs:1507-1614 parse_sao
//   CCP ResScaleVal                            src/decode.rs:4257 (log2_res_scale_abs_plus1, sign)
//   grid placement / crop window               src/lib.rs:1610-1653, :1754-1817
//   scaling lists (sizeId 3 chroma, flat flag) src/config.rs:409, :558-563
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "../../include/hpvcd.h"
#include "gen.h"
#include "hevc_ps.h"
#include "internal.h"

namespace hpvcd {

uint64_t tile_seed(uint64_t base, uint32_t config_id, uint32_t tile) {
    uint64_t s = base ^ ((uint64_t)config_id << 32) ^ (uint64_t)tile;
    return splitmix64_step(s);
}

// The walker's parameters for the generator's pictures.  The transform
// depth is capped at its SPS range and the PCM size range at what the SPS
// can signal; neither changes which syntax is present.
PicParams gen_params(const hpvcd_gen_config& c) {
    PicParams P;
    P.w = c.tile_width;
    P.h = c.tile_height;
    P.chroma_idc = c.chroma_idc;
    P.bd = P.bd_c = c.bit_depth;
    P.log2_ctb = c.log2_ctb;
    P.log2_min_cb = c.log2_min_cb;
    P.log2_min_tb = c.log2_min_tb;
    P.log2_max_tb = c.log2_max_tb;
    P.max_depth_intra = P.max_depth_inter = std::min<int>(c.max_transform_depth_intra, c.log2_ctb - c.log2_min_tb);
    P.tqb_enabled = c.p_tqb > 0.f;
    P.transform_skip = c.transform_skip != 0;
    P.log2_max_ts = c.log2_max_ts;
    P.implicit_rdpcm = P.explicit_rdpcm = c.implicit_rdpcm != 0;
    P.cross_component = c.cross_component != 0 && c.chroma_idc == 3;
    P.extended_precision = c.extended_precision != 0;
    P.persistent_rice = c.persistent_rice != 0;
    P.ts_context = c.ts_context != 0;
    P.bypass_alignment = c.bypass_alignment != 0;
    P.sign_hiding = c.sign_hiding != 0;
    P.act = c.act != 0;
    P.pcm = c.pcm != 0;
    P.pcm_bd = P.pcm_bd_c = c.pcm_bit_depth;
    P.log2_min_pcm = std::max(3, std::min<int>(c.log2_min_cb, 5));
    P.log2_max_pcm = std::min(5, (int)c.log2_ctb);
    P.cu_qp_delta = c.cu_qp_delta != 0;
    P.diff_cu_qp_delta_depth = c.diff_cu_qp_delta_depth;
    P.chroma_qp_off[0] = c.pps_cb_qp_offset + c.slice_cb_qp_offset;
    P.chroma_qp_off[1] = c.pps_cr_qp_offset + c.slice_cr_qp_offset;
    P.sao = c.sao != 0;
    P.sao_luma = true;
    P.sao_chroma = c.chroma_idc != 0;
    P.amp = c.amp != 0;
    P.wpp = c.wpp != 0;
    return P;
}

// Uniformly spaced tile boundaries in CTBs (uniform_spacing_flag, H.265
// eq. 6-3/6-4) and the tile-scan order of the CTBs (CtbAddrTsToRs).
TileLayout::TileLayout(int cols, int rows, int ntc, int ntr) {
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

// Slices of the generator's pictures: equal runs of the tile scan.
std::vector<uint8_t> gen_slices(const hpvcd_gen_config& cfg, const TileLayout& T) {
    const int nctb = (int)T.ts_to_rs.size();
    const int nslices = std::max(1, std::min<int>(std::max<int>(1, cfg.num_slices), std::min(nctb, HPVCD_MAX_SLICES)));
    std::vector<uint8_t> ctb_slice((size_t)nctb, 0);
    for (int ts = 0; ts < nctb; ts++) ctb_slice[T.ts_to_rs[ts]] = (uint8_t)((int64_t)ts * nslices / nctb);
    return ctb_slice;
}

// Deblocking parameters of the slices of picture `index`: the configured
// offsets, or per slice a draw of disable, offsets and the across-slices
// flag.  Values a slice header cannot carry take what a parse infers: a
// disabled slice has zero offsets, and the across-slices flag of a slice
// without SAO and without deblocking is the PPS flag.
void gen_slice_params(const hpvcd_gen_config& c, uint32_t index, int nslices, hpvcd_slice_params* out) {
    Rng prng{tile_seed(c.base_seed ^ 0x5b1cull, c.config_id, index)};
    for (int sl = 0; sl < nslices; sl++) {
        hpvcd_slice_params& S = out[sl];
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
        if (S.disabled) S.beta_offset_div2 = S.tc_offset_div2 = 0;
        if (!c.sao && S.disabled) S.lf_across = c.lf_across_slices ? 1 : 0;
    }
}

namespace {

// Seeded decisions: the coding tree, modes, CBFs and coefficient levels of
// synthetic pictures drawn from the configured distributions.
class RandomSource : public Source {
public:
    RandomSource(const hpvcd_gen_config& c, uint64_t seed) : cfg(c), r{seed} {}
    void begin_slice(int, int) override {}
    void begin_substream(int) override {}
    int end_ctb(bool last_of_slice, bool) override { return last_of_slice ? 1 : 0; }
    // merged from the left or above CTB (where available) or drawn; offsets
    // |v| <= 7 with edge-offset signs fixed (src/decode.rs:1507-1614)
    void sao(hpvcd_sao& s, bool left_avail, bool up_avail, const hpvcd_sao* left, const hpvcd_sao* up) override {
        std::memset(&s, 0, sizeof(s));
        if (cfg.p_sao_merge > 0.f) {
            if (left_avail && r.bern(cfg.p_sao_merge)) { s = *left; return; }
            if (up_avail && r.bern(cfg.p_sao_merge)) { s = *up; return; }
        }
        const int ncomp = cfg.chroma_idc ? 3 : 1;
        for (int c = 0; c < ncomp; c++) {
            const int type = c == 2 ? s.type_idx[1] : (int)r.uni(3);
            s.type_idx[c] = (uint8_t)type;
            if (!type) continue;
            int v[4];
            for (int k = 0; k < 4; k++) v[k] = (int)r.uni(8);
            if (type == 1) {
                for (int k = 0; k < 4; k++)
                    if (v[k] && (r.next() & 1)) v[k] = -v[k];
                s.band_pos[c] = (uint8_t)r.uni(32);
            } else {
                v[2] = -v[2];
                v[3] = -v[3];
                s.eo_class[c] = c == 2 ? s.eo_class[1] : (uint8_t)r.uni(4);
            }
            for (int k = 0; k < 4; k++) s.offsets[c][k] = (int16_t)v[k];
        }
    }
    bool split_cu(int) override { return r.bern(cfg.p_split_cu); }
    bool tqb() override { return r.bern(cfg.p_tqb); }
    bool intra_nxn() override { return r.bern(cfg.p_nxn); }
    bool pcm_flag() override { return r.bern(cfg.p_pcm); }
    void pcm_samples(int n_luma, int n_chroma, int bd_y, int bd_c, uint32_t* out) override {
        for (int i = 0; i < n_luma; i++) out[i] = r.uni(1u << bd_y);
        for (int i = 0; i < 2 * n_chroma; i++) out[n_luma + i] = r.uni(1u << bd_c);
    }
    void luma_modes(int npu, FnRef<void(int, uint8_t*)>,
                    FnRef<void(int, uint8_t)> commit) override {
        for (int k = 0; k < npu; k++) commit(k, (uint8_t)r.uni(35));
    }
    int chroma_mode_syntax() override { return (int)r.uni(5); }
    bool split_tf(int) override { return r.bern(cfg.p_split_tu); }
    bool cbf_chroma(int) override { return r.bern(cfg.p_cbf); }
    bool cbf_luma(int) override { return r.bern(cfg.p_cbf); }
    bool act_flag() override { return r.bern(cfg.p_act); }
    int qp_delta() override { return r.sint(-(int)cfg.qp_delta_max, (int)cfg.qp_delta_max); }
    // ResScaleVal = +-(1 << (log2_res_scale_abs_plus1 - 1)) or 0
    int res_scale(int) override {
        const int m = cfg.log2_res_scale_max;
        const int v = r.sint(-m, m);
        if (v == 0) return 0;
        const int mag = 1 << (std::abs(v) - 1);
        return v < 0 ? -mag : mag;
    }
    // transform_skip_flag, explicit RDPCM, then the levels in raster order:
    // Bernoulli significance, geometric magnitude >= 1, random sign; at
    // least one nonzero (the coded last position exists whenever cbf is
    // set).  With sign hiding the hidden signs follow the magnitudes.
    void residual(const ResidualParams& p, int32_t* lv, ResidualOut& o) override {
        o.ts = p.ts_allowed && r.bern(cfg.p_ts);
        o.rdpcm = -1;
        if (p.explicit_rdpcm && (o.ts || p.tqb)) o.rdpcm = (int)r.uni(3) - 1;
        const int n = 1 << p.log2n;
        const int lmax = cfg.extended_precision ? HPVCD_COEFF_LEVEL_MAX : 32767;
        bool any = false;
        for (int pos = 0; pos < n * n; pos++) {
            lv[pos] = 0;
            if (!r.bern(cfg.p_sig)) continue;
            int level = 1;
            while (level < lmax && !r.bern(cfg.p_level_geom)) level++;
            if (r.next() & 1) level = -level;
            lv[pos] = level;
            any = true;
        }
        if (!any) {
            const int pos = (int)r.uni((uint32_t)(n * n));
            lv[pos] = (r.next() & 1) ? 1 : -1;
        }
        if (p.sign_hiding) residual_hide_signs(p, o, lv);
    }
    // inter decisions: skip, intra CUs, partitions (NxN at the smallest CU
    // above 8x8, AMP above the smallest), merge, prediction lists (no bi for
    // 8x4 / 4x8), reference indices, and vectors ~ N(0, mv_sigma) keeping the
    // block within 64 samples of the picture, coded against a random
    // predictor of the two
    bool cu_skip(int) override { return r.bern(cfg.p_skip); }
    bool pred_mode_intra() override { return r.bern(cfg.p_intra_cu); }
    int part_mode_inter(int log2, bool min_size, bool amp) override {
        int opts[8], n = 0;
        opts[n++] = PART_2Nx2N;
        opts[n++] = PART_2NxN;
        opts[n++] = PART_Nx2N;
        if (min_size && log2 > 3) opts[n++] = PART_NxN;
        if (amp)
            for (int k = PART_2NxnU; k <= PART_nRx2N; k++) opts[n++] = k;
        return opts[r.uni((uint32_t)n)];
    }
    bool merge_flag() override { return r.bern(cfg.p_merge); }
    int merge_idx(int max_cand) override { return (int)r.uni((uint32_t)max_cand); }
    int inter_pred_idc(int, bool no_bi) override {
        if (!no_bi && r.bern(cfg.p_bi)) return 2;
        return (r.next() & 1) ? 0 : 1;
    }
    int ref_idx(int num) override { return (int)r.uni((uint32_t)num); }
    void mv_syntax(const PuGeom& pu, int, const int16_t cand[2][2], bool mvd_zero, int16_t mvd[2],
                   int& mvp_flag) override {
        mvp_flag = (int)(r.next() & 1);
        if (mvd_zero) {
            mvd[0] = mvd[1] = 0;
            return;
        }
        const double sig = cfg.mv_sigma * 4.0;
        int mvx = (int)std::lround(gauss() * sig), mvy = (int)std::lround(gauss() * sig);
        mvx = std::clamp(mvx, (-64 - pu.x - pu.w) * 4, (cfg.tile_width + 64 - pu.x) * 4);
        mvy = std::clamp(mvy, (-64 - pu.y - pu.h) * 4, (cfg.tile_height + 64 - pu.y) * 4);
        mvd[0] = (int16_t)(uint16_t)(uint32_t)(mvx - cand[mvp_flag][0]);
        mvd[1] = (int16_t)(uint16_t)(uint32_t)(mvy - cand[mvp_flag][1]);
    }
    bool rqt_root_cbf() override { return r.bern(0.85f); }

private:
    double gauss() {
        const double u1 = ((double)(r.next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = (double)(r.next() >> 11) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    const hpvcd_gen_config& cfg;
    Rng r;
};

}  // namespace

int32_t gen_poc(const hpvcd_gen_config& c, uint32_t index) {
    if (!c.inter || c.gop <= 1) return (int32_t)index;
    const std::vector<int> order = gop_decode_order(c.grid_cols * c.grid_rows, c.gop);
    return index < order.size() ? order[index] : (int32_t)index;
}

void gen_lists(const hpvcd_gen_config& c, uint32_t index, std::vector<uint32_t>& l0, std::vector<uint32_t>& l1) {
    l0.clear();
    l1.clear();
    if (!c.inter || index == 0) return;
    const int nr = c.num_refs ? c.num_refs : 1;
    if (c.gop > 1) {
        gop_ref_lists(c.grid_cols * c.grid_rows, c.gop, nr, index, l0, l1);
    } else {
        const int n = std::max(1, std::min(nr, (int)index));
        for (int i = 0; i < n; i++) {
            l0.push_back(index - 1 - (uint32_t)i);
            l1.push_back(index - (uint32_t)n + (uint32_t)i);
        }
    }
    if (c.p_bi <= 0.f) l1.clear();
}

// Each list entry is a pred_weight_table row (src/inter.rs:141): luma and
// chroma weights each signalled with probability 0.8 (both chroma
// components under one flag), weights in [32, 96] of denominator 64,
// offsets in [-16, 16] 8-bit units scaled to the bit depth.
void gen_inter_slice(const hpvcd_gen_config& c, uint32_t index, InterSlice& s) {
    s = InterSlice();
    Rng rng{tile_seed(c.base_seed ^ 0x1e7e5ull, c.config_id, index)};
    std::vector<uint32_t> lists[2];
    gen_lists(c, index, lists[0], lists[1]);
    const bool b = c.p_bi > 0.f;
    s.slice_type = b ? 0 : 1;
    for (int l = 0; l < 2; l++)
        for (uint32_t pic : lists[l]) {
            RefPic rp;
            rp.poc = gen_poc(c, pic);
            rp.pic = (int32_t)pic;
            s.list[l].push_back(rp);
            hpvcd_ref e;
            std::memset(&e, 0, sizeof(e));
            e.picture = pic;
            e.weighted = c.weighted ? 1 : 0;
            e.log2_denom[0] = e.log2_denom[1] = 6;
            const bool luma = c.weighted && rng.bern(0.8f);
            const bool chroma = c.weighted && c.chroma_idc && rng.bern(0.8f);
            for (int k = 0; k < 3; k++) {
                const bool signalled = k == 0 ? luma : chroma;
                e.weight[k] = (int16_t)(signalled ? rng.sint(32, 96) : 64);
                e.offset[k] = (int16_t)(signalled ? rng.sint(-16, 16) * (1 << (c.bit_depth - 8)) : 0);
            }
            // a 4:0:0 table has no chroma denominator: ChromaLog2WeightDenom 0
            // and the neutral weight 1 (src/inter.rs:170)
            if (c.weighted && !c.chroma_idc) {
                e.log2_denom[1] = 0;
                e.weight[1] = e.weight[2] = 1;
            }
            s.rows.push_back(e);
        }
    s.tmvp = c.tmvp != 0;
    s.col_from_l0 = b ? rng.bern(0.5f) : true;
    const size_t ncol = s.list[s.col_from_l0 ? 0 : 1].size();
    s.col_ref_idx = ncol > 1 ? (int)rng.uni((uint32_t)ncol) : 0;
    s.max_merge = c.merge_cand ? std::min<int>(c.merge_cand, 5) : 5;
    s.par_mrg_level = c.par_merge_level ? std::min<int>(c.par_merge_level, c.log2_ctb) : 2;
    s.mvd_l1_zero = b && c.mvd_l1_zero;
    s.cabac_init = c.cabac_init && rng.bern(0.5f);
}

int gen_col_picture(const hpvcd_gen_config& c, uint32_t index) {
    if (!c.inter || !c.tmvp || index == 0) return -1;
    InterSlice s;
    gen_inter_slice(c, index, s);
    const std::vector<RefPic>& l = s.list[s.col_from_l0 ? 0 : 1];
    return (size_t)s.col_ref_idx < l.size() ? l[(size_t)s.col_ref_idx].pic : -1;
}

void gen_for_pictures(const hpvcd_gen_config& c, const std::function<void(uint32_t)>& fn) {
    const uint32_t n = (uint32_t)(c.grid_cols * c.grid_rows);
    // waves: a picture one after its collocated picture's
    std::vector<int> wave(n, 0);
    int nwaves = 1;
    for (uint32_t i = 0; i < n; i++) {
        const int col = gen_col_picture(c, i);
        wave[i] = col >= 0 && (uint32_t)col < i ? wave[(size_t)col] + 1 : 0;
        nwaves = std::max(nwaves, wave[i] + 1);
    }
    const unsigned nth = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    for (int wv = 0; wv < nwaves; wv++) {
        std::vector<uint32_t> todo;
        for (uint32_t i = 0; i < n; i++)
            if (wave[i] == wv) todo.push_back(i);
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth && t < todo.size(); t++)
            th.emplace_back([&, t]() {
                for (size_t k = t; k < todo.size(); k += nth) fn(todo[k]);
            });
        for (auto& x : th) x.join();
    }
}

// One generated picture: records, and the coded slice data when `enc` is set.
void gen_picture(const hpvcd_gen_config& cfg, uint32_t index, PicRecords& out, PicCoder* enc, const ColField* col) {
    const PicParams P = gen_params(cfg);
    RandomSource rs(cfg, tile_seed(cfg.base_seed, cfg.config_id, index));
    Source* src = &rs;
    std::unique_ptr<Source> coded;
    if (enc) {
        coded = enc->wrap(rs, P);
        src = coded.get();
    }
    PicWalker g(P, *src, out, gen_poc(cfg, index));
    InterSlice inter;
    if (cfg.inter && index > 0) {
        gen_inter_slice(cfg, index, inter);
        inter.col = col;
        g.set_inter(&inter);
    }
    const int ctb = 1 << cfg.log2_ctb;
    const int cols = (P.w + ctb - 1) / ctb, rows = (P.h + ctb - 1) / ctb;
    const int nctb = cols * rows;
    // CTBs in tile-scan order; slices are equal runs of that order
    TileLayout T(cols, rows, std::max<int>(1, cfg.tile_cols), std::max<int>(1, cfg.tile_rows));
    const std::vector<uint8_t> ctb_slice = gen_slices(cfg, T);
    g.set_layout(ctb_slice, T.tile_of);
    out.num_slices = ctb_slice.empty() ? 1 : ctb_slice[T.ts_to_rs[nctb - 1]] + 1;
    out.num_tiles = T.num_tiles();
    // substreams: one per tile, and with WPP one per CTB row of a tile
    auto sub_start = [&](int t) {
        const uint32_t r = T.ts_to_rs[t];
        if (t == 0 || T.tile_of[r] != T.tile_of[T.ts_to_rs[t - 1]]) return true;
        return P.wpp && (r % (uint32_t)cols == 0 || T.tile_of[r - 1] != T.tile_of[r]);
    };
    for (int ts = 0; ts < nctb; ts++) {
        const uint32_t rs_addr = T.ts_to_rs[ts];
        const bool slice_start = ts == 0 || ctb_slice[rs_addr] != ctb_slice[T.ts_to_rs[ts - 1]];
        const bool tile_start = sub_start(ts);
        const bool slice_end = ts == nctb - 1 || ctb_slice[rs_addr] != ctb_slice[T.ts_to_rs[ts + 1]];
        const bool tile_end = ts == nctb - 1 || sub_start(ts + 1);
        g.ctb((int)(rs_addr % cols), (int)(rs_addr / cols), slice_start, tile_start, slice_end, tile_end, cfg.qp);
    }
    g.finish();
}

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

hpvcd_status gen_validate(const hpvcd_gen_config& c) {
    if (c.tile_width <= 0 || c.tile_height <= 0 || c.grid_cols <= 0 || c.grid_rows <= 0 ||
        c.tile_width % (1 << c.log2_min_cb) || c.tile_height % (1 << c.log2_min_cb) || c.tile_width > 8192 ||
        c.tile_height > 8192)
        return set_error(HPVCD_E_BAD_DIMENSIONS, "generator: tile size must be a positive multiple of the min CU");
    if (c.chroma_idc == 2) return set_error(HPVCD_E_UNSUPPORTED_CHROMA, "generator: 4:2:2 not generated");
    if (c.gop > 64 || (c.gop > 1 && !c.inter)) return set_error(HPVCD_E_PARAMSET, "generator: GOP size 2..64 of an inter sequence");
    if (c.bit_depth < 8 || c.bit_depth > 12) return set_error(HPVCD_E_UNSUPPORTED_BITDEPTH, "generator: bit depth");
    if (c.merge_cand > 5 || c.par_merge_level == 1 || c.par_merge_level > c.log2_ctb || c.num_refs > 8)
        return set_error(HPVCD_E_PARAMSET, "generator: merge candidates 0..5, merge level 0 or 2..log2_ctb, 0..8 references");
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
    return HPVCD_OK;
}

hpvcd_status gen_assemble(const hpvcd_gen_config& c, std::vector<PicRecords>& tiles, hpvcd_records** out) {
    const uint32_t ntiles = (uint32_t)tiles.size();
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
        if (HPVCD_OUT_IS_YUV(c.out_kind)) {
            ho.channels = HPVCD_OUT_YUV_CHROMA(c.out_kind) ? 3 : 1;
            ho.bytes_per_sample = HPVCD_OUT_YUV_WIDE(c.out_kind) ? 2 : 1;
        } else {
            ho.channels = (c.out_kind == HPVCD_OUT_LUMA8 || c.out_kind == HPVCD_OUT_LUMA16) ? 1 : 3;
            ho.bytes_per_sample = (c.out_kind == HPVCD_OUT_RGB16 || c.out_kind == HPVCD_OUT_LUMA16) ? 2 : 1;
        }
        ho.orientation = c.orientation;
        r->outputs.push_back(ho);
    }
    // output of each picture: decode order is display order unless the
    // pictures form hierarchical GOPs
    std::vector<int> display(ntiles);
    for (uint32_t i = 0; i < ntiles; i++) display[i] = (int)i;
    if (c.inter && c.gop > 1) display = gop_decode_order((int)ntiles, c.gop);
    for (uint32_t i = 0; i < ntiles; i++) {
        PicRecords& t = tiles[i];
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
        gen_slice_params(c, i, t.num_slices, p.slices);
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
        p.out_index = nout > 1 ? (uint32_t)display[i] % (uint32_t)nout : 0u;
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

}  // namespace hpvcd

using namespace hpvcd;

extern "C" hpvcd_status hpvcd_gen_records(const hpvcd_gen_config* cfg, hpvcd_records** out) {
    if (!cfg || !out) return set_error(HPVCD_E_INVALID, "null argument");
    *out = nullptr;
    const hpvcd_gen_config& c = *cfg;
    if (hpvcd_status st = gen_validate(c)) return st;
    const uint32_t ntiles = (uint32_t)(c.grid_cols * c.grid_rows);
    std::vector<PicRecords> tiles(ntiles);
    gen_for_pictures(c, [&](uint32_t i) {
        const int col = gen_col_picture(c, i);
        gen_picture(c, i, tiles[i], nullptr, col >= 0 ? &tiles[(size_t)col].col : nullptr);
    });
    return gen_assemble(c, tiles, out);
}


static_assert(sizeof(hpvcd_gen_config) == 176, "hpvcd_gen_config layout");
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
    c->p_merge = 0.4f;
    c->config_id = 0;
    c->base_seed = 0x48505643ull;
    c->qp_delta_max = 4;
    c->lf_across_slices = 1;
    c->lf_across_tiles = 1;
    c->num_slices = 1;
    c->tile_cols = 1;
    c->tile_rows = 1;
}

