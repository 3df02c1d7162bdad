// gen_hevc.cpp — the generator's pictures as coded HEVC.
//
// The seeded decisions of gen.cpp are CABAC-coded as the walker asks for
// them (CodedSource), which makes a bitstream whose parse yields the
// generator's records.  Syntax and binarisations follow H.265 7.3.8 and
// 9.3.3 as the reference decoder reads them: SAO (src/decode.rs:1507),
// split_cu_flag (:2540), part_mode / pcm_flag (:2620-2660), intra modes
// (:3204, :3248), transform tree and unit (:3315, :3478), cu_qp_delta
// (:3528), tu_residual_act_flag (:4229), cross_comp_pred (:4257),
// residual_coding (src/cabac/residual_full.rs:392), end_of_slice_segment_flag
// and end_of_subset_one_bit (src/decode.rs:1319-1435).
#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hpvcd.h"
#include "cabac.h"
#include "gen.h"
#include "gen_hevc.h"
#include "hevc_ps.h"
#include "internal.h"

namespace hpvcd {

namespace {

// k-th order Exp-Golomb, bypass bins (9.3.3.3)
void enc_egk(CabacEnc& e, uint32_t v, int k) {
    for (;;) {
        if (v >= (1u << k)) {
            e.bypass(1);
            v -= 1u << k;
            k++;
        } else {
            e.bypass(0);
            while (k--) e.bypass((v >> k) & 1);
            return;
        }
    }
}

// The coded form of every decision of `in`: one substream per tile of a
// slice, each ending in its terminating bin and byte alignment.
class CodedSource : public Source {
public:
    CodedSource(Source& in, const PicParams& P, PicStream& out) : in_(in), P_(P), out_(out) {}

    void begin_slice(int slice_qp, int init_type) override {
        in_.begin_slice(slice_qp, init_type);
        slice_qp_ = slice_qp;
        init_type_ = init_type;
        out_.slices.emplace_back();
        open_substream();
    }
    void begin_substream(int sync_row) override {
        in_.begin_substream(sync_row);
        open_substream();
        if (sync_row >= 0 && (size_t)sync_row < saved_.size()) c_ = saved_[(size_t)sync_row];
    }
    void wpp_save(int row) override {
        if (saved_.size() <= (size_t)row) saved_.resize((size_t)row + 1);
        saved_[(size_t)row] = c_;
    }
    int end_ctb(bool last_of_slice, bool last_of_tile) override {
        const int r = in_.end_ctb(last_of_slice, last_of_tile);
        e_.terminate(last_of_slice ? 1 : 0);  // end_of_slice_segment_flag
        if (last_of_slice) {
            w_.align_zero();  // the flush wrote rbsp_stop_one_bit
            close_substream();
        } else if (last_of_tile) {
            e_.terminate(1);  // end_of_subset_one_bit
            w_.align_zero();  // byte_alignment(): the flush wrote its one bit
            close_substream();
        }
        return r;
    }

    void sao(hpvcd_sao& s, bool left_avail, bool up_avail, const hpvcd_sao* left, const hpvcd_sao* up) override {
        in_.sao(s, left_avail, up_avail, left, up);
        // a parameter set equal to a candidate's is coded as a merge
        if (left_avail) {
            const bool m = std::memcmp(&s, left, sizeof(s)) == 0;
            e_.encode(c_.s[CTX_SAO_MERGE], m);
            if (m) return;
        }
        if (up_avail) {
            const bool m = std::memcmp(&s, up, sizeof(s)) == 0;
            e_.encode(c_.s[CTX_SAO_MERGE], m);
            if (m) return;
        }
        const int ncomp = P_.chroma_idc ? 3 : 1;
        for (int c = 0; c < ncomp; c++) {
            if ((c == 0 && !P_.sao_luma) || (c > 0 && !P_.sao_chroma)) continue;
            const int type = s.type_idx[c];
            if (c < 2) {  // sao_type_idx_luma / _chroma: TR cMax 2, bin 0 coded, bin 1 bypass
                e_.encode(c_.s[CTX_SAO_TYPE], type != 0);
                if (type) e_.bypass(type == 2);
            }
            if (!type) continue;
            const int bd = c ? P_.bd_c : P_.bd;
            const int cmax = (1 << (std::min(bd, 10) - 5)) - 1;
            for (int k = 0; k < 4; k++) {
                const int a = std::abs((int)s.offsets[c][k]);
                if (a > cmax) fail("SAO offset beyond sao_offset_abs range");
                for (int i = 0; i < a; i++) e_.bypass(1);
                if (a < cmax) e_.bypass(0);
            }
            if (type == 1) {
                for (int k = 0; k < 4; k++)
                    if (s.offsets[c][k]) e_.bypass(s.offsets[c][k] < 0);
                e_.bypass_bits(s.band_pos[c], 5);
            } else if (c < 2) {
                e_.bypass_bits(s.eo_class[c], 2);
            }
        }
    }
    bool split_cu(int ctx_inc) override {
        const bool v = in_.split_cu(ctx_inc);
        e_.encode(c_.s[CTX_SPLIT_CU + ctx_inc], v);
        return v;
    }
    bool tqb() override {
        const bool v = in_.tqb();
        e_.encode(c_.s[CTX_TQB], v);
        return v;
    }
    bool intra_nxn() override {  // part_mode bin 0 of an intra CU: 1 = PART_2Nx2N
        const bool v = in_.intra_nxn();
        e_.encode(c_.s[CTX_PART_MODE], v ? 0 : 1);
        return v;
    }
    bool pcm_flag() override {
        const bool v = in_.pcm_flag();
        e_.terminate(v);
        if (v) w_.align_zero();  // pcm_alignment_zero_bit
        return v;
    }
    void pcm_samples(int n_luma, int n_chroma, int bd_y, int bd_c, uint32_t* out) override {
        in_.pcm_samples(n_luma, n_chroma, bd_y, bd_c, out);
        for (int i = 0; i < n_luma; i++) w_.bits(out[i], bd_y);
        for (int i = 0; i < 2 * n_chroma; i++) w_.bits(out[n_luma + i], bd_c);
        e_.start(&w_);  // 9.3.2.5 after the PCM samples
    }
    void luma_modes(int npu, FnRef<void(int, uint8_t*)> mpm,
                    FnRef<void(int, uint8_t)> commit) override {
        int flag[4] = {0, 0, 0, 0}, val[4] = {0, 0, 0, 0};
        in_.luma_modes(npu, mpm, [&](int k, uint8_t mode) {
            uint8_t m[3];
            mpm(k, m);
            flag[k] = 0;
            for (int i = 0; i < 3; i++)
                if (m[i] == mode) {
                    flag[k] = 1;
                    val[k] = i;
                }
            if (!flag[k]) {  // rem_intra_luma_pred_mode: the mode among the 32 non-candidates
                int rem = mode;
                for (int i = 0; i < 3; i++) rem -= m[i] < mode ? 1 : 0;
                val[k] = rem;
            }
            commit(k, mode);
        });
        for (int k = 0; k < npu; k++) e_.encode(c_.s[CTX_PREV_INTRA], flag[k]);
        for (int k = 0; k < npu; k++) {
            if (flag[k]) {  // mpm_idx: TR cMax 2, bypass
                e_.bypass(val[k] > 0);
                if (val[k] > 0) e_.bypass(val[k] > 1);
            } else {
                e_.bypass_bits((uint32_t)val[k], 5);
            }
        }
    }
    int chroma_mode_syntax() override {
        const int v = in_.chroma_mode_syntax();
        e_.encode(c_.s[CTX_CHROMA_MODE], v != 4);
        if (v != 4) e_.bypass_bits((uint32_t)v, 2);
        return v;
    }
    bool split_tf(int log2) override {
        const bool v = in_.split_tf(log2);
        e_.encode(c_.s[CTX_SPLIT_TF + 5 - log2], v);
        return v;
    }
    bool cbf_chroma(int depth) override {
        const bool v = in_.cbf_chroma(depth);
        e_.encode(c_.s[CTX_CBF_CHROMA + depth], v);
        return v;
    }
    bool cbf_luma(int depth) override {
        const bool v = in_.cbf_luma(depth);
        e_.encode(c_.s[CTX_CBF_LUMA + (depth == 0 ? 1 : 0)], v);
        return v;
    }
    bool act_flag() override {
        const bool v = in_.act_flag();
        e_.encode(c_.s[CTX_ACT], v);
        return v;
    }
    // cu_qp_delta_abs: TU prefix cMax 5 (bin 0 context 0, others 1), EG0
    // suffix; cu_qp_delta_sign_flag bypass
    int qp_delta() override {
        const int v = in_.qp_delta();
        const int a = std::abs(v), pre = std::min(a, 5);
        for (int i = 0; i < pre; i++) e_.encode(c_.s[CTX_QP_DELTA + (i ? 1 : 0)], 1);
        if (pre < 5) e_.encode(c_.s[CTX_QP_DELTA + (pre ? 1 : 0)], 0);
        else enc_egk(e_, (uint32_t)(a - 5), 0);
        if (a) e_.bypass(v < 0);
        return v;
    }
    // log2_res_scale_abs_plus1: TR cMax 4, context 4c + binIdx; sign context c
    int res_scale(int c) override {
        const int v = in_.res_scale(c);
        int a = 0;
        while (a < 4 && (1 << a) <= std::abs(v)) a++;
        if (v && (1 << (a - 1)) != std::abs(v)) fail("ResScaleVal not a power of two");
        for (int i = 0; i < a; i++) e_.encode(c_.s[CTX_RES_SCALE + 4 * c + i], 1);
        if (a < 4) e_.encode(c_.s[CTX_RES_SCALE + 4 * c + a], 0);
        if (a) e_.encode(c_.s[CTX_RES_SIGN + c], v < 0);
        return v;
    }
    void residual(const ResidualParams& p, int32_t* levels, ResidualOut& r) override {
        in_.residual(p, levels, r);
        // coeff_abs_level_remaining covers |level| <= 32768 (2^log2TransformRange
        // with extended precision)
        const int range = p.extended_precision ? std::max(15, p.bit_depth + 6) : 15;
        const int64_t lim = (int64_t)1 << range;
        const int n = 1 << p.log2n;
        for (int i = 0; i < n * n; i++)
            if (std::abs((int64_t)levels[i]) > lim) fail("coefficient level beyond the coded range");
        residual_encode(e_, c_, p, levels, r);
    }
    // inter syntax (binarisations as src/decode.rs:6290, :6519, :6571, :6771,
    // :6878, :6929 read them)
    bool cu_skip(int ctx_inc) override {
        const bool v = in_.cu_skip(ctx_inc);
        e_.encode(c_.s[CTX_SKIP + ctx_inc], v);
        return v;
    }
    bool pred_mode_intra() override {
        const bool v = in_.pred_mode_intra();
        e_.encode(c_.s[CTX_PRED_MODE], v);
        return v;
    }
    // part_mode: 1 for 2Nx2N; else 0, then the direction (1 horizontal);
    // above the smallest CU with AMP a symmetric bin (context 3) and the AMP
    // position in bypass; at the smallest CU above 8x8 the Nx2N / NxN bin
    // (context 2)
    int part_mode_inter(int log2, bool min_size, bool amp) override {
        const int v = in_.part_mode_inter(log2, min_size, amp);
        e_.encode(c_.s[CTX_PART_MODE], v == PART_2Nx2N);
        if (v == PART_2Nx2N) return v;
        const bool horizontal = v == PART_2NxN || v == PART_2NxnU || v == PART_2NxnD;
        e_.encode(c_.s[CTX_PART_MODE + 1], horizontal);
        if (!min_size) {
            if (v == PART_NxN || (!amp && v >= PART_2NxnU)) fail("part_mode not codable at this CU size");
            if (!amp) return v;
            const bool sym = v == PART_2NxN || v == PART_Nx2N;
            e_.encode(c_.s[CTX_PART_MODE + 3], sym);
            if (!sym) e_.bypass(v == PART_2NxnD || v == PART_nRx2N);
        } else {
            if (v >= PART_2NxnU) fail("AMP at the smallest CU");
            if (horizontal || log2 == 3) {
                if (v == PART_NxN) fail("NxN inter partition of an 8x8 CU");
                return v;
            }
            e_.encode(c_.s[CTX_PART_MODE + 2], v == PART_Nx2N);
        }
        return v;
    }
    bool merge_flag() override {
        const bool v = in_.merge_flag();
        e_.encode(c_.s[CTX_MERGE_FLAG], v);
        return v;
    }
    // merge_idx: TR cMax MaxNumMergeCand - 1, bin 0 context coded
    int merge_idx(int max_cand) override {
        const int v = in_.merge_idx(max_cand);
        e_.encode(c_.s[CTX_MERGE_IDX], v > 0);
        if (v > 0) {
            for (int i = 1; i < v; i++) e_.bypass(1);
            if (v < max_cand - 1) e_.bypass(0);
        }
        return v;
    }
    // inter_pred_idc: bi bin (context CtDepth) unless 8x4 / 4x8, then the
    // list bin (context 4)
    int inter_pred_idc(int ct_depth, bool no_bi) override {
        const int v = in_.inter_pred_idc(ct_depth, no_bi);
        if (!no_bi) e_.encode(c_.s[CTX_INTER_PRED + ct_depth], v == 2);
        else if (v == 2) fail("bi-prediction of an 8x4 / 4x8 PU");
        if (v != 2) e_.encode(c_.s[CTX_INTER_PRED + 4], v == 1);
        return v;
    }
    // ref_idx_lX: TR cMax num - 1, bins 0 and 1 context coded
    int ref_idx(int num) override {
        const int v = in_.ref_idx(num);
        e_.encode(c_.s[CTX_REF_IDX], v > 0);
        if (v == 0 || num == 2) return v;
        e_.encode(c_.s[CTX_REF_IDX + 1], v > 1);
        if (v == 1) return v;
        for (int i = 2; i < v; i++) e_.bypass(1);
        if (v < num - 1) e_.bypass(0);
        return v;
    }
    // mvd_coding: greater0 flags (context 0), greater1 flags (context 1),
    // then per component abs_mvd_minus2 (EG1) and the sign; mvp_lX_flag
    void mv_syntax(const PuGeom& pu, int list, const int16_t cand[2][2], bool mvd_zero, int16_t mvd[2],
                   int& mvp_flag) override {
        in_.mv_syntax(pu, list, cand, mvd_zero, mvd, mvp_flag);
        if (!mvd_zero) {
            const int a[2] = {std::abs((int)mvd[0]), std::abs((int)mvd[1])};
            for (int c = 0; c < 2; c++) e_.encode(c_.s[CTX_MVD], a[c] > 0);
            for (int c = 0; c < 2; c++)
                if (a[c] > 0) e_.encode(c_.s[CTX_MVD + 1], a[c] > 1);
            for (int c = 0; c < 2; c++) {
                if (!a[c]) continue;
                if (a[c] > 1) enc_egk(e_, (uint32_t)(a[c] - 2), 1);
                e_.bypass(mvd[c] < 0);
            }
        }
        e_.encode(c_.s[CTX_MVP_FLAG], mvp_flag);
    }
    bool rqt_root_cbf() override {
        const bool v = in_.rqt_root_cbf();
        e_.encode(c_.s[CTX_RQT_ROOT_CBF], v);
        return v;
    }

private:
    void open_substream() {
        c_.init(init_type_, slice_qp_);
        w_ = BitWriter();
        e_.start(&w_);
    }
    void close_substream() { out_.slices.back().substreams.push_back(std::move(w_.buf)); }
    void fail(const char* why) {
        if (!out_.failed) out_.why = why;
        out_.failed = true;
    }

    Source& in_;
    const PicParams P_;
    PicStream& out_;
    Contexts c_;
    std::vector<Contexts> saved_;  // WPP storage per CTB row
    CabacEnc e_;
    BitWriter w_;
    int slice_qp_ = 26, init_type_ = 0;
};

struct HevcCoder : PicCoder {
    PicStream stream;
    std::unique_ptr<Source> wrap(Source& inner, const PicParams& P) override {
        return std::make_unique<CodedSource>(inner, P, stream);
    }
};

// emulation_prevention_three_byte insertions inside one substream; a
// substream ends in a nonzero byte, so none straddles two of them
size_t ep_bytes(const std::vector<uint8_t>& b) {
    size_t n = 0;
    int zeros = 0;
    for (uint8_t x : b) {
        if (zeros >= 2 && x <= 3) {
            n++;
            zeros = 0;
        }
        zeros = x == 0 ? zeros + 1 : 0;
    }
    return n;
}

}  // namespace

SeqPlan gen_plan(const hpvcd_gen_config& c) {
    SeqPlan plan;
    const uint32_t n = (uint32_t)(c.grid_cols * c.grid_rows);
    plan.poc.resize(n);
    plan.rps.assign(n, StRps());
    plan.rps_idx.assign(n, 0);
    std::vector<std::vector<uint32_t>> refs(n);
    for (uint32_t i = 0; i < n; i++) {
        plan.poc[i] = gen_poc(c, i);
        std::vector<uint32_t> l0, l1;
        gen_lists(c, i, l0, l1);
        refs[i] = l0;
        for (uint32_t x : l1)
            if (std::find(refs[i].begin(), refs[i].end(), x) == refs[i].end()) refs[i].push_back(x);
    }
    // still needed: referenced by this picture or by one decoded later
    std::vector<uint8_t> later(n, 0);
    int max_jump = 1;
    for (uint32_t i = n; i-- > 1;) {
        std::vector<std::pair<int32_t, uint8_t>> neg, pos;
        for (uint32_t j = 0; j < i; j++) {
            const bool cur = std::find(refs[i].begin(), refs[i].end(), j) != refs[i].end();
            if (!cur && !later[j]) continue;
            const int32_t d = plan.poc[j] - plan.poc[i];
            (d < 0 ? neg : pos).emplace_back(d, cur ? 1 : 0);
        }
        std::sort(neg.begin(), neg.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        std::sort(pos.begin(), pos.end());
        StRps& r = plan.rps[i];
        for (auto& e : neg) { r.d0.push_back(e.first); r.u0.push_back(e.second); }
        for (auto& e : pos) { r.d1.push_back(e.first); r.u1.push_back(e.second); }
        for (uint32_t x : refs[i]) later[x] = 1;
        plan.max_dec = std::max(plan.max_dec, r.num() + 1);
        max_jump = std::max(max_jump, std::abs(plan.poc[i] - plan.poc[i - 1]));
    }
    for (uint32_t i = 0; i < n; i++) {
        int reorder = 0;
        for (uint32_t j = 0; j < i; j++) reorder += plan.poc[j] > plan.poc[i] ? 1 : 0;
        plan.max_reorder = std::max(plan.max_reorder, reorder);
    }
    // POC LSBs: half their range beyond any step between consecutive pictures
    while (plan.log2_max_poc_lsb < 16 && (1 << (plan.log2_max_poc_lsb - 1)) <= max_jump) plan.log2_max_poc_lsb++;
    if (c.rps_in_sps)
        for (uint32_t i = 1; i < n; i++) {
            size_t k = 0;
            while (k < plan.sps_sets.size() && !(plan.sps_sets[k] == plan.rps[i])) k++;
            if (k == plan.sps_sets.size() && k < 64) plan.sps_sets.push_back(plan.rps[i]);
            plan.rps_idx[i] = k < plan.sps_sets.size() ? (int)k : -1;   // -1: coded in the slice header
        }
    return plan;
}

// Parameter sets of the generator's configuration (one SPS / PPS for every
// picture of it).
void gen_parameter_sets(const hpvcd_gen_config& c, Sps& s, Pps& p, const SeqPlan* plan) {
    const PicParams P = gen_params(c);
    s = Sps();
    s.chroma_format_idc = c.chroma_idc;
    s.width = c.tile_width;
    s.height = c.tile_height;
    for (int i = 0; i < 4; i++) s.conf_win[i] = c.conf_win[i];
    s.bit_depth_luma = s.bit_depth_chroma = c.bit_depth;
    s.log2_min_cb = c.log2_min_cb;
    s.log2_ctb = c.log2_ctb;
    s.log2_min_tb = c.log2_min_tb;
    s.log2_max_tb = c.log2_max_tb;
    s.max_th_depth_intra = s.max_th_depth_inter = P.max_depth_intra;
    s.scaling_list_enabled = c.scaling_list != 0;
    s.scaling_list_data_present = c.scaling_list == 2;
    if (s.scaling_list_enabled) make_scaling_set(c, &s.scaling.e[0][0][0]);
    s.amp = c.amp != 0;
    s.sao = c.sao != 0;
    s.pcm = c.pcm != 0;
    s.pcm_bit_depth_luma = s.pcm_bit_depth_chroma = std::max<int>(1, c.pcm_bit_depth);
    s.log2_min_pcm_cb = P.log2_min_pcm;
    s.log2_max_pcm_cb = P.log2_max_pcm;
    s.pcm_loop_filter_disabled = c.pcm_lf_disabled != 0;
    s.strong_intra_smoothing = c.strong_intra_smoothing != 0;
    s.vui = true;
    s.colour_description = true;
    s.full_range = c.full_range != 0;
    s.colour_primaries = 2;
    s.transfer = 2;
    s.matrix = c.matrix;
    s.ts_rotation = c.ts_rotation != 0;
    s.ts_context = c.ts_context != 0;
    s.implicit_rdpcm = s.explicit_rdpcm = c.implicit_rdpcm != 0;
    s.extended_precision = c.extended_precision != 0;
    s.persistent_rice = c.persistent_rice != 0;
    s.bypass_alignment = c.bypass_alignment != 0;

    p = Pps();
    p.sign_hiding = c.sign_hiding != 0;
    p.init_qp = c.qp;
    p.transform_skip = c.transform_skip != 0;
    p.log2_max_ts = c.log2_max_ts;
    p.cu_qp_delta = c.cu_qp_delta != 0;
    p.diff_cu_qp_delta_depth = c.diff_cu_qp_delta_depth;
    p.cb_qp_offset = c.pps_cb_qp_offset;
    p.cr_qp_offset = c.pps_cr_qp_offset;
    p.slice_chroma_qp_offsets_present = c.slice_cb_qp_offset != 0 || c.slice_cr_qp_offset != 0;
    p.transquant_bypass = P.tqb_enabled;
    const int ctb = 1 << c.log2_ctb;
    const int cols = (c.tile_width + ctb - 1) / ctb, rows = (c.tile_height + ctb - 1) / ctb;
    TileLayout T(cols, rows, std::max<int>(1, c.tile_cols), std::max<int>(1, c.tile_rows));
    p.tiles = T.num_tiles() > 1;
    p.tile_cols = (int)T.col_bd.size() - 1;
    p.tile_rows = (int)T.row_bd.size() - 1;
    p.uniform_spacing = true;
    p.entropy_sync = c.wpp != 0;
    p.lf_across_tiles = c.lf_across_tiles != 0;
    p.lf_across_slices = c.lf_across_slices != 0;
    p.deblocking_control = true;
    p.deblocking_override = c.random_slice_params != 0;
    p.deblocking_disabled = c.deblock == 0;
    p.beta_offset_div2 = p.deblocking_disabled ? 0 : c.beta_offset_div2;
    p.tc_offset_div2 = p.deblocking_disabled ? 0 : c.tc_offset_div2;
    p.cross_component = P.cross_component;
    p.act = c.act != 0 && c.chroma_idc == 3;
    if (c.inter && plan) {
        s.log2_max_poc_lsb = plan->log2_max_poc_lsb;
        s.st_rps = plan->sps_sets;
        s.max_dec_pic_buffering = std::max(plan->max_dec, 1);
        s.max_num_reorder = plan->max_reorder;
        s.temporal_mvp = c.tmvp != 0;
        p.num_ref_idx_default[0] = p.num_ref_idx_default[1] = std::max<int>(1, c.num_refs);
        p.weighted_pred = p.weighted_bipred = c.weighted != 0;
        // low-delay B lists name the same pictures in opposite orders: list 1
        // is coded as a modification
        p.lists_modification = c.gop <= 1 && c.p_bi > 0.f;
        p.cabac_init_present = c.cabac_init != 0;
        p.log2_parallel_merge_level = c.par_merge_level ? std::min<int>(c.par_merge_level, c.log2_ctb) : 2;
    }
}

namespace {

// RefPicListTempX of a slice (8.3.4): StCurrBefore then StCurrAfter for
// list 0, the reverse for list 1, as picture indices
void default_lists(const StRps& rps, const std::vector<int32_t>& poc_of, int32_t cur_poc, std::vector<uint32_t>& t0,
                   std::vector<uint32_t>& t1) {
    std::vector<uint32_t> before, after;
    auto find = [&](int32_t poc) {
        for (size_t j = 0; j < poc_of.size(); j++)
            if (poc_of[j] == poc) return (uint32_t)j;
        return (uint32_t)0;
    };
    for (size_t k = 0; k < rps.d0.size(); k++)
        if (rps.u0[k]) before.push_back(find(cur_poc + rps.d0[k]));
    for (size_t k = 0; k < rps.d1.size(); k++)
        if (rps.u1[k]) after.push_back(find(cur_poc + rps.d1[k]));
    t0 = before;
    t0.insert(t0.end(), after.begin(), after.end());
    t1 = after;
    t1.insert(t1.end(), before.begin(), before.end());
}

// pred_weight_table syntax of a slice from its record rows (weights of
// denominator 64, offsets multiples of 1 << (bit depth - 8))
void weights_of(const InterSlice& is, int bd, bool chroma, PredWeights& t) {
    t = PredWeights();
    t.luma_log2_denom = 6;
    t.chroma_log2_denom = chroma ? 6 : 0;
    size_t row = 0;
    for (int l = 0; l < 2; l++)
        for (size_t i = 0; i < is.list[l].size(); i++, row++) {
            const hpvcd_ref& e = is.rows[row];
            WpEntry w;
            w.luma = e.weight[0] != 64 || e.offset[0] != 0;
            w.delta_weight = e.weight[0] - 64;
            w.offset = e.offset[0] >> (bd - 8);
            w.chroma = chroma && (e.weight[1] != 64 || e.offset[1] != 0 || e.weight[2] != 64 || e.offset[2] != 0);
            for (int j = 0; j < 2; j++) {
                const int wc = e.weight[1 + j], u = e.offset[1 + j] >> (bd - 8);
                w.delta_cweight[j] = wc - 64;
                w.delta_coffset[j] = u - (128 - ((128 * wc) >> 6));
            }
            t.e[l].push_back(w);
        }
}

}  // namespace

// Slice segment NAL units (header bytes included, escaped) of one coded picture.
std::vector<std::vector<uint8_t>> gen_slice_nals(const hpvcd_gen_config& c, const Sps& sps, const Pps& pps,
                                                 uint32_t index, const PicStream& st, const SeqPlan* plan) {
    const bool inter = c.inter && index > 0 && plan;
    const uint8_t nal_type = (c.inter && index > 0) ? NAL_TRAIL_R : NAL_IDR_N_LP;
    InterSlice is;
    if (inter) gen_inter_slice(c, index, is);
    std::vector<std::vector<uint8_t>> nals;
    const int ctb = 1 << c.log2_ctb;
    const int cols = (c.tile_width + ctb - 1) / ctb, rows = (c.tile_height + ctb - 1) / ctb;
    TileLayout T(cols, rows, std::max<int>(1, c.tile_cols), std::max<int>(1, c.tile_rows));
    const std::vector<uint8_t> ctb_slice = gen_slices(c, T);
    const int nslices = (int)st.slices.size();
    hpvcd_slice_params sp[HPVCD_MAX_SLICES];
    gen_slice_params(c, index, std::min(nslices, HPVCD_MAX_SLICES), sp);
    int ts = 0;
    for (int k = 0; k < nslices; k++) {
        SliceHeader h;
        h.first_slice = k == 0;
        h.segment_address = (int)T.ts_to_rs[ts];
        h.slice_type = 2;
        if (inter) {
            const int32_t poc = plan->poc[index];
            h.slice_type = is.slice_type;
            h.poc_lsb = poc & ((1 << sps.log2_max_poc_lsb) - 1);
            h.rps = plan->rps[index];
            h.rps_from_sps = c.rps_in_sps && plan->rps_idx[index] >= 0;
            h.rps_idx = h.rps_from_sps ? plan->rps_idx[index] : 0;
            h.tmvp = is.tmvp;
            const bool b = is.slice_type == 0;
            h.num_ref[0] = (int)is.list[0].size();
            h.num_ref[1] = b ? (int)is.list[1].size() : 0;
            h.num_ref_override = h.num_ref[0] != pps.num_ref_idx_default[0] || (b && h.num_ref[1] != pps.num_ref_idx_default[1]);
            // lists that differ from the default construction are coded as
            // modifications (entries index RefPicListTempX)
            std::vector<uint32_t> t[2];
            default_lists(h.rps, plan->poc, poc, t[0], t[1]);
            for (int l = 0; l < (b ? 2 : 1); l++) {
                bool same = true;
                for (size_t i = 0; i < is.list[l].size(); i++)
                    same &= !t[l].empty() && t[l][i % t[l].size()] == (uint32_t)is.list[l][i].pic;
                if (same) continue;
                h.list_mod[l] = true;
                for (size_t i = 0; i < is.list[l].size(); i++) {
                    size_t e = 0;
                    while (e < t[l].size() && t[l][e] != (uint32_t)is.list[l][i].pic) e++;
                    h.list_entry[l].push_back((uint32_t)e);
                }
            }
            h.mvd_l1_zero = is.mvd_l1_zero;
            h.cabac_init = is.cabac_init;
            h.col_from_l0 = is.col_from_l0;
            h.col_ref_idx = is.col_ref_idx;
            if (slice_weighted(h, pps)) weights_of(is, c.bit_depth, c.chroma_idc != 0, h.wp);
            h.max_merge = is.max_merge;
        }
        h.sao_luma = c.sao != 0;
        h.sao_chroma = c.sao != 0 && c.chroma_idc != 0;
        h.slice_qp_delta = 0;
        h.cb_qp_offset = c.slice_cb_qp_offset;
        h.cr_qp_offset = c.slice_cr_qp_offset;
        const hpvcd_slice_params& S = sp[std::min(k, HPVCD_MAX_SLICES - 1)];
        if (pps.deblocking_override) {
            h.deblocking_override = true;
            h.deblocking_disabled = S.disabled != 0;
            h.beta_offset_div2 = S.beta_offset_div2;
            h.tc_offset_div2 = S.tc_offset_div2;
        }
        h.lf_across_slices = S.lf_across != 0;
        const auto& subs = st.slices[k].substreams;
        for (size_t i = 0; i + 1 < subs.size(); i++)
            h.entry_points.push_back((uint32_t)(subs[i].size() + ep_bytes(subs[i])));
        BitWriter w;
        write_slice_header(w, h, sps, pps, nal_type);
        for (const auto& b : subs) w.buf.insert(w.buf.end(), b.begin(), b.end());
        nals.push_back(nal_from_rbsp(nal_type, w.buf));
        while (ts < (int)ctb_slice.size() && ctb_slice[T.ts_to_rs[ts]] == k) ts++;
    }
    return nals;
}

// Every picture of the configuration: records and coded slices.
hpvcd_status gen_coded(const hpvcd_gen_config& c, std::vector<PicRecords>& recs, std::vector<PicStream>& streams) {
    if (hpvcd_status st = gen_validate(c)) return st;
    if (c.cu_qp_delta && c.qp_delta_max > 25)
        return set_error(HPVCD_E_PARAMSET, "coded output: CuQpDeltaVal beyond [-26, 25]");
    const uint32_t n = (uint32_t)(c.grid_cols * c.grid_rows);
    recs.assign(n, PicRecords());
    streams.assign(n, PicStream());
    gen_for_pictures(c, [&](uint32_t i) {
        HevcCoder coder;
        const int col = gen_col_picture(c, i);
        gen_picture(c, i, recs[i], &coder, col >= 0 ? &recs[(size_t)col].col : nullptr);
        streams[i] = std::move(coder.stream);
    });
    for (uint32_t i = 0; i < n; i++)
        if (streams[i].failed) {
            const std::string m = "coded output: " + streams[i].why;
            return set_error(HPVCD_E_UNSUPPORTED, m.c_str());
        }
    return HPVCD_OK;
}

uint8_t* copy_out(const std::vector<uint8_t>& v, size_t* size) {
    uint8_t* p = (uint8_t*)std::malloc(v.empty() ? 1 : v.size());
    if (!p) return nullptr;
    if (!v.empty()) std::memcpy(p, v.data(), v.size());
    *size = v.size();
    return p;
}

}  // namespace hpvcd

using namespace hpvcd;

extern "C" hpvcd_status hpvcd_gen_hevc(const hpvcd_gen_config* cfg, uint8_t** data, size_t* size) {
    if (!cfg || !data || !size) return set_error(HPVCD_E_INVALID, "null argument");
    *data = nullptr;
    *size = 0;
    std::vector<PicRecords> recs;
    std::vector<PicStream> streams;
    if (hpvcd_status st = gen_coded(*cfg, recs, streams)) return st;
    const SeqPlan plan = gen_plan(*cfg);
    Sps sps;
    Pps pps;
    gen_parameter_sets(*cfg, sps, pps, &plan);
    std::vector<uint8_t> out;
    auto put = [&](const std::vector<uint8_t>& nal) {
        static const uint8_t sc[4] = {0, 0, 0, 1};
        out.insert(out.end(), sc, sc + 4);
        out.insert(out.end(), nal.begin(), nal.end());
    };
    put(write_vps(sps));
    put(write_sps(sps));
    put(write_pps(pps, sps));
    for (uint32_t i = 0; i < (uint32_t)streams.size(); i++)
        for (const auto& nal : gen_slice_nals(*cfg, sps, pps, i, streams[i], &plan)) put(nal);
    *data = copy_out(out, size);
    return *data ? HPVCD_OK : set_error(HPVCD_E_ALLOC, "coded output: out of memory");
}

extern "C" void hpvcd_bytes_free(uint8_t* data) { std::free(data); }
