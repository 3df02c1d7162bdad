// parse.cpp — CABAC syntax parse of intra pictures into reconstruction
// records.
//
// The decisions of the coding-tree walk (walker.cpp) are read from the slice
// data the way the reference decoder reads them: SAO (src/decode.rs:1507),
// split_cu_flag (:2540), cu_transquant_bypass_flag / part_mode / pcm_flag
// (:2585-2660) and the PCM samples with the engine re-initialised after them
// (:2768, src/cabac/engine.rs:313), intra luma and chroma modes (:3204,
// :3248), split_transform_flag and the CBFs (:3315-3478), cu_qp_delta
// (:3528), tu_residual_act_flag (:4229), cross_comp_pred (:4257),
// residual_coding (src/cabac/residual_full.rs:392), and the slice data loop
// with its end_of_slice_segment_flag and end_of_subset_one_bit
// (src/decode.rs:1319-1435).  Slice segment headers come from hevc_ps.cpp.
#include "parse.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <new>
#include <string>
#include <thread>

#include "cabac.h"
#include "internal.h"

namespace hpvcd {

namespace {

// k-th order Exp-Golomb, bypass bins (9.3.3.3); false past 31 prefix ones
bool dec_egk(CabacDec& d, int k, uint32_t& v) {
    uint64_t abs = 0;
    while (d.bypass()) {
        abs += 1ull << k;
        if (++k > 31) return false;
    }
    abs += d.bypass_bits(k);
    if (abs > 0xffffffffull) return false;
    v = (uint32_t)abs;
    return true;
}

class CabacSource : public Source {
public:
    CabacSource(const Pps& p, const PicParams& P) : pps_(p), P_(P) {}

    // sub_starts: RBSP offsets of the slice's substreams after the first,
    // from its entry points
    void set_slice(const uint8_t* data, size_t len, const SliceHeader* h, std::vector<size_t> sub_starts) {
        data_ = data;
        len_ = len;
        h_ = h;
        sub_starts_ = std::move(sub_starts);
        substream_ = 0;
    }
    bool bad() const { return bad_ || d_.failed(); }
    const char* why() const { return d_.failed() && !bad_ ? "slice data ends inside a coding tree unit" : why_; }

    void begin_slice(int slice_qp, int init_type) override {
        qp_ = slice_qp;
        init_type_ = init_type;
        c_.init(init_type, slice_qp);
        d_.start(data_, len_);
    }
    void begin_substream(int sync_row) override {
        c_.init(init_type_, qp_);
        if (sync_row >= 0 && (size_t)sync_row < saved_.size()) c_ = saved_[(size_t)sync_row];
        // the engine starts at the substream's entry point when the slice
        // header gives one, else where the previous substream ended
        // (src/decode.rs:1184-1196)
        const size_t k = substream_++;
        if (k < sub_starts_.size() && sub_starts_[k] <= len_) d_.start(data_ + sub_starts_[k], len_ - sub_starts_[k]);
        else d_.init_engine();
    }
    void wpp_save(int row) override {
        if (saved_.size() <= (size_t)row) saved_.resize((size_t)row + 1);
        saved_[(size_t)row] = c_;
    }
    int end_ctb(bool, bool last_of_tile) override {
        if (d_.terminate()) return 1;  // end_of_slice_segment_flag
        if (last_of_tile) {
            if (!d_.terminate()) fail("end_of_subset_one_bit is not one");
            d_.byte_align();
        }
        return 0;
    }
    void sao(hpvcd_sao& s, bool left_avail, bool up_avail, const hpvcd_sao* left, const hpvcd_sao* up) override {
        std::memset(&s, 0, sizeof(s));
        if (!h_->sao_luma && !h_->sao_chroma) return;
        if (left_avail && d_.decode(c_.s[CTX_SAO_MERGE])) {
            s = *left;
            return;
        }
        if (up_avail && d_.decode(c_.s[CTX_SAO_MERGE])) {
            s = *up;
            return;
        }
        const int ncomp = P_.chroma_idc ? 3 : 1;
        for (int c = 0; c < ncomp; c++) {
            if ((c == 0 && !h_->sao_luma) || (c > 0 && !h_->sao_chroma)) continue;
            int type;
            if (c < 2) type = d_.decode(c_.s[CTX_SAO_TYPE]) ? (d_.bypass() ? 2 : 1) : 0;
            else type = s.type_idx[1];
            s.type_idx[c] = (uint8_t)type;
            if (!type) continue;
            const int bd = c ? P_.bd_c : P_.bd;
            const int cmax = (1 << (std::min(bd, 10) - 5)) - 1;
            const int scale = c ? pps_.log2_sao_offset_scale_chroma : pps_.log2_sao_offset_scale_luma;
            int a[4];
            for (int k = 0; k < 4; k++) {
                a[k] = 0;
                while (a[k] < cmax && d_.bypass()) a[k]++;
            }
            if (type == 1) {
                for (int k = 0; k < 4; k++)
                    if (a[k] && d_.bypass()) a[k] = -a[k];
                s.band_pos[c] = (uint8_t)d_.bypass_bits(5);
            } else {
                a[2] = -a[2];
                a[3] = -a[3];
                s.eo_class[c] = c < 2 ? (uint8_t)d_.bypass_bits(2) : s.eo_class[1];
            }
            for (int k = 0; k < 4; k++) s.offsets[c][k] = (int16_t)(a[k] * (1 << scale));
        }
    }
    bool split_cu(int ctx_inc) override { return d_.decode(c_.s[CTX_SPLIT_CU + ctx_inc]) != 0; }
    bool tqb() override { return d_.decode(c_.s[CTX_TQB]) != 0; }
    bool intra_nxn() override { return d_.decode(c_.s[CTX_PART_MODE]) == 0; }
    bool pcm_flag() override {
        const bool v = d_.terminate() != 0;
        if (v) d_.byte_align();  // pcm_alignment_zero_bit
        return v;
    }
    void pcm_samples(int n_luma, int n_chroma, int bd_y, int bd_c, uint32_t* out) override {
        for (int i = 0; i < n_luma; i++) out[i] = d_.read_bits(bd_y);
        for (int i = 0; i < 2 * n_chroma; i++) out[n_luma + i] = d_.read_bits(bd_c);
        d_.init_engine();
    }
    void luma_modes(int npu, FnRef<void(int, uint8_t*)> mpm,
                    FnRef<void(int, uint8_t)> commit) override {
        int flag[4];
        for (int k = 0; k < npu; k++) flag[k] = d_.decode(c_.s[CTX_PREV_INTRA]);
        for (int k = 0; k < npu; k++) {
            uint8_t m[3];
            mpm(k, m);
            uint8_t mode;
            if (flag[k]) {
                const int idx = d_.bypass() ? (d_.bypass() ? 2 : 1) : 0;
                mode = m[idx];
            } else {  // 8.4.2: rem_intra_luma_pred_mode over the sorted candidates
                std::sort(m, m + 3);
                int v = (int)d_.bypass_bits(5);
                for (int i = 0; i < 3; i++)
                    if (v >= m[i]) v++;
                mode = (uint8_t)v;
            }
            commit(k, mode);
        }
    }
    int chroma_mode_syntax() override { return d_.decode(c_.s[CTX_CHROMA_MODE]) ? (int)d_.bypass_bits(2) : 4; }
    bool split_tf(int log2) override { return d_.decode(c_.s[CTX_SPLIT_TF + 5 - log2]) != 0; }
    bool cbf_chroma(int depth) override { return d_.decode(c_.s[CTX_CBF_CHROMA + depth]) != 0; }
    bool cbf_luma(int depth) override { return d_.decode(c_.s[CTX_CBF_LUMA + (depth == 0 ? 1 : 0)]) != 0; }
    bool act_flag() override { return d_.decode(c_.s[CTX_ACT]) != 0; }
    int qp_delta() override {
        int a = 0;
        while (a < 5 && d_.decode(c_.s[CTX_QP_DELTA + (a ? 1 : 0)])) a++;
        if (a == 5) {
            uint32_t suffix;
            if (!dec_egk(d_, 0, suffix) || suffix > 64) {
                fail("cu_qp_delta_abs out of range");
                return 0;
            }
            a += (int)suffix;
        }
        const int v = (a && d_.bypass()) ? -a : a;
        // 7.4.9.14: CuQpDeltaVal in [-(26 + QpBdOffsetY / 2), 25 + QpBdOffsetY / 2]
        const int half = 3 * (P_.bd - 8);
        if (v < -(26 + half) || v > 25 + half) {
            fail("CuQpDeltaVal out of range");
            return 0;
        }
        return v;
    }
    int res_scale(int c) override {
        int a = 0;
        while (a < 4 && d_.decode(c_.s[CTX_RES_SCALE + 4 * c + a])) a++;
        if (!a) return 0;
        const bool neg = d_.decode(c_.s[CTX_RES_SIGN + c]) != 0;
        return neg ? -(1 << (a - 1)) : (1 << (a - 1));
    }
    void residual(const ResidualParams& p, int32_t* levels, ResidualOut& r) override {
        residual_decode(d_, c_, p, levels, r);
        for (int i = 0; i < r.nnz; i++) {
            const int32_t v = levels[r.nz[i]];
            if (v > HPVCD_COEFF_LEVEL_MAX || v < HPVCD_COEFF_LEVEL_MIN) fail("coefficient level beyond the record range");
        }
    }
    // inter syntax (src/decode.rs:6290 cu_skip_flag, :6519 part_mode, :6571
    // merge_idx, :6771 inter_pred_idc, :6878 ref_idx, :6929 mvd_coding)
    bool cu_skip(int ctx_inc) override { return d_.decode(c_.s[CTX_SKIP + ctx_inc]) != 0; }
    bool pred_mode_intra() override { return d_.decode(c_.s[CTX_PRED_MODE]) != 0; }
    int part_mode_inter(int log2, bool min_size, bool amp) override {
        if (d_.decode(c_.s[CTX_PART_MODE])) return PART_2Nx2N;
        const bool horizontal = d_.decode(c_.s[CTX_PART_MODE + 1]) != 0;
        if (!min_size) {
            if (!amp || d_.decode(c_.s[CTX_PART_MODE + 3])) return horizontal ? PART_2NxN : PART_Nx2N;
            const bool far = d_.bypass() != 0;
            return horizontal ? (far ? PART_2NxnD : PART_2NxnU) : (far ? PART_nRx2N : PART_nLx2N);
        }
        if (horizontal) return PART_2NxN;
        if (log2 == 3) return PART_Nx2N;
        return d_.decode(c_.s[CTX_PART_MODE + 2]) ? PART_Nx2N : PART_NxN;
    }
    bool merge_flag() override { return d_.decode(c_.s[CTX_MERGE_FLAG]) != 0; }
    int merge_idx(int max_cand) override {
        if (!d_.decode(c_.s[CTX_MERGE_IDX])) return 0;
        int idx = 1;
        while (idx < max_cand - 1 && d_.bypass()) idx++;
        return idx;
    }
    int inter_pred_idc(int ct_depth, bool no_bi) override {
        if (!no_bi && d_.decode(c_.s[CTX_INTER_PRED + ct_depth])) return 2;
        return d_.decode(c_.s[CTX_INTER_PRED + 4]) ? 1 : 0;
    }
    int ref_idx(int num) override {
        if (!d_.decode(c_.s[CTX_REF_IDX])) return 0;
        if (num == 2) return 1;
        if (!d_.decode(c_.s[CTX_REF_IDX + 1])) return 1;
        int idx = 2;
        while (idx < num - 1 && d_.bypass()) idx++;
        return idx;
    }
    void mv_syntax(const PuGeom&, int, const int16_t (*)[2], bool mvd_zero, int16_t mvd[2], int& mvp_flag) override {
        mvd[0] = mvd[1] = 0;
        if (!mvd_zero) {
            bool g0[2], g1[2] = {false, false};
            for (int c = 0; c < 2; c++) g0[c] = d_.decode(c_.s[CTX_MVD]) != 0;
            for (int c = 0; c < 2; c++)
                if (g0[c]) g1[c] = d_.decode(c_.s[CTX_MVD + 1]) != 0;
            for (int c = 0; c < 2; c++) {
                if (!g0[c]) continue;
                uint32_t a = 1;
                if (g1[c]) {
                    uint32_t v;
                    if (!dec_egk(d_, 1, v) || v > 65534) {
                        fail("abs_mvd_minus2 out of range");
                        return;
                    }
                    a = v + 2;
                }
                // the value wraps to 16 bits as the reference's i16 cast does
                mvd[c] = (int16_t)(uint16_t)(d_.bypass() ? 0u - a : a);
            }
        }
        mvp_flag = d_.decode(c_.s[CTX_MVP_FLAG]);
    }
    bool rqt_root_cbf() override { return d_.decode(c_.s[CTX_RQT_ROOT_CBF]) != 0; }

private:
    void fail(const char* why) {
        if (!bad_) why_ = why;
        bad_ = true;
    }
    const Pps& pps_;
    const PicParams P_;
    const SliceHeader* h_ = nullptr;
    const uint8_t* data_ = nullptr;
    size_t len_ = 0;
    Contexts c_;
    std::vector<Contexts> saved_;  // WPP storage per CTB row
    std::vector<size_t> sub_starts_;
    size_t substream_ = 0;
    CabacDec d_;
    int qp_ = 26, init_type_ = 0;
    bool bad_ = false;
    const char* why_ = "";
};

PicParams params_of(const Sps& s, const Pps& p) {
    PicParams P;
    P.w = s.width;
    P.h = s.height;
    P.chroma_idc = s.chroma_format_idc;
    P.bd = s.bit_depth_luma;
    P.bd_c = s.bit_depth_chroma;
    P.log2_ctb = s.log2_ctb;
    P.log2_min_cb = s.log2_min_cb;
    P.log2_min_tb = s.log2_min_tb;
    P.log2_max_tb = s.log2_max_tb;
    P.max_depth_intra = s.max_th_depth_intra;
    P.max_depth_inter = s.max_th_depth_inter;
    P.tqb_enabled = p.transquant_bypass;
    P.transform_skip = p.transform_skip;
    P.log2_max_ts = p.log2_max_ts;
    P.implicit_rdpcm = s.implicit_rdpcm;
    P.explicit_rdpcm = s.explicit_rdpcm;
    P.cross_component = p.cross_component;
    P.extended_precision = s.extended_precision;
    P.persistent_rice = s.persistent_rice;
    P.ts_context = s.ts_context;
    P.bypass_alignment = s.bypass_alignment;
    P.sign_hiding = p.sign_hiding;
    P.act = p.act;
    P.act_qp[0] = p.act_y_qp_offset;
    P.act_qp[1] = p.act_cb_qp_offset;
    P.act_qp[2] = p.act_cr_qp_offset;
    P.pcm = s.pcm;
    P.pcm_bd = s.pcm_bit_depth_luma;
    P.pcm_bd_c = s.pcm_bit_depth_chroma;
    P.log2_min_pcm = s.log2_min_pcm_cb;
    P.log2_max_pcm = s.log2_max_pcm_cb;
    P.cu_qp_delta = p.cu_qp_delta;
    P.diff_cu_qp_delta_depth = p.diff_cu_qp_delta_depth;
    P.chroma_qp_off[0] = p.cb_qp_offset;
    P.chroma_qp_off[1] = p.cr_qp_offset;
    P.sao = s.sao;
    P.amp = s.amp;
    P.constrained_intra = p.constrained_intra;
    P.wpp = p.entropy_sync;
    return P;
}

// the back end's format range (validate_batch) and the syntax the walker reads
hpvcd_status check_params(const Sps& s, const Pps& p) {
    if (s.chroma_format_idc == 2) return set_error(HPVCD_E_UNSUPPORTED_CHROMA, "4:2:2 pictures are not supported");
    if (s.bit_depth_luma < 8 || s.bit_depth_luma > 12 || s.bit_depth_chroma < 8 || s.bit_depth_chroma > 12)
        return set_error(HPVCD_E_UNSUPPORTED_BITDEPTH, "bit depth must be 8..12");
    if ((s.bit_depth_luma > 8) != (s.bit_depth_chroma > 8))
        return set_error(HPVCD_E_UNSUPPORTED_BITDEPTH, "luma and chroma bit depths of different storage classes");
    if (s.log2_ctb < 4 || s.log2_ctb > 6 || s.log2_min_cb > s.log2_ctb || s.log2_min_tb >= s.log2_min_cb ||
        s.log2_max_tb > std::min(5, s.log2_ctb))
        return set_error(HPVCD_E_PARAMSET, "coding / transform block size limits");
    if (s.width % (1 << s.log2_min_cb) || s.height % (1 << s.log2_min_cb))
        return set_error(HPVCD_E_BAD_DIMENSIONS, "picture size not a multiple of the minimum coding block");
    if (s.width > 16384 || s.height > 16384) return set_error(HPVCD_E_LIMIT, "picture size");
    if (s.max_th_depth_intra > s.log2_ctb - s.log2_min_tb) return set_error(HPVCD_E_PARAMSET, "transform depth");
    if (s.pcm && (s.pcm_bit_depth_luma > s.bit_depth_luma || s.pcm_bit_depth_chroma > s.bit_depth_chroma ||
                  s.log2_min_pcm_cb < 3 || s.log2_max_pcm_cb > std::min(5, s.log2_ctb) ||
                  s.log2_min_pcm_cb > s.log2_max_pcm_cb))
        return set_error(HPVCD_E_PARAMSET, "PCM parameters");
    if (p.init_qp < -6 * (s.bit_depth_luma - 8) || p.init_qp > 51) return set_error(HPVCD_E_PARAMSET, "init_qp");
    if (p.cb_qp_offset < -12 || p.cb_qp_offset > 12 || p.cr_qp_offset < -12 || p.cr_qp_offset > 12)
        return set_error(HPVCD_E_PARAMSET, "pps chroma QP offsets");
    if (p.cu_qp_delta && p.diff_cu_qp_delta_depth > s.log2_ctb - s.log2_min_cb)
        return set_error(HPVCD_E_PARAMSET, "diff_cu_qp_delta_depth");
    if (p.transform_skip && (p.log2_max_ts < 2 || p.log2_max_ts > 5)) return set_error(HPVCD_E_PARAMSET, "log2_max_ts");
    if (p.log2_sao_offset_scale_luma > std::max(0, s.bit_depth_luma - 10) ||
        p.log2_sao_offset_scale_chroma > std::max(0, s.bit_depth_chroma - 10))
        return set_error(HPVCD_E_PARAMSET, "log2_sao_offset_scale");
    if (p.cross_component && s.chroma_format_idc != 3) return set_error(HPVCD_E_PARAMSET, "cross-component prediction");
    if (p.act && s.chroma_format_idc != 3) return set_error(HPVCD_E_PARAMSET, "adaptive colour transform");
    const int ctb = 1 << s.log2_ctb;
    const int cols = (s.width + ctb - 1) / ctb, rows = (s.height + ctb - 1) / ctb;
    if (p.tiles) {
        if (p.tile_cols > cols || p.tile_rows > rows || p.tile_cols * p.tile_rows > 255)
            return set_error(HPVCD_E_PARAMSET, "tile grid");
        if (!p.uniform_spacing) {
            int sc = 0, sr = 0;
            for (int v : p.col_widths) sc += v;
            for (int v : p.row_heights) sr += v;
            if (sc >= cols || sr >= rows) return set_error(HPVCD_E_PARAMSET, "tile column / row sizes");
        }
    }
    return HPVCD_OK;
}

}  // namespace

hpvcd_status parse_headers(const Sps& sps, const Pps& pps, const CodedPicture& cp, std::vector<SliceHeader>& hdr,
                           bool allow_inter) {
    const size_t nsl = cp.slices.size();
    if (nsl == 0) return set_error(HPVCD_E_BITSTREAM, "picture without slices");
    if (nsl > HPVCD_MAX_SLICES) return set_error(HPVCD_E_UNSUPPORTED, "more than 8 slices in a picture");
    hdr.assign(nsl, SliceHeader());
    for (size_t k = 0; k < nsl; k++) {
        const auto& rb = cp.slices[k];
        if (rb.size() < 3) return set_error(HPVCD_E_BITSTREAM, "slice NAL unit truncated");
        const uint8_t type = (rb[0] >> 1) & 63;
        if (int st = parse_slice_header(rb, sps, pps, type, hdr[k], allow_inter)) return (hpvcd_status)st;
        if ((k == 0) != hdr[k].first_slice) return set_error(HPVCD_E_BITSTREAM, "first_slice_segment_in_pic_flag");
    }
    return HPVCD_OK;
}

hpvcd_status parse_picture(const Sps& sps, const Pps& pps, const CodedPicture& cp, ParsedPicture& out,
                           const std::vector<SliceHeader>* headers, const std::vector<const InterSlice*>* inter,
                           int32_t poc) {
    if (hpvcd_status st = check_params(sps, pps)) return st;
    std::vector<SliceHeader> own;
    if (!headers) {
        if (hpvcd_status st = parse_headers(sps, pps, cp, own, false)) return st;
        headers = &own;
    }
    const std::vector<SliceHeader>& hdr = *headers;
    const size_t nsl = hdr.size();
    if (nsl != cp.slices.size() || (inter && inter->size() != nsl))
        return set_error(HPVCD_E_INVALID, "slice headers of another picture");
    // CTB layout: tile scan (6.5.1) and each CTB's slice
    const int ctb = 1 << sps.log2_ctb;
    const int cols = (sps.width + ctb - 1) / ctb, rows = (sps.height + ctb - 1) / ctb;
    const int nctb = cols * rows;
    std::vector<int> col_bd, row_bd;
    tile_boundaries(pps, cols, rows, col_bd, row_bd);
    std::vector<uint32_t> ts_to_rs;
    std::vector<int> rs_to_ts((size_t)nctb, 0);
    std::vector<uint8_t> tile_of((size_t)nctb, 0);
    const int ntc = (int)col_bd.size() - 1, ntr = (int)row_bd.size() - 1;
    for (int tj = 0; tj < ntr; tj++)
        for (int ti = 0; ti < ntc; ti++)
            for (int ry = row_bd[tj]; ry < row_bd[tj + 1]; ry++)
                for (int rx = col_bd[ti]; rx < col_bd[ti + 1]; rx++) {
                    rs_to_ts[(size_t)ry * cols + rx] = (int)ts_to_rs.size();
                    ts_to_rs.push_back((uint32_t)(ry * cols + rx));
                    tile_of[(size_t)ry * cols + rx] = (uint8_t)(tj * ntc + ti);
                }
    std::vector<int> start(nsl + 1, nctb);
    for (size_t k = 0; k < nsl; k++) {
        start[k] = rs_to_ts[(size_t)hdr[k].segment_address];
        if (k && start[k] <= start[k - 1]) return set_error(HPVCD_E_BITSTREAM, "slice segment addresses out of order");
    }
    std::vector<uint8_t> ctb_slice((size_t)nctb, 0);
    for (size_t k = 0; k < nsl; k++)
        for (int ts = start[k]; ts < start[k + 1]; ts++) ctb_slice[ts_to_rs[ts]] = (uint8_t)k;

    PicParams P = params_of(sps, pps);
    for (const SliceHeader& h : hdr) {
        P.sao_luma |= h.sao_luma;
        P.sao_chroma |= h.sao_chroma;
    }
    CabacSource src(pps, P);
    out.rec = PicRecords();
    PicWalker walker(P, src, out.rec, poc);
    walker.set_layout(ctb_slice, tile_of);
    for (size_t k = 0; k < nsl; k++) {
        const SliceHeader& h = hdr[k];
        walker.set_inter(inter ? (*inter)[k] : nullptr);
        const int qp = pps.init_qp + h.slice_qp_delta;
        if (qp < -6 * (sps.bit_depth_luma - 8) || qp > 51) return set_error(HPVCD_E_BITSTREAM, "SliceQpY");
        const int cb = pps.cb_qp_offset + h.cb_qp_offset, cr = pps.cr_qp_offset + h.cr_qp_offset;
        if (cb < -12 || cb > 12 || cr < -12 || cr > 12) return set_error(HPVCD_E_BITSTREAM, "slice chroma QP offsets");
        walker.set_chroma_offsets(cb, cr);
        const auto& rb = cp.slices[k];
        // entry points: byte sizes of the substreams in the NAL unit, where
        // emulation prevention bytes count; as RBSP offsets from the slice data
        std::vector<size_t> sub_starts;
        {
            const std::vector<uint32_t>& ep = cp.removed.size() > k ? cp.removed[k] : std::vector<uint32_t>();
            auto nal_of = [&](size_t r) {  // NAL position of RBSP byte r
                size_t n = r;
                for (size_t i = 0; i < ep.size() && (size_t)ep[i] - i <= r; i++) n++;
                return n;
            };
            auto rbsp_of = [&](size_t nal) {  // RBSP position of NAL byte nal
                size_t c = 0;
                while (c < ep.size() && ep[c] < nal) c++;
                return nal - c;
            };
            size_t at = nal_of(h.data_offset);
            for (uint32_t e : h.entry_points) {
                at += e;
                const size_t r = rbsp_of(at);
                sub_starts.push_back(r >= h.data_offset ? r - h.data_offset : 0);
            }
        }
        src.set_slice(rb.data() + h.data_offset, rb.size() - h.data_offset, &h, std::move(sub_starts));
        // substreams: one per tile, and with WPP one per CTB row of a tile
        auto sub_start = [&](int t) {
            const uint32_t r = ts_to_rs[t];
            if (t == 0 || tile_of[r] != tile_of[ts_to_rs[t - 1]]) return true;
            return pps.entropy_sync && (r % (uint32_t)cols == 0 || tile_of[r - 1] != tile_of[r]);
        };
        for (int ts = start[k]; ts < start[k + 1]; ts++) {
            const uint32_t rs = ts_to_rs[ts];
            const bool first_tile = ts == start[k] || sub_start(ts);
            const bool last_slice = ts == start[k + 1] - 1;
            const bool last_tile = ts == nctb - 1 || sub_start(ts + 1);
            const int eos = walker.ctb((int)(rs % cols), (int)(rs / cols), ts == start[k], first_tile, last_slice,
                                       last_tile, qp);
            if (src.bad()) return set_error(HPVCD_E_CABAC_DESYNC, src.why());
            if (walker.failed()) return set_error(HPVCD_E_BITSTREAM, "merge candidate beyond its reference list");
            if (eos != (last_slice ? 1 : 0))
                return set_error(HPVCD_E_CABAC_DESYNC, eos ? "slice ends before its last coding tree unit"
                                                          : "slice data continues past the next slice's address");
        }
    }
    walker.finish();

    hpvcd_picture& p = out.pic;
    std::memset(&p, 0, sizeof(p));
    p.width = sps.width;
    p.height = sps.height;
    p.chroma_idc = (uint8_t)sps.chroma_format_idc;
    p.bit_depth_luma = (uint8_t)sps.bit_depth_luma;
    p.bit_depth_chroma = (uint8_t)sps.bit_depth_chroma;
    p.log2_ctb = (uint8_t)sps.log2_ctb;
    uint32_t f = pps.lf_across_slices ? HPVCD_PF_LF_ACROSS_SLICES : 0u;
    if (sps.strong_intra_smoothing) f |= HPVCD_PF_STRONG_INTRA_SMOOTHING;
    if (sps.intra_smoothing_disabled) f |= HPVCD_PF_INTRA_SMOOTHING_DISABLED;
    if (sps.intra_boundary_filter_disabled) f |= HPVCD_PF_INTRA_BOUNDARY_FILTER_OFF;
    if (sps.extended_precision) f |= HPVCD_PF_EXTENDED_PRECISION;
    if (sps.implicit_rdpcm) f |= HPVCD_PF_IMPLICIT_RDPCM;
    if (sps.ts_rotation) f |= HPVCD_PF_TS_ROTATION;
    if (sps.sao && P.sao_luma) f |= HPVCD_PF_SAO_LUMA;
    if (sps.sao && P.sao_chroma && sps.chroma_format_idc) f |= HPVCD_PF_SAO_CHROMA;
    if (sps.full_range) f |= HPVCD_PF_FULL_RANGE;
    if (pps.deblocking_disabled) f |= HPVCD_PF_DEBLOCK_DISABLED;
    if (sps.pcm && sps.pcm_loop_filter_disabled) f |= HPVCD_PF_PCM_LF_DISABLED;
    if (pps.constrained_intra) f |= HPVCD_PF_CONSTRAINED_INTRA;
    if (ntc * ntr > 1) f |= HPVCD_PF_TILES | (pps.lf_across_tiles ? HPVCD_PF_LF_ACROSS_TILES : 0u);
    p.flags = f;
    p.pps_cb_qp_offset = (int8_t)pps.cb_qp_offset;
    p.pps_cr_qp_offset = (int8_t)pps.cr_qp_offset;
    p.matrix = (uint8_t)sps.matrix;
    p.num_slices = (uint8_t)nsl;
    p.log2_sao_offset_scale_luma = (int8_t)pps.log2_sao_offset_scale_luma;
    p.log2_sao_offset_scale_chroma = (int8_t)pps.log2_sao_offset_scale_chroma;
    for (size_t k = 0; k < nsl; k++) {
        hpvcd_slice_params& S = p.slices[k];
        S.disabled = hdr[k].deblocking_disabled ? 1 : 0;
        S.beta_offset_div2 = S.disabled ? 0 : (int8_t)hdr[k].beta_offset_div2;
        S.tc_offset_div2 = S.disabled ? 0 : (int8_t)hdr[k].tc_offset_div2;
        S.lf_across = hdr[k].lf_across_slices ? 1 : 0;
    }
    p.slice_qp = pps.init_qp + hdr[0].slice_qp_delta;
    // the visible window: the SPS conformance window
    p.crop_left = sps.conf_win[0];
    p.crop_top = sps.conf_win[2];
    p.win_w = sps.width - sps.conf_win[0] - sps.conf_win[1];
    p.win_h = sps.height - sps.conf_win[2] - sps.conf_win[3];
    if (p.win_w <= 0 || p.win_h <= 0) return set_error(HPVCD_E_BAD_DIMENSIONS, "conformance window leaves no picture");
    p.scaling_index = -1;
    out.scaling.clear();
    if (sps.scaling_list_enabled) {
        const ScalingLists& L = pps.scaling_list_data_present ? pps.scaling : sps.scaling;
        out.scaling.assign(&L.e[0][0][0], &L.e[0][0][0] + HPVCD_SCALING_SET_BYTES);
    }
    return HPVCD_OK;
}

hpvcd_status split_nals(const uint8_t* data, size_t size, int size_len,
                        std::vector<std::pair<const uint8_t*, size_t>>& out) {
    out.clear();
    if (size_len) {
        size_t pos = 0;
        while (pos < size) {
            if (size - pos < (size_t)size_len) return set_error(HPVCD_E_BITSTREAM, "NAL length field truncated");
            size_t n = 0;
            for (int i = 0; i < size_len; i++) n = (n << 8) | data[pos + i];
            pos += size_len;
            if (n > size - pos) return set_error(HPVCD_E_BITSTREAM, "NAL unit extends past the sample");
            if (n) out.emplace_back(data + pos, n);
            pos += n;
        }
        return HPVCD_OK;
    }
    // Annex B: start codes 00 00 01 (a leading zero byte of 00 00 00 01 is
    // trailing_zero_8bits of the previous unit)
    size_t i = 0, begin = SIZE_MAX;
    while (i + 2 < size) {
        if (data[i] == 0 && data[i + 1] == 0 && data[i + 2] == 1) {
            if (begin != SIZE_MAX) {
                size_t end = i;
                while (end > begin && data[end - 1] == 0) end--;
                if (end > begin) out.emplace_back(data + begin, end - begin);
            }
            i += 3;
            begin = i;
        } else {
            i++;
        }
    }
    if (begin == SIZE_MAX) return set_error(HPVCD_E_BITSTREAM, "no start code");
    size_t end = size;
    while (end > begin && data[end - 1] == 0) end--;
    if (end > begin) out.emplace_back(data + begin, end - begin);
    return HPVCD_OK;
}

hpvcd_status assemble_records(std::vector<ParsedPicture>& pics, const std::vector<hpvcd_output>& outputs,
                              hpvcd_records** out) {
    *out = nullptr;
    hpvcd_records* r = new (std::nothrow) hpvcd_records();
    if (!r) return set_error(HPVCD_E_ALLOC, "out of memory");
    r->outputs = outputs;
    // every array at its final size once: no reallocation while appending
    size_t ntb = 0, nco = 0, nsao = 0, ncell = 0, npu = 0, nref = 0;
    for (const ParsedPicture& pp : pics) {
        ntb += pp.rec.tbs.size();
        nco += pp.rec.coeffs.size();
        nsao += pp.rec.sao.size();
        ncell += pp.rec.cells.size();
        npu += pp.rec.pus.size();
        nref += pp.rec.refs.size();
    }
    r->pictures.reserve(pics.size());
    r->tbs.reserve(ntb);
    r->coeffs.reserve(nco);
    r->sao.reserve(nsao);
    r->cells.reserve(ncell);
    r->pus.reserve(npu);
    r->refs.reserve(nref);
    std::vector<const uint8_t*> sets;
    for (ParsedPicture& pp : pics) {
        hpvcd_picture p = pp.pic;
        p.scaling_index = -1;
        if (!pp.scaling.empty()) {
            int found = -1;
            for (size_t s = 0; s < sets.size() && found < 0; s++)
                if (std::memcmp(sets[s], pp.scaling.data(), HPVCD_SCALING_SET_BYTES) == 0) found = (int)s;
            if (found < 0) {
                found = (int)sets.size();
                sets.push_back(pp.scaling.data());
            }
            p.scaling_index = found;
        }
        PicRecords& t = pp.rec;
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
        r->tbs.insert(r->tbs.end(), t.tbs.begin(), t.tbs.end());
        r->coeffs.insert(r->coeffs.end(), t.coeffs.begin(), t.coeffs.end());
        r->sao.insert(r->sao.end(), t.sao.begin(), t.sao.end());
        r->cells.insert(r->cells.end(), t.cells.begin(), t.cells.end());
        r->pus.insert(r->pus.end(), t.pus.begin(), t.pus.end());
        r->refs.insert(r->refs.end(), t.refs.begin(), t.refs.end());
    }
    for (const uint8_t* s : sets) r->scaling.insert(r->scaling.end(), s, s + HPVCD_SCALING_SET_BYTES);
    *out = r;
    return HPVCD_OK;
}

}  // namespace hpvcd

using namespace hpvcd;

namespace hpvcd {
namespace {

// A decoded picture in the decoded picture buffer: its POC, batch index and
// reference marking.
struct DpbPic {
    int32_t poc;
    uint32_t index;
    bool ref;
};

// The reference pictures of a P / B slice (src/dpb.rs:217 apply_rps_lt,
// :298 build_ref_lists, :465 finalize_list): RefPicListTemp0 is
// StCurrBefore then StCurrAfter (pictures the DPB holds as references;
// others are left out), RefPicListTemp1 the reverse; each is repeated to
// max(num_ref_idx_active, its length) entries and then taken in order or
// through list_entry_lX.
hpvcd_status slice_lists(const SliceHeader& h, int32_t cur_poc, const std::vector<DpbPic>& dpb, InterSlice& s) {
    auto find = [&](int32_t poc) -> const DpbPic* {
        for (const DpbPic& d : dpb)
            if (d.ref && d.poc == poc) return &d;
        return nullptr;
    };
    std::vector<const DpbPic*> before, after;
    for (size_t k = 0; k < h.rps.d0.size(); k++)
        if (h.rps.u0[k])
            if (const DpbPic* d = find(cur_poc + h.rps.d0[k])) before.push_back(d);
    for (size_t k = 0; k < h.rps.d1.size(); k++)
        if (h.rps.u1[k])
            if (const DpbPic* d = find(cur_poc + h.rps.d1[k])) after.push_back(d);
    std::vector<const DpbPic*> temp[2];
    temp[0] = before;
    temp[0].insert(temp[0].end(), after.begin(), after.end());
    temp[1] = after;
    temp[1].insert(temp[1].end(), before.begin(), before.end());
    for (int l = 0; l < (h.slice_type == 0 ? 2 : 1); l++) {
        const int num = h.num_ref[l];
        if (temp[l].empty() || num == 0)
            return set_error(HPVCD_E_BITSTREAM, "reference picture list without pictures");
        std::vector<const DpbPic*> ext;
        const size_t len = std::max((size_t)num, temp[l].size());
        while (ext.size() < len)
            for (const DpbPic* d : temp[l])
                if (ext.size() < len) ext.push_back(d);
        for (int i = 0; i < num; i++) {
            size_t e = (size_t)i;
            if (h.list_mod[l]) {
                e = h.list_entry[l][(size_t)i];
                if (e >= ext.size()) return set_error(HPVCD_E_BITSTREAM, "list_entry beyond the reference pictures");
            }
            RefPic rp;
            rp.poc = ext[e]->poc;
            rp.pic = (int32_t)ext[e]->index;
            s.list[l].push_back(rp);
        }
    }
    return HPVCD_OK;
}

// The record rows of a slice's entries: the pred_weight_table values
// (src/inter.rs:159 PredWeightTable::parse, :60 normalize_luma_offset, :68
// derive_chroma_offset), or neutral rows of denominator 64 without one.
void slice_rows(const SliceHeader& h, const Sps& sps, const Pps& pps, InterSlice& s) {
    const bool weighted = slice_weighted(h, pps);
    const PredWeights& t = h.wp;
    for (int l = 0; l < 2; l++)
        for (size_t i = 0; i < s.list[l].size(); i++) {
            hpvcd_ref e;
            std::memset(&e, 0, sizeof(e));
            e.picture = (uint32_t)s.list[l][i].pic;
            e.weighted = weighted ? 1 : 0;
            e.log2_denom[0] = e.log2_denom[1] = 6;
            for (int k = 0; k < 3; k++) e.weight[k] = 64;
            if (weighted) {
                const WpEntry& w = t.e[l][i];
                e.log2_denom[0] = (uint8_t)t.luma_log2_denom;
                e.log2_denom[1] = (uint8_t)t.chroma_log2_denom;
                const int bd_y = sps.bit_depth_luma, bd_c = sps.bit_depth_chroma;
                const bool hp = sps.high_precision_offsets;
                const int range_y = hp ? 1 << (bd_y - 1) : 128, scale_y = hp ? 1 : 1 << (bd_y - 8);
                const int range_c = hp ? 1 << (bd_c - 1) : 128, scale_c = hp ? 1 : 1 << (bd_c - 8);
                e.weight[0] = (int16_t)((1 << t.luma_log2_denom) + (w.luma ? w.delta_weight : 0));
                e.offset[0] = (int16_t)(w.luma ? std::clamp(w.offset, -range_y, range_y - 1) * scale_y : 0);
                for (int j = 0; j < 2; j++) {
                    const int wc = (1 << t.chroma_log2_denom) + (w.chroma ? w.delta_cweight[j] : 0);
                    e.weight[1 + j] = (int16_t)wc;
                    int off = 0;
                    if (w.chroma) {
                        const int64_t pred = range_c - (((int64_t)range_c * wc) >> t.chroma_log2_denom);
                        off = (int)std::clamp<int64_t>(w.delta_coffset[j] + pred, -range_c, range_c - 1) * scale_c;
                    }
                    e.offset[1 + j] = (int16_t)off;
                }
            }
            s.rows.push_back(e);
        }
}

// 8.3.1 (src/rps.rs:208 derive_poc): the MSB from the previous TemporalId-0
// picture
int32_t derive_poc(int poc_lsb, int32_t prev, int log2_max) {
    const int32_t max = 1 << log2_max;
    const int32_t prev_lsb = ((prev % max) + max) % max, prev_msb = prev - prev_lsb;
    int32_t msb = prev_msb;
    if (poc_lsb < prev_lsb && prev_lsb - poc_lsb >= max / 2) msb = prev_msb + max;
    else if (poc_lsb > prev_lsb && poc_lsb - prev_lsb > max / 2) msb = prev_msb - max;
    return msb + poc_lsb;
}

}  // namespace
}  // namespace hpvcd

using namespace hpvcd;

// Annex B stream -> records, pictures in decode order, one output image per
// picture (its conformance window) in output order: by coded video
// sequence, then POC.  Kind `out_kind`, or when negative the kind
// decode_heic gives (RGB or luma by chroma format, 8 or 16 bit by depth).
// P and B pictures follow the reference's video decoder
// (src/video.rs:1045 decode_access_unit): POC, RPS marking, reference
// lists, weights and the collocated picture per slice; RASL pictures of a
// random-access point that starts the stream are skipped.
extern "C" hpvcd_status hpvcd_parse_hevc(const uint8_t* data, size_t size, int32_t out_kind, hpvcd_records** out) {
    if (!data || !out) return set_error(HPVCD_E_INVALID, "null argument");
    *out = nullptr;
    std::vector<std::pair<const uint8_t*, size_t>> nals;
    if (hpvcd_status st = split_nals(data, size, 0, nals)) return st;
    Sps sps;
    Pps pps;
    bool have_sps = false, have_pps = false;
    std::vector<CodedPicture> coded;
    std::vector<std::pair<Sps, Pps>> active;
    for (const auto& n : nals) {
        if (n.second < 2) return set_error(HPVCD_E_BITSTREAM, "NAL unit header truncated");
        const uint8_t type = (n.first[0] >> 1) & 63;
        if ((n.first[0] & 1) || (n.first[1] >> 3)) continue;  // nuh_layer_id > 0
        std::vector<uint32_t> removed;
        std::vector<uint8_t> rb = rbsp_from_nal(n.first, n.second, &removed);
        if (type == NAL_SPS) {
            if (int st = parse_sps(rb, sps)) return (hpvcd_status)st;
            have_sps = true;
        } else if (type == NAL_PPS) {
            if (!have_sps) return set_error(HPVCD_E_PARAMSET, "PPS before its SPS");
            if (int st = parse_pps(rb, sps, pps)) return (hpvcd_status)st;
            have_pps = true;
        } else if (type <= 21) {
            if (type > 9 && type < 16) return set_error(HPVCD_E_UNSUPPORTED, "reserved VCL NAL unit type");
            if (!have_pps) return set_error(HPVCD_E_PARAMSET, "slice before its parameter sets");
            if (rb.size() < 3) return set_error(HPVCD_E_BITSTREAM, "slice NAL unit truncated");
            const bool first = (rb[2] & 0x80) != 0;
            if (first || coded.empty()) {
                coded.emplace_back();
                active.emplace_back(sps, pps);
            }
            coded.back().slices.push_back(std::move(rb));
            coded.back().removed.push_back(std::move(removed));
        }
    }
    if (coded.empty()) return set_error(HPVCD_E_BITSTREAM, "no coded picture");
    // Phase 1, in decode order: slice headers, POC, RPS marking, reference
    // lists and weights of every picture (cheap, serial).  Phase 2: the
    // slice data of the pictures on host threads, a picture after the
    // pictures whose motion it reads as collocated motion.
    struct PicJob {
        size_t coded;
        int32_t poc;
        std::vector<SliceHeader> hdr;
        std::vector<InterSlice> slices;
        std::vector<int> col_of;   // per slice: the collocated batch picture, or -1
        int wave = 0;
    };
    std::vector<PicJob> jobs;
    std::vector<std::pair<int, int32_t>> out_key;   // (coded video sequence, POC) per picture
    std::vector<uint8_t> shown;
    std::vector<DpbPic> dpb;
    int32_t prev_tid0 = 0;
    int cvs = -1;
    bool rasl_skip = false;
    for (size_t i = 0; i < coded.size(); i++) {
        const Sps& S = active[i].first;
        const Pps& Q = active[i].second;
        PicJob job;
        job.coded = i;
        if (hpvcd_status st = parse_headers(S, Q, coded[i], job.hdr, true)) return st;
        const std::vector<SliceHeader>& hdr = job.hdr;
        const uint8_t nal_type = (coded[i].slices[0][0] >> 1) & 63;
        const uint8_t tid = (uint8_t)((coded[i].slices[0][1] & 7) - 1);
        const bool irap = nal_type >= 16 && nal_type <= 23, idr = nal_type == 19 || nal_type == 20;
        const bool rasl = nal_type == 8 || nal_type == 9, radl = nal_type == 6 || nal_type == 7;
        // NoRaslOutputFlag: IDR, BLA, or a CRA that starts the stream
        const bool no_rasl = irap && (nal_type < 21 || cvs < 0);
        if (rasl && rasl_skip) continue;
        int32_t poc;
        if (irap && no_rasl) {
            poc = idr ? 0 : hdr[0].poc_lsb;
            for (DpbPic& d : dpb) d.ref = false;
            cvs++;
            rasl_skip = true;
        } else {
            if (cvs < 0) return set_error(HPVCD_E_BITSTREAM, "stream does not start with a random-access picture");
            poc = derive_poc(hdr[0].poc_lsb, prev_tid0, S.log2_max_poc_lsb);
        }
        if (irap && !no_rasl) rasl_skip = false;
        // RPS marking (8.3.2): pictures outside the set are no longer references
        if (!idr) {
            for (DpbPic& d : dpb) {
                if (!d.ref) continue;
                bool keep = false;
                for (int32_t dp : hdr[0].rps.d0) keep |= d.poc == poc + dp;
                for (int32_t dp : hdr[0].rps.d1) keep |= d.poc == poc + dp;
                d.ref = keep;
            }
        }
        job.poc = poc;
        job.slices.assign(hdr.size(), InterSlice());
        job.col_of.assign(hdr.size(), -1);
        for (size_t k = 0; k < hdr.size(); k++) {
            const SliceHeader& h = hdr[k];
            if (h.slice_type == 2) continue;
            InterSlice& sl = job.slices[k];
            sl.slice_type = h.slice_type;
            sl.cabac_init = h.cabac_init;
            if (hpvcd_status st = slice_lists(h, poc, dpb, sl)) return st;
            slice_rows(h, S, Q, sl);
            sl.tmvp = h.tmvp;
            sl.col_from_l0 = h.col_from_l0;
            sl.col_ref_idx = h.col_ref_idx;
            if (h.tmvp) {
                const std::vector<RefPic>& l = sl.list[h.col_from_l0 ? 0 : 1];
                if ((size_t)h.col_ref_idx < l.size()) job.col_of[k] = l[(size_t)h.col_ref_idx].pic;
            }
            sl.max_merge = h.max_merge;
            sl.par_mrg_level = Q.log2_parallel_merge_level;
            sl.mvd_l1_zero = h.mvd_l1_zero;
        }
        const uint32_t index = (uint32_t)jobs.size();
        for (int c : job.col_of)
            if (c >= 0) job.wave = std::max(job.wave, jobs[(size_t)c].wave + 1);
        dpb.push_back(DpbPic{poc, index, true});
        // pictures that serve as references no more leave the bookkeeping
        dpb.erase(std::remove_if(dpb.begin(), dpb.end(), [](const DpbPic& d) { return !d.ref; }), dpb.end());
        out_key.emplace_back(cvs, poc);
        shown.push_back(hdr[0].pic_output ? 1 : 0);
        if (tid == 0 && !rasl && !radl && !(nal_type <= 14 && nal_type % 2 == 0)) prev_tid0 = poc;
        jobs.push_back(std::move(job));
    }
    std::vector<ParsedPicture> pics(jobs.size());
    for (PicJob& job : jobs)
        for (size_t k = 0; k < job.slices.size(); k++)
            if (job.col_of[k] >= 0) job.slices[k].col = &pics[(size_t)job.col_of[k]].rec.col;
    std::vector<hpvcd_status> sts(jobs.size(), HPVCD_OK);
    std::vector<std::string> why(jobs.size());
    int nwaves = 0;
    for (const PicJob& job : jobs) nwaves = std::max(nwaves, job.wave + 1);
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (int wv = 0; wv < nwaves; wv++) {
        std::vector<size_t> todo;
        for (size_t i = 0; i < jobs.size(); i++)
            if (jobs[i].wave == wv) todo.push_back(i);
        // each thread takes the wave's next unparsed picture (pictures of a
        // wave can differ much in size)
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t k = next++; k < todo.size(); k = next++) {
                PicJob& job = jobs[todo[k]];
                std::vector<const InterSlice*> inter(job.slices.size(), nullptr);
                for (size_t s2 = 0; s2 < job.slices.size(); s2++)
                    if (job.hdr[s2].slice_type != 2) inter[s2] = &job.slices[s2];
                const size_t c = job.coded;
                sts[todo[k]] = parse_picture(active[c].first, active[c].second, coded[c], pics[todo[k]], &job.hdr,
                                             &inter, job.poc);
                if (sts[todo[k]]) why[todo[k]] = hpvcd_last_error();
            }
        };
        const size_t nw = std::min<size_t>(nth, todo.size());
        if (nw <= 1) {
            work();
        } else {
            std::vector<std::thread> th;
            for (size_t t = 0; t < nw; t++) th.emplace_back(work);
            for (auto& x : th) x.join();
        }
    }
    for (size_t i = 0; i < jobs.size(); i++)
        if (sts[i]) return set_error(sts[i], why[i].c_str());
    // output order: coded video sequence, then POC
    std::vector<uint32_t> order;
    for (uint32_t i = 0; i < (uint32_t)pics.size(); i++)
        if (shown[i]) order.push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return out_key[a] < out_key[b]; });
    std::vector<hpvcd_output> outs;
    std::vector<uint32_t> rank(pics.size(), 0);
    for (uint32_t r = 0; r < (uint32_t)order.size(); r++) rank[order[r]] = r;
    for (size_t i = 0; i < pics.size(); i++) {
        hpvcd_picture& p = pics[i].pic;
        const bool mono = p.chroma_idc == 0, deep = p.bit_depth_luma > 8;
        const int kind = out_kind >= 0 ? out_kind
                                       : (mono ? (deep ? HPVCD_OUT_LUMA16 : HPVCD_OUT_LUMA8)
                                               : (deep ? HPVCD_OUT_RGB16 : HPVCD_OUT_RGB8));
        p.out_kind = (uint8_t)kind;
        p.dst_x = p.dst_y = 0;
        p.copy_w = shown[i] ? p.win_w : 0;
        p.copy_h = shown[i] ? p.win_h : 0;
        p.out_index = rank[i];
    }
    for (uint32_t r = 0; r < (uint32_t)order.size(); r++) {
        const hpvcd_picture& p = pics[order[r]].pic;
        const int kind = p.out_kind;
        hpvcd_output o;
        std::memset(&o, 0, sizeof(o));
        o.width = p.win_w;
        o.height = p.win_h;
        o.kind = (uint8_t)kind;
        o.channels = (kind == HPVCD_OUT_LUMA8 || kind == HPVCD_OUT_LUMA16) ? 1 : 3;
        o.bytes_per_sample = (kind == HPVCD_OUT_RGB16 || kind == HPVCD_OUT_LUMA16) ? 2 : 1;
        outs.push_back(o);
    }
    return assemble_records(pics, outs, out);
}
