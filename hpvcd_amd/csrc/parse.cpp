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
#include <cstring>
#include <new>

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

    void begin_slice(int slice_qp) override {
        qp_ = slice_qp;
        c_.init(0, slice_qp);
        d_.start(data_, len_);
    }
    void begin_substream(int sync_row) override {
        c_.init(0, qp_);
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
    void luma_modes(int npu, const std::function<void(int, uint8_t*)>& mpm,
                    const std::function<void(int, uint8_t)>& commit) override {
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
    int qp_ = 26;
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

hpvcd_status parse_picture(const Sps& sps, const Pps& pps, const CodedPicture& cp, ParsedPicture& out) {
    if (hpvcd_status st = check_params(sps, pps)) return st;
    const size_t nsl = cp.slices.size();
    if (nsl == 0) return set_error(HPVCD_E_BITSTREAM, "picture without slices");
    if (nsl > HPVCD_MAX_SLICES) return set_error(HPVCD_E_UNSUPPORTED, "more than 8 slices in a picture");
    std::vector<SliceHeader> hdr(nsl);
    for (size_t k = 0; k < nsl; k++) {
        const auto& rb = cp.slices[k];
        if (rb.size() < 3) return set_error(HPVCD_E_BITSTREAM, "slice NAL unit truncated");
        const uint8_t type = (rb[0] >> 1) & 63;
        if (int st = parse_slice_header(rb, sps, pps, type, hdr[k])) return (hpvcd_status)st;
        if ((k == 0) != hdr[k].first_slice) return set_error(HPVCD_E_BITSTREAM, "first_slice_segment_in_pic_flag");
    }
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
    PicWalker walker(P, src, out.rec);
    walker.set_layout(ctb_slice, tile_of);
    for (size_t k = 0; k < nsl; k++) {
        const SliceHeader& h = hdr[k];
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

// Annex B stream of intra pictures -> records, one output image per picture
// (its conformance window, kind `out_kind` or, when negative, the kind
// decode_heic gives: RGB or luma by chroma format, 8 or 16 bit by depth).
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
            if (type != 19 && type != 20 && type != 21 && type > 9)
                return set_error(HPVCD_E_UNSUPPORTED, "reserved VCL NAL unit type");
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
    std::vector<ParsedPicture> pics(coded.size());
    std::vector<hpvcd_output> outs;
    for (size_t i = 0; i < coded.size(); i++) {
        if (hpvcd_status st = parse_picture(active[i].first, active[i].second, coded[i], pics[i])) return st;
        hpvcd_picture& p = pics[i].pic;
        const bool mono = p.chroma_idc == 0, deep = p.bit_depth_luma > 8;
        const int kind = out_kind >= 0 ? out_kind
                                       : (mono ? (deep ? HPVCD_OUT_LUMA16 : HPVCD_OUT_LUMA8)
                                               : (deep ? HPVCD_OUT_RGB16 : HPVCD_OUT_RGB8));
        hpvcd_output o;
        std::memset(&o, 0, sizeof(o));
        o.width = p.win_w;
        o.height = p.win_h;
        o.kind = (uint8_t)kind;
        o.channels = (kind == HPVCD_OUT_LUMA8 || kind == HPVCD_OUT_LUMA16) ? 1 : 3;
        o.bytes_per_sample = (kind == HPVCD_OUT_RGB16 || kind == HPVCD_OUT_LUMA16) ? 2 : 1;
        outs.push_back(o);
        p.out_kind = (uint8_t)kind;
        p.dst_x = p.dst_y = 0;
        p.copy_w = p.win_w;
        p.copy_h = p.win_h;
        p.out_index = (uint32_t)i;
    }
    return assemble_records(pics, outs, out);
}
