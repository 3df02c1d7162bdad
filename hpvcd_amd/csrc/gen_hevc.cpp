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

    Rng* rng() override { return in_.rng(); }

    void begin_slice(int slice_qp) override {
        in_.begin_slice(slice_qp);
        slice_qp_ = slice_qp;
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
    void luma_modes(int npu, const std::function<void(int, uint8_t*)>& mpm,
                    const std::function<void(int, uint8_t)>& commit) override {
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

private:
    void open_substream() {
        c_.init(0, slice_qp_);
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
    int slice_qp_ = 26;
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

// Parameter sets of the generator's configuration (one SPS / PPS for every
// picture of it).
void gen_parameter_sets(const hpvcd_gen_config& c, Sps& s, Pps& p) {
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
}

// Slice segment NAL units (header bytes included, escaped) of one coded picture.
std::vector<std::vector<uint8_t>> gen_slice_nals(const hpvcd_gen_config& c, const Sps& sps, const Pps& pps,
                                                 uint32_t index, const PicStream& st) {
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
        write_slice_header(w, h, sps, pps, NAL_IDR_N_LP);
        for (const auto& b : subs) w.buf.insert(w.buf.end(), b.begin(), b.end());
        nals.push_back(nal_from_rbsp(NAL_IDR_N_LP, w.buf));
        while (ts < (int)ctb_slice.size() && ctb_slice[T.ts_to_rs[ts]] == k) ts++;
    }
    return nals;
}

// Every picture of the configuration: records and coded slices.
hpvcd_status gen_coded(const hpvcd_gen_config& c, std::vector<PicRecords>& recs, std::vector<PicStream>& streams) {
    if (hpvcd_status st = gen_validate(c)) return st;
    if (c.inter) return set_error(HPVCD_E_UNSUPPORTED, "coded output: inter pictures are records only");
    if (c.cu_qp_delta && c.qp_delta_max > 25)
        return set_error(HPVCD_E_PARAMSET, "coded output: CuQpDeltaVal beyond [-26, 25]");
    const uint32_t n = (uint32_t)(c.grid_cols * c.grid_rows);
    recs.assign(n, PicRecords());
    streams.assign(n, PicStream());
    const unsigned nth = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; t++)
        th.emplace_back([&, t]() {
            for (uint32_t i = t; i < n; i += nth) {
                HevcCoder coder;
                gen_picture(c, i, recs[i], &coder);
                streams[i] = std::move(coder.stream);
            }
        });
    for (auto& x : th) x.join();
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
    Sps sps;
    Pps pps;
    gen_parameter_sets(*cfg, sps, pps);
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
        for (const auto& nal : gen_slice_nals(*cfg, sps, pps, i, streams[i])) put(nal);
    *data = copy_out(out, size);
    return *data ? HPVCD_OK : set_error(HPVCD_E_ALLOC, "coded output: out of memory");
}

extern "C" void hpvcd_bytes_free(uint8_t* data) { std::free(data); }
