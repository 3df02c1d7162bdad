// hevc_ps.cpp — parameter sets and slice segment headers (see hevc_ps.h).
#include "hevc_ps.h"

#include <algorithm>
#include <cstring>

#include "../../include/hpvcd.h"
#include "internal.h"

namespace hpvcd {

// ---------------------------------------------------------------------------
// scaling lists (7.3.4; src/config.rs:469-563)
// ---------------------------------------------------------------------------
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

void ScalingLists::refresh_flat(int size_id, int m) {
    const int used = size_id == 0 ? 16 : 64;
    bool flat = size_id < 2 || e[size_id][m][64] == 16;
    for (int k = 0; k < used; k++) flat &= e[size_id][m][k] == 16;
    e[size_id][m][65] = flat ? 1 : 0;
}

void ScalingLists::set_default() {
    for (int s = 0; s < 4; s++)
        for (int m = 0; m < 6; m++) {
            std::memset(e[s][m], 16, 64);
            if (s > 0) std::memcpy(e[s][m], m < 3 ? kDefaultIntra8x8 : kDefaultInter8x8, 64);
            e[s][m][64] = 16;
            refresh_flat(s, m);
        }
}

// raster position of the k-th coefficient of the up-right diagonal scan of a size x size list
static int diag_pos(int size, int k) {
    int x = 0, y = 0, seen = 0;
    for (;;) {
        while (y >= 0) {
            if (x < size && y < size) {
                if (seen == k) return y * size + x;
                seen++;
            }
            y--;
            x++;
        }
        y = x;
        x = 0;
    }
}

static void write_scaling_list_data(BitWriter& w, const ScalingLists& L) {
    for (int s = 0; s < 4; s++)
        for (int m = 0; m < 6; m += (s == 3 ? 3 : 1)) {
            // every list coded explicitly (scaling_list_pred_mode_flag 1)
            w.flag(true);
            const int size = s == 0 ? 4 : 8, coefs = size * size;
            int next = 8;
            if (s > 1) {
                w.se((int)L.e[s][m][64] - 8);
                next = L.e[s][m][64];
            }
            for (int k = 0; k < coefs; k++) {
                const int v = L.e[s][m][diag_pos(size, k)];
                int d = v - next;
                if (d > 127) d -= 256;
                if (d < -128) d += 256;
                w.se(d);
                next = v;
            }
        }
}

static int parse_scaling_list_data(BitReader& r, ScalingLists& L) {
    L.set_default();
    for (int s = 0; s < 4; s++)
        for (int m = 0; m < 6; m += (s == 3 ? 3 : 1)) {
            if (!r.flag()) {
                const uint32_t delta = r.ue();
                if (delta == 0) {
                    std::memset(L.e[s][m], 16, 64);
                    if (s > 0) std::memcpy(L.e[s][m], m < 3 ? kDefaultIntra8x8 : kDefaultInter8x8, 64);
                    L.e[s][m][64] = 16;
                    L.refresh_flat(s, m);
                } else {
                    const int ref = m - (int)delta * (s == 3 ? 3 : 1);
                    if (ref < 0) return set_error(HPVCD_E_PARAMSET, "scaling_list_pred_matrix_id_delta");
                    std::memcpy(L.e[s][m], L.e[s][ref], 66);
                }
            } else {
                const int size = s == 0 ? 4 : 8, coefs = size * size;
                int next = 8;
                if (s > 1) {
                    next = r.se() + 8;
                    if (next < 0 || next > 255) return set_error(HPVCD_E_PARAMSET, "scaling_list_dc_coef");
                    L.e[s][m][64] = (uint8_t)next;
                } else {
                    L.e[s][m][64] = 16;
                }
                for (int k = 0; k < coefs; k++) {
                    next = (next + r.se() + 256) & 255;
                    L.e[s][m][diag_pos(size, k)] = (uint8_t)next;
                }
                L.refresh_flat(s, m);
            }
        }
    for (int m : {1, 2, 4, 5}) std::memcpy(L.e[3][m], L.e[2][m], 66);
    return r.overrun ? set_error(HPVCD_E_BITSTREAM, "scaling_list_data truncated") : 0;
}

// ---------------------------------------------------------------------------
// tiles (6.5.1)
// ---------------------------------------------------------------------------
void tile_boundaries(const Pps& pps, int ctb_cols, int ctb_rows, std::vector<int>& col_bd, std::vector<int>& row_bd) {
    col_bd.assign(1, 0);
    row_bd.assign(1, 0);
    const int nc = pps.tiles ? pps.tile_cols : 1, nr = pps.tiles ? pps.tile_rows : 1;
    for (int i = 0; i < nc; i++) {
        int wdt;
        if (!pps.tiles || pps.uniform_spacing) wdt = ((i + 1) * ctb_cols) / nc - (i * ctb_cols) / nc;
        else wdt = i < nc - 1 ? pps.col_widths[i] : ctb_cols - col_bd.back();
        col_bd.push_back(col_bd.back() + wdt);
    }
    for (int j = 0; j < nr; j++) {
        int hgt;
        if (!pps.tiles || pps.uniform_spacing) hgt = ((j + 1) * ctb_rows) / nr - (j * ctb_rows) / nr;
        else hgt = j < nr - 1 ? pps.row_heights[j] : ctb_rows - row_bd.back();
        row_bd.push_back(row_bd.back() + hgt);
    }
}

// ---------------------------------------------------------------------------
// writers
// ---------------------------------------------------------------------------
static int profile_idc(const Sps& s) {
    if (s.chroma_format_idc == 1 && s.bit_depth_luma == 8 && s.bit_depth_chroma == 8) return 1;   // Main
    if (s.chroma_format_idc == 1 && s.bit_depth_luma <= 10 && s.bit_depth_chroma <= 10) return 2;  // Main 10
    return 4;                                                                                      // RExt
}

static void write_ptl(BitWriter& w, const Sps& s) {
    w.bits(0, 2);                     // general_profile_space
    w.bit(0);                         // general_tier_flag
    const int p = profile_idc(s);
    w.bits((uint32_t)p, 5);
    for (int j = 0; j < 32; j++) w.bit(j == p ? 1 : 0);   // general_profile_compatibility_flag
    w.bit(1);                         // progressive_source_flag
    w.bit(0);                         // interlaced_source_flag
    w.bit(0);                         // non_packed_constraint_flag
    w.bit(1);                         // frame_only_constraint_flag
    w.bits(0, 32);                    // 43 reserved / constraint bits
    w.bits(0, 11);
    w.bit(0);                         // general_inbld_flag
    w.bits(120, 8);                   // general_level_idc (level 4)
}

static void skip_ptl(BitReader& r, int max_sub_layers_minus1) {
    r.bits(8);
    r.bits(32);
    r.bits(32);
    r.bits(16);
    r.bits(8);  // general_level_idc
    std::vector<int> prof(max_sub_layers_minus1), lev(max_sub_layers_minus1);
    for (int i = 0; i < max_sub_layers_minus1; i++) {
        prof[i] = (int)r.bit();
        lev[i] = (int)r.bit();
    }
    if (max_sub_layers_minus1 > 0)
        for (int i = max_sub_layers_minus1; i < 8; i++) r.bits(2);
    for (int i = 0; i < max_sub_layers_minus1; i++) {
        if (prof[i]) {
            r.bits(8);
            r.bits(32);
            r.bits(32);
            r.bits(16);
        }
        if (lev[i]) r.bits(8);
    }
}

std::vector<uint8_t> write_vps(const Sps& sps) {
    BitWriter w;
    w.bits(0, 4);        // vps_video_parameter_set_id
    w.bit(1);            // vps_base_layer_internal_flag
    w.bit(1);            // vps_base_layer_available_flag
    w.bits(0, 6);        // vps_max_layers_minus1
    w.bits(0, 3);        // vps_max_sub_layers_minus1
    w.bit(1);            // vps_temporal_id_nesting_flag
    w.bits(0xffff, 16);  // vps_reserved_0xffff_16bits
    write_ptl(w, sps);
    w.bit(0);            // vps_sub_layer_ordering_info_present_flag
    w.ue(4);             // vps_max_dec_pic_buffering_minus1
    w.ue(0);             // vps_max_num_reorder_pics
    w.ue(0);             // vps_max_latency_increase_plus1
    w.bits(0, 6);        // vps_max_layer_id
    w.ue(0);             // vps_num_layer_sets_minus1
    w.bit(0);            // vps_timing_info_present_flag
    w.bit(0);            // vps_extension_flag
    w.align_one_zero();
    return nal_from_rbsp(NAL_VPS, w.buf);
}

std::vector<uint8_t> write_sps(const Sps& s) {
    BitWriter w;
    w.bits(0, 4);   // sps_video_parameter_set_id
    w.bits(0, 3);   // sps_max_sub_layers_minus1
    w.bit(1);       // sps_temporal_id_nesting_flag
    write_ptl(w, s);
    w.ue((uint32_t)s.sps_id);
    w.ue((uint32_t)s.chroma_format_idc);
    if (s.chroma_format_idc == 3) w.bit(0);  // separate_colour_plane_flag
    w.ue((uint32_t)s.width);
    w.ue((uint32_t)s.height);
    const bool win = s.conf_win[0] || s.conf_win[1] || s.conf_win[2] || s.conf_win[3];
    w.flag(win);
    if (win) {
        const int cu_x = (s.chroma_format_idc == 1 || s.chroma_format_idc == 2) ? 2 : 1;
        const int cu_y = s.chroma_format_idc == 1 ? 2 : 1;
        w.ue((uint32_t)(s.conf_win[0] / cu_x));
        w.ue((uint32_t)(s.conf_win[1] / cu_x));
        w.ue((uint32_t)(s.conf_win[2] / cu_y));
        w.ue((uint32_t)(s.conf_win[3] / cu_y));
    }
    w.ue((uint32_t)(s.bit_depth_luma - 8));
    w.ue((uint32_t)(s.bit_depth_chroma - 8));
    w.ue((uint32_t)(s.log2_max_poc_lsb - 4));
    w.bit(0);       // sps_sub_layer_ordering_info_present_flag
    w.ue((uint32_t)(s.max_dec_pic_buffering - 1));
    w.ue((uint32_t)s.max_num_reorder);
    w.ue(0);        // sps_max_latency_increase_plus1
    w.ue((uint32_t)(s.log2_min_cb - 3));
    w.ue((uint32_t)(s.log2_ctb - s.log2_min_cb));
    w.ue((uint32_t)(s.log2_min_tb - 2));
    w.ue((uint32_t)(s.log2_max_tb - s.log2_min_tb));
    w.ue((uint32_t)s.max_th_depth_inter);
    w.ue((uint32_t)s.max_th_depth_intra);
    w.flag(s.scaling_list_enabled);
    if (s.scaling_list_enabled) {
        w.flag(s.scaling_list_data_present);
        if (s.scaling_list_data_present) write_scaling_list_data(w, s.scaling);
    }
    w.flag(s.amp);
    w.flag(s.sao);
    w.flag(s.pcm);
    if (s.pcm) {
        w.bits((uint32_t)(s.pcm_bit_depth_luma - 1), 4);
        w.bits((uint32_t)(s.pcm_bit_depth_chroma - 1), 4);
        w.ue((uint32_t)(s.log2_min_pcm_cb - 3));
        w.ue((uint32_t)(s.log2_max_pcm_cb - s.log2_min_pcm_cb));
        w.flag(s.pcm_loop_filter_disabled);
    }
    w.ue((uint32_t)s.st_rps.size());
    for (size_t i = 0; i < s.st_rps.size(); i++) write_st_rps(w, (int)i, (int)s.st_rps.size(), s.st_rps, s.st_rps[i]);
    w.flag(s.long_term_refs);
    if (s.long_term_refs) w.ue(0);  // num_long_term_ref_pics_sps
    w.flag(s.temporal_mvp);
    w.flag(s.strong_intra_smoothing);
    w.flag(s.vui);
    if (s.vui) {
        w.bit(0);  // aspect_ratio_info_present_flag
        w.bit(0);  // overscan_info_present_flag
        w.bit(1);  // video_signal_type_present_flag
        w.bits(5, 3);  // video_format: unspecified
        w.flag(s.full_range);
        w.flag(s.colour_description);
        if (s.colour_description) {
            w.bits((uint32_t)s.colour_primaries, 8);
            w.bits((uint32_t)s.transfer, 8);
            w.bits((uint32_t)s.matrix, 8);
        }
        w.bit(0);  // chroma_loc_info_present_flag
        w.bit(0);  // neutral_chroma_indication_flag
        w.bit(0);  // field_seq_flag
        w.bit(0);  // frame_field_info_present_flag
        w.bit(0);  // default_display_window_flag
        w.bit(0);  // vui_timing_info_present_flag
        w.bit(0);  // bitstream_restriction_flag
    }
    const bool range_ext = s.ts_rotation || s.ts_context || s.implicit_rdpcm || s.explicit_rdpcm ||
                           s.extended_precision || s.intra_smoothing_disabled || s.high_precision_offsets ||
                           s.persistent_rice || s.bypass_alignment;
    w.flag(range_ext);  // sps_extension_present_flag
    if (range_ext) {
        w.bit(1);       // sps_range_extension_flag
        w.bits(0, 7);   // multilayer, 3d, scc, 4bits
        w.flag(s.ts_rotation);
        w.flag(s.ts_context);
        w.flag(s.implicit_rdpcm);
        w.flag(s.explicit_rdpcm);
        w.flag(s.extended_precision);
        w.flag(s.intra_smoothing_disabled);
        w.flag(s.high_precision_offsets);
        w.flag(s.persistent_rice);
        w.flag(s.bypass_alignment);
    }
    w.align_one_zero();
    return nal_from_rbsp(NAL_SPS, w.buf);
}

std::vector<uint8_t> write_pps(const Pps& p, const Sps& s) {
    (void)s;
    BitWriter w;
    w.ue((uint32_t)p.pps_id);
    w.ue((uint32_t)p.sps_id);
    w.flag(p.dependent_slices);
    w.flag(p.output_flag_present);
    w.bits((uint32_t)p.num_extra_slice_header_bits, 3);
    w.flag(p.sign_hiding);
    w.flag(p.cabac_init_present);
    w.ue((uint32_t)(p.num_ref_idx_default[0] - 1));
    w.ue((uint32_t)(p.num_ref_idx_default[1] - 1));
    w.se(p.init_qp - 26);
    w.flag(p.constrained_intra);
    w.flag(p.transform_skip);
    w.flag(p.cu_qp_delta);
    if (p.cu_qp_delta) w.ue((uint32_t)p.diff_cu_qp_delta_depth);
    w.se(p.cb_qp_offset);
    w.se(p.cr_qp_offset);
    w.flag(p.slice_chroma_qp_offsets_present);
    w.flag(p.weighted_pred);
    w.flag(p.weighted_bipred);
    w.flag(p.transquant_bypass);
    w.flag(p.tiles);
    w.flag(p.entropy_sync);
    if (p.tiles) {
        w.ue((uint32_t)(p.tile_cols - 1));
        w.ue((uint32_t)(p.tile_rows - 1));
        w.flag(p.uniform_spacing);
        if (!p.uniform_spacing) {
            for (int i = 0; i < p.tile_cols - 1; i++) w.ue((uint32_t)(p.col_widths[i] - 1));
            for (int j = 0; j < p.tile_rows - 1; j++) w.ue((uint32_t)(p.row_heights[j] - 1));
        }
        w.flag(p.lf_across_tiles);
    }
    w.flag(p.lf_across_slices);
    w.flag(p.deblocking_control);
    if (p.deblocking_control) {
        w.flag(p.deblocking_override);
        w.flag(p.deblocking_disabled);
        if (!p.deblocking_disabled) {
            w.se(p.beta_offset_div2);
            w.se(p.tc_offset_div2);
        }
    }
    w.flag(p.scaling_list_data_present);
    if (p.scaling_list_data_present) write_scaling_list_data(w, p.scaling);
    w.flag(p.lists_modification);
    w.ue((uint32_t)(p.log2_parallel_merge_level - 2));
    w.flag(p.slice_header_extension);
    const bool range_ext = (p.transform_skip && p.log2_max_ts != 2) || p.cross_component || p.chroma_qp_offset_list ||
                           p.log2_sao_offset_scale_luma || p.log2_sao_offset_scale_chroma;
    const bool scc_ext = p.act;
    w.flag(range_ext || scc_ext);
    if (range_ext || scc_ext) {
        w.flag(range_ext);
        w.bit(0);   // pps_multilayer_extension_flag
        w.bit(0);   // pps_3d_extension_flag
        w.flag(scc_ext);
        w.bits(0, 4);
        if (range_ext) {
            if (p.transform_skip) w.ue((uint32_t)(p.log2_max_ts - 2));
            w.flag(p.cross_component);
            w.flag(false);  // chroma_qp_offset_list_enabled_flag
            w.ue((uint32_t)p.log2_sao_offset_scale_luma);
            w.ue((uint32_t)p.log2_sao_offset_scale_chroma);
        }
        if (scc_ext) {
            w.bit(0);   // pps_curr_pic_ref_enabled_flag
            w.flag(p.act);
            if (p.act) {
                w.flag(p.slice_act_qp_offsets_present);
                w.se(p.act_y_qp_offset + 5);
                w.se(p.act_cb_qp_offset + 5);
                w.se(p.act_cr_qp_offset + 3);
            }
            w.bit(0);   // pps_palette_predictor_initializers_present_flag
        }
    }
    w.align_one_zero();
    return nal_from_rbsp(NAL_PPS, w.buf);
}

static int ceil_log2(int n) {
    int b = 0;
    while ((1 << b) < n) b++;
    return b;
}

// 7.4.8 (src/rps.rs:148): the new S0 takes, in decreasing POC, the
// reference's S1 entries (reversed), the reference picture itself and the
// reference's S0 entries that land before the current picture; S1 the
// mirror image.  Candidates are indexed S0 entries, S1 entries, then the
// reference picture.
StRps derive_predicted_rps(const StRps& ref, int delta_rps, const std::vector<uint8_t>& used,
                           const std::vector<uint8_t>& use_delta) {
    StRps out;
    const int nn = (int)ref.d0.size(), np = (int)ref.d1.size(), n = nn + np;
    auto take = [&](int idx) { return idx < (int)use_delta.size() && use_delta[(size_t)idx]; };
    auto u = [&](int idx) { return (uint8_t)(idx < (int)used.size() ? used[(size_t)idx] : 0); };
    for (int j = np - 1; j >= 0; j--) {
        const int d = ref.d1[(size_t)j] + delta_rps;
        if (d < 0 && take(nn + j)) { out.d0.push_back(d); out.u0.push_back(u(nn + j)); }
    }
    if (delta_rps < 0 && take(n)) { out.d0.push_back(delta_rps); out.u0.push_back(u(n)); }
    for (int j = 0; j < nn; j++) {
        const int d = ref.d0[(size_t)j] + delta_rps;
        if (d < 0 && take(j)) { out.d0.push_back(d); out.u0.push_back(u(j)); }
    }
    for (int j = nn - 1; j >= 0; j--) {
        const int d = ref.d0[(size_t)j] + delta_rps;
        if (d > 0 && take(j)) { out.d1.push_back(d); out.u1.push_back(u(j)); }
    }
    if (delta_rps > 0 && take(n)) { out.d1.push_back(delta_rps); out.u1.push_back(u(n)); }
    for (int j = 0; j < np; j++) {
        const int d = ref.d1[(size_t)j] + delta_rps;
        if (d > 0 && take(nn + j)) { out.d1.push_back(d); out.u1.push_back(u(nn + j)); }
    }
    return out;
}

int parse_st_rps(BitReader& r, int idx, int num_sets, const std::vector<StRps>& sets, StRps& out) {
    out = StRps();
    const bool predicted = idx != 0 && r.flag();
    if (predicted) {
        int delta_idx = 1;
        if (idx == num_sets) delta_idx = (int)r.ue() + 1;
        if (delta_idx > idx || idx - delta_idx >= (int)sets.size())
            return set_error(HPVCD_E_BITSTREAM, "delta_idx_minus1 of a predicted RPS");
        const StRps& ref = sets[(size_t)(idx - delta_idx)];
        const bool neg = r.flag();
        const uint32_t abs_minus1 = r.ue();
        if (abs_minus1 >= 32768) return set_error(HPVCD_E_BITSTREAM, "abs_delta_rps_minus1");
        const int delta_rps = neg ? -(int)(abs_minus1 + 1) : (int)(abs_minus1 + 1);
        const int n = ref.num();
        std::vector<uint8_t> used((size_t)n + 1), use_delta((size_t)n + 1, 1);
        for (int j = 0; j <= n; j++) {
            used[(size_t)j] = r.flag() ? 1 : 0;
            if (!used[(size_t)j]) use_delta[(size_t)j] = r.flag() ? 1 : 0;
        }
        out = derive_predicted_rps(ref, delta_rps, used, use_delta);
    } else {
        const uint32_t nn = r.ue(), np = r.ue();
        if (nn > 16 || np > 16) return set_error(HPVCD_E_BITSTREAM, "short-term RPS size");
        int prev = 0;
        for (uint32_t i = 0; i < nn; i++) {
            prev -= (int)r.ue() + 1;
            out.d0.push_back(prev);
            out.u0.push_back(r.flag() ? 1 : 0);
        }
        prev = 0;
        for (uint32_t i = 0; i < np; i++) {
            prev += (int)r.ue() + 1;
            out.d1.push_back(prev);
            out.u1.push_back(r.flag() ? 1 : 0);
        }
    }
    if (r.overrun) return set_error(HPVCD_E_BITSTREAM, "short-term RPS truncated");
    if (out.num() > 16) return set_error(HPVCD_E_BITSTREAM, "short-term RPS size");
    return 0;
}

void write_st_rps(BitWriter& w, int idx, int num_sets, const std::vector<StRps>& sets, const StRps& rps) {
    if (idx != 0) {
        // prediction from the set just before (delta_idx_minus1 = 0): the
        // first delta POC that makes every entry of `rps` derivable
        const StRps& ref = sets[(size_t)idx - 1];
        const int n = ref.num();
        for (int mag = 1; mag <= 64; mag++)
            for (int sg = 0; sg < 2; sg++) {
                const int delta = sg ? -mag : mag;
                std::vector<uint8_t> used((size_t)n + 1, 0), use((size_t)n + 1, 0);
                for (int j = 0; j <= n; j++) {
                    const int d = j == n ? delta : (j < (int)ref.d0.size() ? ref.d0[(size_t)j] : ref.d1[(size_t)j - ref.d0.size()]) + delta;
                    for (size_t k = 0; k < rps.d0.size(); k++)
                        if (rps.d0[k] == d) { use[(size_t)j] = 1; used[(size_t)j] = rps.u0[k]; }
                    for (size_t k = 0; k < rps.d1.size(); k++)
                        if (rps.d1[k] == d) { use[(size_t)j] = 1; used[(size_t)j] = rps.u1[k]; }
                }
                if (!(derive_predicted_rps(ref, delta, used, use) == rps)) continue;
                w.bit(1);  // inter_ref_pic_set_prediction_flag
                if (idx == num_sets) w.ue(0);  // delta_idx_minus1
                w.flag(delta < 0);
                w.ue((uint32_t)(mag - 1));
                for (int j = 0; j <= n; j++) {
                    w.flag(used[(size_t)j] != 0);
                    if (!used[(size_t)j]) w.flag(use[(size_t)j] != 0);
                }
                return;
            }
        w.bit(0);
    }
    w.ue((uint32_t)rps.d0.size());
    w.ue((uint32_t)rps.d1.size());
    int prev = 0;
    for (size_t i = 0; i < rps.d0.size(); i++) {
        w.ue((uint32_t)(prev - rps.d0[i] - 1));
        prev = rps.d0[i];
        w.flag(rps.u0[i] != 0);
    }
    prev = 0;
    for (size_t i = 0; i < rps.d1.size(); i++) {
        w.ue((uint32_t)(rps.d1[i] - prev - 1));
        prev = rps.d1[i];
        w.flag(rps.u1[i] != 0);
    }
}

int num_pic_total_curr(const SliceHeader& h) {
    int n = 0;
    for (uint8_t u : h.rps.u0) n += u;
    for (uint8_t u : h.rps.u1) n += u;
    return n;
}

bool slice_weighted(const SliceHeader& h, const Pps& p) {
    return (h.slice_type == 1 && p.weighted_pred) || (h.slice_type == 0 && p.weighted_bipred);
}

static void write_pred_weights(BitWriter& w, const SliceHeader& h, const Sps& s) {
    const PredWeights& t = h.wp;
    w.ue((uint32_t)t.luma_log2_denom);
    if (s.chroma_format_idc) w.se(t.chroma_log2_denom - t.luma_log2_denom);
    for (int l = 0; l < (h.slice_type == 0 ? 2 : 1); l++) {
        const int n = h.num_ref[l];
        for (int i = 0; i < n; i++) w.flag(t.e[l][(size_t)i].luma);
        if (s.chroma_format_idc)
            for (int i = 0; i < n; i++) w.flag(t.e[l][(size_t)i].chroma);
        for (int i = 0; i < n; i++) {
            const WpEntry& e = t.e[l][(size_t)i];
            if (e.luma) {
                w.se(e.delta_weight);
                w.se(e.offset);
            }
            if (e.chroma)
                for (int j = 0; j < 2; j++) {
                    w.se(e.delta_cweight[j]);
                    w.se(e.delta_coffset[j]);
                }
        }
    }
}

static int parse_pred_weights(BitReader& r, SliceHeader& h, const Sps& s) {
    PredWeights& t = h.wp;
    t = PredWeights();
    t.luma_log2_denom = (int)r.ue();
    if (t.luma_log2_denom > 7) return set_error(HPVCD_E_BITSTREAM, "luma_log2_weight_denom");
    t.chroma_log2_denom = 0;
    if (s.chroma_format_idc) t.chroma_log2_denom = std::clamp(t.luma_log2_denom + r.se(), 0, 7);
    for (int l = 0; l < (h.slice_type == 0 ? 2 : 1); l++) {
        const int n = h.num_ref[l];
        t.e[l].assign((size_t)n, WpEntry());
        for (int i = 0; i < n; i++) t.e[l][(size_t)i].luma = r.flag();
        if (s.chroma_format_idc)
            for (int i = 0; i < n; i++) t.e[l][(size_t)i].chroma = r.flag();
        for (int i = 0; i < n; i++) {
            WpEntry& e = t.e[l][(size_t)i];
            if (e.luma) {
                e.delta_weight = r.se();
                e.offset = r.se();
            }
            if (e.chroma)
                for (int j = 0; j < 2; j++) {
                    e.delta_cweight[j] = r.se();
                    e.delta_coffset[j] = r.se();
                }
        }
    }
    return 0;
}

void write_slice_header(BitWriter& w, const SliceHeader& h, const Sps& s, const Pps& p, uint8_t nal_type) {
    w.flag(h.first_slice);
    if (nal_type >= 16 && nal_type <= 23) w.bit(0);  // no_output_of_prior_pics_flag
    w.ue((uint32_t)p.pps_id);
    const int ctb = 1 << s.log2_ctb;
    const int nctb = ((s.width + ctb - 1) / ctb) * ((s.height + ctb - 1) / ctb);
    if (!h.first_slice) {
        if (p.dependent_slices) w.flag(h.dependent);
        w.bits((uint32_t)h.segment_address, ceil_log2(nctb));
    }
    if (!h.dependent) {
        w.bits(0, p.num_extra_slice_header_bits);
        w.ue((uint32_t)h.slice_type);
        if (p.output_flag_present) w.flag(h.pic_output);
        if (nal_type != NAL_IDR_W_RADL && nal_type != NAL_IDR_N_LP) {
            w.bits((uint32_t)h.poc_lsb, s.log2_max_poc_lsb);
            w.flag(h.rps_from_sps);
            if (!h.rps_from_sps) write_st_rps(w, (int)s.st_rps.size(), (int)s.st_rps.size(), s.st_rps, h.rps);
            else if (s.st_rps.size() > 1) w.bits((uint32_t)h.rps_idx, ceil_log2((int)s.st_rps.size()));
            if (s.long_term_refs) w.ue(0);  // num_long_term_pics
            if (s.temporal_mvp) w.flag(h.tmvp);
        }
        if (s.sao) {
            w.flag(h.sao_luma);
            if (s.chroma_format_idc) w.flag(h.sao_chroma);
        }
        if (h.slice_type != 2) {
            const bool b = h.slice_type == 0;
            w.flag(h.num_ref_override);
            if (h.num_ref_override) {
                w.ue((uint32_t)(h.num_ref[0] - 1));
                if (b) w.ue((uint32_t)(h.num_ref[1] - 1));
            }
            const int total = num_pic_total_curr(h);
            if (p.lists_modification && total > 1) {
                for (int l = 0; l < (b ? 2 : 1); l++) {
                    w.flag(h.list_mod[l]);
                    if (h.list_mod[l])
                        for (int i = 0; i < h.num_ref[l]; i++) w.bits(h.list_entry[l][(size_t)i], ceil_log2(total));
                }
            }
            if (b) w.flag(h.mvd_l1_zero);
            if (p.cabac_init_present) w.flag(h.cabac_init);
            if (h.tmvp) {
                if (b) w.flag(h.col_from_l0);
                if (h.num_ref[h.col_from_l0 ? 0 : 1] > 1) w.ue((uint32_t)h.col_ref_idx);
            }
            if (slice_weighted(h, p)) write_pred_weights(w, h, s);
            w.ue((uint32_t)(5 - h.max_merge));
        }
        w.se(h.slice_qp_delta);
        if (p.slice_chroma_qp_offsets_present) {
            w.se(h.cb_qp_offset);
            w.se(h.cr_qp_offset);
        }
        if (p.deblocking_override) w.flag(h.deblocking_override);
        if (h.deblocking_override) {
            w.flag(h.deblocking_disabled);
            if (!h.deblocking_disabled) {
                w.se(h.beta_offset_div2);
                w.se(h.tc_offset_div2);
            }
        }
        const bool disabled = h.deblocking_override ? h.deblocking_disabled : p.deblocking_disabled;
        if (p.lf_across_slices && (h.sao_luma || h.sao_chroma || !disabled)) w.flag(h.lf_across_slices);
    }
    if (p.tiles || p.entropy_sync) {
        w.ue((uint32_t)h.entry_points.size());
        if (!h.entry_points.empty()) {
            uint32_t mx = 1;
            for (uint32_t e : h.entry_points) mx = std::max(mx, e);
            int len = 1;
            while (len < 32 && ((mx - 1) >> len)) len++;
            w.ue((uint32_t)(len - 1));
            for (uint32_t e : h.entry_points) w.bits(e - 1, len);
        }
    }
    if (p.slice_header_extension) w.ue(0);
    w.align_one_zero();
}

// ---------------------------------------------------------------------------
// parsers
// ---------------------------------------------------------------------------
int parse_sps(const std::vector<uint8_t>& rbsp, Sps& s) {
    BitReader r(rbsp.data(), rbsp.size());
    s = Sps();
    r.bits(16);  // NAL header
    r.bits(4);   // sps_video_parameter_set_id
    const int max_sub_layers_minus1 = (int)r.bits(3);
    r.bit();
    skip_ptl(r, max_sub_layers_minus1);
    s.sps_id = (int)r.ue();
    s.chroma_format_idc = (int)r.ue();
    if (s.chroma_format_idc > 3) return set_error(HPVCD_E_UNSUPPORTED_CHROMA, "chroma_format_idc");
    if (s.chroma_format_idc == 3 && r.flag()) return set_error(HPVCD_E_UNSUPPORTED, "separate_colour_plane_flag");
    s.width = (int)r.ue();
    s.height = (int)r.ue();
    if (r.flag()) {
        const int cu_x = (s.chroma_format_idc == 1 || s.chroma_format_idc == 2) ? 2 : 1;
        const int cu_y = s.chroma_format_idc == 1 ? 2 : 1;
        s.conf_win[0] = (int)r.ue() * cu_x;
        s.conf_win[1] = (int)r.ue() * cu_x;
        s.conf_win[2] = (int)r.ue() * cu_y;
        s.conf_win[3] = (int)r.ue() * cu_y;
    }
    s.bit_depth_luma = (int)r.ue() + 8;
    s.bit_depth_chroma = (int)r.ue() + 8;
    if (s.bit_depth_luma > 16 || s.bit_depth_chroma > 16)
        return set_error(HPVCD_E_UNSUPPORTED_BITDEPTH, "bit depth");
    s.log2_max_poc_lsb = (int)r.ue() + 4;
    if (s.log2_max_poc_lsb > 16) return set_error(HPVCD_E_PARAMSET, "log2_max_pic_order_cnt_lsb");
    const bool ordering = r.flag();
    for (int i = ordering ? 0 : max_sub_layers_minus1; i <= max_sub_layers_minus1; i++) {
        r.ue();
        r.ue();
        r.ue();
    }
    s.log2_min_cb = (int)r.ue() + 3;
    s.log2_ctb = s.log2_min_cb + (int)r.ue();
    s.log2_min_tb = (int)r.ue() + 2;
    s.log2_max_tb = s.log2_min_tb + (int)r.ue();
    if (s.log2_min_cb > 6 || s.log2_ctb > 6 || s.log2_max_tb > 5 || s.log2_min_tb > s.log2_max_tb)
        return set_error(HPVCD_E_PARAMSET, "block size limits");
    s.max_th_depth_inter = (int)r.ue();
    s.max_th_depth_intra = (int)r.ue();
    s.scaling_list_enabled = r.flag();
    if (s.scaling_list_enabled) {
        s.scaling_list_data_present = r.flag();
        if (s.scaling_list_data_present) {
            if (int st = parse_scaling_list_data(r, s.scaling)) return st;
        } else {
            s.scaling.set_default();
        }
    }
    s.amp = r.flag();
    s.sao = r.flag();
    s.pcm = r.flag();
    if (s.pcm) {
        s.pcm_bit_depth_luma = (int)r.bits(4) + 1;
        s.pcm_bit_depth_chroma = (int)r.bits(4) + 1;
        s.log2_min_pcm_cb = (int)r.ue() + 3;
        s.log2_max_pcm_cb = s.log2_min_pcm_cb + (int)r.ue();
        s.pcm_loop_filter_disabled = r.flag();
    }
    {
        const uint32_t n = r.ue();
        if (n > 64) return set_error(HPVCD_E_PARAMSET, "num_short_term_ref_pic_sets");
        for (uint32_t i = 0; i < n; i++) {
            StRps rps;
            if (int st = parse_st_rps(r, (int)i, (int)n, s.st_rps, rps)) return st;
            s.st_rps.push_back(rps);
        }
    }
    s.long_term_refs = r.flag();
    if (s.long_term_refs) {
        const uint32_t n = r.ue();
        if (n > 32) return set_error(HPVCD_E_PARAMSET, "num_long_term_ref_pics_sps");
        s.num_lt_sps = (int)n;
        for (uint32_t i = 0; i < n; i++) {
            r.bits(s.log2_max_poc_lsb);
            r.bit();
        }
    }
    s.temporal_mvp = r.flag();
    s.strong_intra_smoothing = r.flag();
    s.vui = r.flag();
    if (s.vui) {
        if (r.flag() && r.bits(8) == 255) r.bits(32);   // aspect ratio, extended SAR
        if (r.flag()) r.bit();                          // overscan
        if (r.flag()) {                                 // video signal type
            r.bits(3);
            s.full_range = r.flag();
            s.colour_description = r.flag();
            if (s.colour_description) {
                s.colour_primaries = (int)r.bits(8);
                s.transfer = (int)r.bits(8);
                s.matrix = (int)r.bits(8);
            }
        }
        if (r.flag()) {  // chroma location
            r.ue();
            r.ue();
        }
        r.bit();
        r.bit();
        r.bit();
        if (r.flag()) for (int i = 0; i < 4; i++) r.ue();   // default display window
        if (r.flag()) {                                    // timing
            r.bits(32);
            r.bits(32);
            if (r.flag()) r.ue();
            if (r.flag()) return set_error(HPVCD_E_UNSUPPORTED, "VUI HRD parameters");
        }
        if (r.flag()) {  // bitstream restriction
            r.bit();
            r.bit();
            r.bit();
            for (int i = 0; i < 5; i++) r.ue();
        }
    }
    if (r.flag()) {  // sps_extension_present_flag
        const bool range_ext = r.flag(), multilayer = r.flag(), ext3d = r.flag(), scc = r.flag();
        r.bits(4);
        if (range_ext) {
            s.ts_rotation = r.flag();
            s.ts_context = r.flag();
            s.implicit_rdpcm = r.flag();
            s.explicit_rdpcm = r.flag();
            s.extended_precision = r.flag();
            s.intra_smoothing_disabled = r.flag();
            s.high_precision_offsets = r.flag();
            s.persistent_rice = r.flag();
            s.bypass_alignment = r.flag();
        }
        if (multilayer) r.bit();
        if (ext3d) return set_error(HPVCD_E_UNSUPPORTED, "SPS 3D extension");
        if (scc) {
            if (r.flag()) return set_error(HPVCD_E_UNSUPPORTED, "current-picture referencing");
            if (r.flag()) return set_error(HPVCD_E_UNSUPPORTED, "palette mode");
            if (r.bits(2)) return set_error(HPVCD_E_UNSUPPORTED, "motion vector resolution control");
            s.intra_boundary_filter_disabled = r.flag();
        }
    }
    if (r.overrun) return set_error(HPVCD_E_BITSTREAM, "SPS truncated");
    if (s.width <= 0 || s.height <= 0) return set_error(HPVCD_E_BAD_DIMENSIONS, "SPS picture size");
    return 0;
}

int parse_pps(const std::vector<uint8_t>& rbsp, const Sps& s, Pps& p) {
    (void)s;
    BitReader r(rbsp.data(), rbsp.size());
    p = Pps();
    r.bits(16);
    p.pps_id = (int)r.ue();
    p.sps_id = (int)r.ue();
    p.dependent_slices = r.flag();
    p.output_flag_present = r.flag();
    p.num_extra_slice_header_bits = (int)r.bits(3);
    p.sign_hiding = r.flag();
    p.cabac_init_present = r.flag();
    p.num_ref_idx_default[0] = (int)r.ue() + 1;
    p.num_ref_idx_default[1] = (int)r.ue() + 1;
    p.init_qp = 26 + r.se();
    p.constrained_intra = r.flag();
    p.transform_skip = r.flag();
    p.cu_qp_delta = r.flag();
    if (p.cu_qp_delta) p.diff_cu_qp_delta_depth = (int)r.ue();
    p.cb_qp_offset = r.se();
    p.cr_qp_offset = r.se();
    p.slice_chroma_qp_offsets_present = r.flag();
    p.weighted_pred = r.flag();
    p.weighted_bipred = r.flag();
    p.transquant_bypass = r.flag();
    p.tiles = r.flag();
    p.entropy_sync = r.flag();
    if (p.tiles) {
        p.tile_cols = (int)r.ue() + 1;
        p.tile_rows = (int)r.ue() + 1;
        if (p.tile_cols > 1024 || p.tile_rows > 1024) return set_error(HPVCD_E_PARAMSET, "tile count");
        p.uniform_spacing = r.flag();
        if (!p.uniform_spacing) {
            for (int i = 0; i < p.tile_cols - 1; i++) p.col_widths.push_back((int)r.ue() + 1);
            for (int j = 0; j < p.tile_rows - 1; j++) p.row_heights.push_back((int)r.ue() + 1);
        }
        p.lf_across_tiles = r.flag();
    }
    p.lf_across_slices = r.flag();
    p.deblocking_control = r.flag();
    if (p.deblocking_control) {
        p.deblocking_override = r.flag();
        p.deblocking_disabled = r.flag();
        if (!p.deblocking_disabled) {
            p.beta_offset_div2 = r.se();
            p.tc_offset_div2 = r.se();
        }
    }
    p.scaling_list_data_present = r.flag();
    if (p.scaling_list_data_present) {
        if (int st = parse_scaling_list_data(r, p.scaling)) return st;
    }
    p.lists_modification = r.flag();
    p.log2_parallel_merge_level = (int)r.ue() + 2;
    p.slice_header_extension = r.flag();
    if (r.flag()) {
        const bool range_ext = r.flag(), multilayer = r.flag(), ext3d = r.flag(), scc = r.flag();
        r.bits(4);
        if (range_ext) {
            if (p.transform_skip) p.log2_max_ts = (int)r.ue() + 2;
            p.cross_component = r.flag();
            p.chroma_qp_offset_list = r.flag();
            if (p.chroma_qp_offset_list) return set_error(HPVCD_E_UNSUPPORTED, "CU chroma QP offset lists");
            p.log2_sao_offset_scale_luma = (int)r.ue();
            p.log2_sao_offset_scale_chroma = (int)r.ue();
        }
        if (multilayer || ext3d) return set_error(HPVCD_E_UNSUPPORTED, "PPS multilayer / 3D extension");
        if (scc) {
            if (r.flag()) return set_error(HPVCD_E_UNSUPPORTED, "current-picture referencing");
            p.act = r.flag();
            if (p.act) {
                p.slice_act_qp_offsets_present = r.flag();
                p.act_y_qp_offset = r.se() - 5;
                p.act_cb_qp_offset = r.se() - 5;
                p.act_cr_qp_offset = r.se() - 3;
            }
            if (r.flag()) return set_error(HPVCD_E_UNSUPPORTED, "palette predictor initializers");
        }
    }
    if (r.overrun) return set_error(HPVCD_E_BITSTREAM, "PPS truncated");
    return 0;
}

int parse_slice_header(const std::vector<uint8_t>& rbsp, const Sps& s, const Pps& p, uint8_t nal_type,
                       SliceHeader& h, bool allow_inter) {
    BitReader r(rbsp.data(), rbsp.size());
    h = SliceHeader();
    r.bits(16);
    h.first_slice = r.flag();
    if (nal_type >= 16 && nal_type <= 23) h.no_output_of_prior_pics = r.flag();
    if ((int)r.ue() != p.pps_id) return set_error(HPVCD_E_PARAMSET, "slice refers to another PPS");
    const int ctb = 1 << s.log2_ctb;
    const int nctb = ((s.width + ctb - 1) / ctb) * ((s.height + ctb - 1) / ctb);
    if (!h.first_slice) {
        if (p.dependent_slices) h.dependent = r.flag();
        h.segment_address = (int)r.bits(ceil_log2(nctb));
        if (h.segment_address >= nctb) return set_error(HPVCD_E_BITSTREAM, "slice_segment_address");
    }
    if (h.dependent) return set_error(HPVCD_E_UNSUPPORTED, "dependent slice segments");
    r.bits(p.num_extra_slice_header_bits);
    h.slice_type = (int)r.ue();
    if (h.slice_type > 2) return set_error(HPVCD_E_BITSTREAM, "slice_type");
    if (h.slice_type != 2 && !allow_inter) return set_error(HPVCD_E_UNSUPPORTED, "P/B slices in a still-image item");
    if (p.output_flag_present) h.pic_output = r.flag();
    if (nal_type != NAL_IDR_W_RADL && nal_type != NAL_IDR_N_LP) {
        h.poc_lsb = (int)r.bits(s.log2_max_poc_lsb);
        const int nsets = (int)s.st_rps.size();
        h.rps_from_sps = r.flag();
        if (!h.rps_from_sps) {
            if (int st = parse_st_rps(r, nsets, nsets, s.st_rps, h.rps)) return st;
        } else {
            if (nsets == 0) return set_error(HPVCD_E_BITSTREAM, "short_term_ref_pic_set_sps_flag without SPS sets");
            h.rps_idx = nsets > 1 ? (int)r.bits(ceil_log2(nsets)) : 0;
            if (h.rps_idx >= nsets) return set_error(HPVCD_E_BITSTREAM, "short_term_ref_pic_set_idx");
            h.rps = s.st_rps[(size_t)h.rps_idx];
        }
        if (s.long_term_refs) {
            const uint32_t n_sps = s.num_lt_sps ? r.ue() : 0;
            if (n_sps || r.ue()) return set_error(HPVCD_E_UNSUPPORTED, "long-term reference pictures");
        }
        if (s.temporal_mvp) h.tmvp = r.flag();
    }
    if (s.sao) {
        h.sao_luma = r.flag();
        if (s.chroma_format_idc) h.sao_chroma = r.flag();
    }
    if (h.slice_type != 2) {
        const bool b = h.slice_type == 0;
        h.num_ref[0] = p.num_ref_idx_default[0];
        h.num_ref[1] = b ? p.num_ref_idx_default[1] : 0;
        h.num_ref_override = r.flag();
        if (h.num_ref_override) {
            h.num_ref[0] = (int)r.ue() + 1;
            if (b) h.num_ref[1] = (int)r.ue() + 1;
        }
        if (h.num_ref[0] > 15 || h.num_ref[1] > 15) return set_error(HPVCD_E_BITSTREAM, "num_ref_idx_active");
        const int total = num_pic_total_curr(h);
        if (p.lists_modification && total > 1) {
            for (int l = 0; l < (b ? 2 : 1); l++) {
                h.list_mod[l] = r.flag();
                if (h.list_mod[l])
                    for (int i = 0; i < h.num_ref[l]; i++) h.list_entry[l].push_back(r.bits(ceil_log2(total)));
            }
        }
        if (b) h.mvd_l1_zero = r.flag();
        if (p.cabac_init_present) h.cabac_init = r.flag();
        if (h.tmvp) {
            if (b) h.col_from_l0 = r.flag();
            if (h.num_ref[h.col_from_l0 ? 0 : 1] > 1) h.col_ref_idx = (int)r.ue();
            if (h.col_ref_idx >= h.num_ref[h.col_from_l0 ? 0 : 1])
                return set_error(HPVCD_E_BITSTREAM, "collocated_ref_idx");
        }
        if (slice_weighted(h, p)) {
            if (int st = parse_pred_weights(r, h, s)) return st;
        }
        const uint32_t five_minus = r.ue();
        if (five_minus > 4) return set_error(HPVCD_E_BITSTREAM, "five_minus_max_num_merge_cand");
        h.max_merge = 5 - (int)five_minus;
    }
    h.slice_qp_delta = r.se();
    if (p.slice_chroma_qp_offsets_present) {
        h.cb_qp_offset = r.se();
        h.cr_qp_offset = r.se();
    }
    if (p.slice_act_qp_offsets_present) return set_error(HPVCD_E_UNSUPPORTED, "slice ACT QP offsets");
    h.deblocking_disabled = p.deblocking_disabled;
    h.beta_offset_div2 = p.beta_offset_div2;
    h.tc_offset_div2 = p.tc_offset_div2;
    if (p.deblocking_override) h.deblocking_override = r.flag();
    if (h.deblocking_override) {
        h.deblocking_disabled = r.flag();
        if (!h.deblocking_disabled) {
            h.beta_offset_div2 = r.se();
            h.tc_offset_div2 = r.se();
        }
    }
    h.lf_across_slices = p.lf_across_slices;
    if (p.lf_across_slices && (h.sao_luma || h.sao_chroma || !h.deblocking_disabled)) h.lf_across_slices = r.flag();
    if (p.tiles || p.entropy_sync) {
        const uint32_t n = r.ue();
        if (n > (uint32_t)nctb) return set_error(HPVCD_E_BITSTREAM, "num_entry_point_offsets");
        if (n) {
            const int len = (int)r.ue() + 1;
            if (len > 32) return set_error(HPVCD_E_BITSTREAM, "offset_len_minus1");
            for (uint32_t i = 0; i < n; i++) h.entry_points.push_back(r.bits(len) + 1);
        }
    }
    if (p.slice_header_extension) {
        const uint32_t n = r.ue();
        for (uint32_t i = 0; i < n; i++) r.bits(8);
    }
    if (!r.bit()) return set_error(HPVCD_E_BITSTREAM, "slice header alignment bit");
    r.skip_to_byte();
    if (r.overrun) return set_error(HPVCD_E_BITSTREAM, "slice header truncated");
    h.data_offset = r.pos / 8;
    return 0;
}

}  // namespace hpvcd
