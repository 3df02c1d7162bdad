// hevc_ps.h — HEVC parameter sets and slice segment headers: the fields the
// record producer needs, their writers (the generator's bitstream) and their
// parsers (H.265 7.3.2-7.3.6; the reference's src/config.rs:566 parse_sps,
// :990 parse_pps, :504 scaling_list_data, src/decode.rs:7713
// parse_slice_header_full).
#pragma once

#include <cstdint>
#include <vector>

#include "bitio.h"

namespace hpvcd {

enum NalType : uint8_t {
    NAL_TRAIL_N = 0,
    NAL_TRAIL_R = 1,
    NAL_IDR_W_RADL = 19,
    NAL_IDR_N_LP = 20,
    NAL_CRA = 21,
    NAL_VPS = 32,
    NAL_SPS = 33,
    NAL_PPS = 34,
};

// Tables 7-5 / 7-6 default 8x8 lists in raster order
extern const uint8_t kDefaultIntra8x8[64];
extern const uint8_t kDefaultInter8x8[64];

// Scaling lists in the record layout: [sizeId][matrixId] -> 64 raster
// coefficients (16 for 4x4), DC, flat flag (hpvcd_records.h).
struct ScalingLists {
    uint8_t e[4][6][66];
    void set_default();   // Tables 7-5 / 7-6 (src/config.rs:381 ScalingList::default)
    void refresh_flat(int size_id, int m);
};

// A short-term reference picture set (7.3.7, 7.4.8): delta POCs of the
// pictures before the current one (S0, nearest first) and after it (S1,
// nearest first), each with used_by_curr_pic_flag.
struct StRps {
    std::vector<int32_t> d0, d1;
    std::vector<uint8_t> u0, u1;
    int num() const { return (int)(d0.size() + d1.size()); }
    bool operator==(const StRps& o) const { return d0 == o.d0 && d1 == o.d1 && u0 == o.u0 && u1 == o.u1; }
};

// st_ref_pic_set(idx) with num_sets sets in the SPS (idx == num_sets: the
// slice header's own set); predicted sets derive from sets[idx - delta_idx]
// (src/rps.rs:74 parse_short_term_rps, :148 derive_inter_rps)
int parse_st_rps(BitReader& r, int idx, int num_sets, const std::vector<StRps>& sets, StRps& out);
// writes `rps` predicted from the previous set when a delta POC makes it
// derivable (inter_ref_pic_set_prediction_flag), else explicitly
void write_st_rps(BitWriter& w, int idx, int num_sets, const std::vector<StRps>& sets, const StRps& rps);
// 7.4.8 derivation of a predicted set
StRps derive_predicted_rps(const StRps& ref, int delta_rps, const std::vector<uint8_t>& used,
                           const std::vector<uint8_t>& use_delta);

// pred_weight_table (7.3.6.3): the syntax values of each list entry
struct WpEntry {
    bool luma = false, chroma = false;
    int delta_weight = 0, offset = 0;            // delta_luma_weight_lX, luma_offset_lX
    int delta_cweight[2] = {0, 0}, delta_coffset[2] = {0, 0};
};
struct PredWeights {
    int luma_log2_denom = 0, chroma_log2_denom = 0;
    std::vector<WpEntry> e[2];
};

struct Sps {
    int sps_id = 0;
    int chroma_format_idc = 1;
    int width = 0, height = 0;
    int conf_win[4] = {0, 0, 0, 0};   // left, right, top, bottom in luma samples
    int bit_depth_luma = 8, bit_depth_chroma = 8;
    int log2_max_poc_lsb = 8;
    int log2_min_cb = 3, log2_ctb = 6;
    int log2_min_tb = 2, log2_max_tb = 5;
    int max_th_depth_inter = 1, max_th_depth_intra = 1;
    bool scaling_list_enabled = false, scaling_list_data_present = false;
    ScalingLists scaling;
    bool amp = false, sao = false;
    bool pcm = false;
    int pcm_bit_depth_luma = 8, pcm_bit_depth_chroma = 8;
    int log2_min_pcm_cb = 3, log2_max_pcm_cb = 5;
    bool pcm_loop_filter_disabled = false;
    std::vector<StRps> st_rps;      // num_short_term_ref_pic_sets sets
    int max_dec_pic_buffering = 5, max_num_reorder = 0;
    int num_lt_sps = 0;             // num_long_term_ref_pics_sps
    bool long_term_refs = false, temporal_mvp = false;
    bool strong_intra_smoothing = false;
    // VUI colour description (E.2.1)
    bool vui = false, colour_description = false, full_range = false;
    int colour_primaries = 2, transfer = 2, matrix = 2;
    // range extension (7.3.2.2.2)
    bool ts_rotation = false, ts_context = false, implicit_rdpcm = false, explicit_rdpcm = false;
    bool extended_precision = false, intra_smoothing_disabled = false, high_precision_offsets = false;
    bool persistent_rice = false, bypass_alignment = false;
    // screen content extension (7.3.2.2.3)
    bool intra_boundary_filter_disabled = false;
};

struct Pps {
    int pps_id = 0, sps_id = 0;
    bool dependent_slices = false, output_flag_present = false;
    int num_extra_slice_header_bits = 0;
    bool sign_hiding = false, cabac_init_present = false;
    int num_ref_idx_default[2] = {1, 1};
    int init_qp = 26;
    bool constrained_intra = false, transform_skip = false;
    bool cu_qp_delta = false;
    int diff_cu_qp_delta_depth = 0;
    int cb_qp_offset = 0, cr_qp_offset = 0;
    bool slice_chroma_qp_offsets_present = false;
    bool weighted_pred = false, weighted_bipred = false;
    bool transquant_bypass = false;
    bool tiles = false, entropy_sync = false;
    int tile_cols = 1, tile_rows = 1;
    bool uniform_spacing = true;
    std::vector<int> col_widths, row_heights;   // in CTBs, when not uniform
    bool lf_across_tiles = true;
    bool lf_across_slices = false;
    bool deblocking_control = false, deblocking_override = false, deblocking_disabled = false;
    int beta_offset_div2 = 0, tc_offset_div2 = 0;
    bool scaling_list_data_present = false;
    ScalingLists scaling;
    bool lists_modification = false;
    int log2_parallel_merge_level = 2;
    bool slice_header_extension = false;
    // range extension (7.3.2.3.2)
    int log2_max_ts = 2;
    bool cross_component = false, chroma_qp_offset_list = false;
    int log2_sao_offset_scale_luma = 0, log2_sao_offset_scale_chroma = 0;
    // screen content extension: the adaptive colour transform (7.3.2.3.3)
    bool act = false, slice_act_qp_offsets_present = false;
    int act_y_qp_offset = -5, act_cb_qp_offset = -5, act_cr_qp_offset = -3;
};

struct SliceHeader {
    bool first_slice = true;
    bool dependent = false;
    int segment_address = 0;   // CTB address in raster scan of the first CTB
    int slice_type = 2;        // 0 B, 1 P, 2 I
    int poc_lsb = 0;
    bool sao_luma = false, sao_chroma = false;
    int slice_qp_delta = 0;
    int cb_qp_offset = 0, cr_qp_offset = 0;
    bool deblocking_override = false, deblocking_disabled = false;
    int beta_offset_div2 = 0, tc_offset_div2 = 0;
    bool lf_across_slices = false;
    std::vector<uint32_t> entry_points;   // sub-stream byte lengths (all but the last)
    size_t data_offset = 0;               // byte offset of the slice data in the RBSP (NAL header included)
    bool no_output_of_prior_pics = false, pic_output = true;
    // reference pictures (non-IDR) and P / B fields
    bool rps_from_sps = false;            // short_term_ref_pic_set_sps_flag
    int rps_idx = 0;                      // short_term_ref_pic_set_idx
    StRps rps;                            // the picture's short-term set (inline or the SPS's)
    bool tmvp = false;                    // slice_temporal_mvp_enabled_flag
    bool num_ref_override = false;
    int num_ref[2] = {0, 0};              // active entries (list 1: B slices)
    bool list_mod[2] = {false, false};    // ref_pic_list_modification_flag_lX
    std::vector<uint32_t> list_entry[2];
    bool mvd_l1_zero = false, cabac_init = false;
    bool col_from_l0 = true;
    int col_ref_idx = 0;
    PredWeights wp;                       // when the PPS enables weighting for the slice type
    int max_merge = 5;                    // MaxNumMergeCand
};

// NumPicTotalCurr of a slice (7-55): used entries of its short-term set
int num_pic_total_curr(const SliceHeader& h);
// whether the slice carries a pred_weight_table
bool slice_weighted(const SliceHeader& h, const Pps& pps);

// uniform tile spacing (6-3 / 6-4) or explicit widths; column / row CTB boundaries
void tile_boundaries(const Pps& pps, int ctb_cols, int ctb_rows, std::vector<int>& col_bd, std::vector<int>& row_bd);

std::vector<uint8_t> write_vps(const Sps& sps);
std::vector<uint8_t> write_sps(const Sps& sps);
std::vector<uint8_t> write_pps(const Pps& pps, const Sps& sps);
// slice header through byte_alignment(); the slice data follows in the same writer
void write_slice_header(BitWriter& w, const SliceHeader& h, const Sps& sps, const Pps& pps, uint8_t nal_type);

// parsers over an RBSP with the two NAL header bytes at the front; return
// an hpvcd_status (0 = OK) and set the error message
int parse_sps(const std::vector<uint8_t>& rbsp, Sps& sps);
int parse_pps(const std::vector<uint8_t>& rbsp, const Sps& sps, Pps& pps);
// allow_inter: P and B slices are parsed (a still-image item refuses them)
int parse_slice_header(const std::vector<uint8_t>& rbsp, const Sps& sps, const Pps& pps, uint8_t nal_type,
                       SliceHeader& h, bool allow_inter = false);

}  // namespace hpvcd
