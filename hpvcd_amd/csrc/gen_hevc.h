// gen_hevc.h — coded HEVC of the generator's pictures (gen_hevc.cpp).
#pragma once

#include <string>
#include <vector>

#include "../../include/hpvcd.h"
#include "gen.h"
#include "hevc_ps.h"

namespace hpvcd {

// The coded slice data of one picture: per slice its substreams (one per
// tile the slice touches), each byte-aligned RBSP.
struct PicStream {
    struct Slice {
        std::vector<std::vector<uint8_t>> substreams;
    };
    std::vector<Slice> slices;
    bool failed = false;
    std::string why;
};

// The reference structure of an inter sequence as coded: each picture's
// POC and short-term RPS (every earlier picture that it or a later picture
// refers to, used_by_curr for its own references), the SPS's sets when the
// slices name them, the POC LSB length and the DPB sizes.
struct SeqPlan {
    std::vector<int32_t> poc;
    std::vector<StRps> rps;
    std::vector<StRps> sps_sets;
    std::vector<int> rps_idx;   // per picture, into sps_sets (rps_in_sps)
    int log2_max_poc_lsb = 4;
    int max_dec = 1, max_reorder = 0;
};
SeqPlan gen_plan(const hpvcd_gen_config& c);

void gen_parameter_sets(const hpvcd_gen_config& c, Sps& sps, Pps& pps, const SeqPlan* plan = nullptr);
// slice segment NAL units (two header bytes, escaped payload) of picture
// `index`; `plan` is required for inter sequences
std::vector<std::vector<uint8_t>> gen_slice_nals(const hpvcd_gen_config& c, const Sps& sps, const Pps& pps,
                                                 uint32_t index, const PicStream& st, const SeqPlan* plan = nullptr);
// records and coded slices of every picture of the configuration
hpvcd_status gen_coded(const hpvcd_gen_config& c, std::vector<PicRecords>& recs, std::vector<PicStream>& streams);
// malloc'd copy for the C API (freed with hpvcd_bytes_free)
uint8_t* copy_out(const std::vector<uint8_t>& v, size_t* size);

}  // namespace hpvcd
