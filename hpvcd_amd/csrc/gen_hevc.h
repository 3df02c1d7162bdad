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

void gen_parameter_sets(const hpvcd_gen_config& c, Sps& sps, Pps& pps);
// slice segment NAL units (two header bytes, escaped payload) of picture `index`
std::vector<std::vector<uint8_t>> gen_slice_nals(const hpvcd_gen_config& c, const Sps& sps, const Pps& pps,
                                                 uint32_t index, const PicStream& st);
// records and coded slices of every picture of the configuration
hpvcd_status gen_coded(const hpvcd_gen_config& c, std::vector<PicRecords>& recs, std::vector<PicStream>& streams);
// malloc'd copy for the C API (freed with hpvcd_bytes_free)
uint8_t* copy_out(const std::vector<uint8_t>& v, size_t* size);

}  // namespace hpvcd
