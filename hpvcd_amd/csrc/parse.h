// parse.h — HEVC syntax parse into reconstruction records (parse.cpp).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "../../include/hpvcd.h"
#include "hevc_ps.h"
#include "walker.h"

namespace hpvcd {

// One coded picture's slice segment NAL units, each as RBSP with its two
// NAL header bytes at the front, and the NAL positions of the emulation
// prevention bytes removed from each (entry points count them).
struct CodedPicture {
    std::vector<std::vector<uint8_t>> slices;
    std::vector<std::vector<uint32_t>> removed;
};

// A parsed picture: its records and the picture record's coding fields
// (placement, output and array ranges are the caller's).
struct ParsedPicture {
    PicRecords rec;
    hpvcd_picture pic;
    std::vector<uint8_t> scaling;  // HPVCD_SCALING_SET_BYTES when scaling lists are on
};

// slice segment headers of a coded picture (allow_inter: P and B slices)
hpvcd_status parse_headers(const Sps& sps, const Pps& pps, const CodedPicture& cp, std::vector<SliceHeader>& hdr,
                           bool allow_inter);
// The picture's slice data into records.  Without `headers` the headers are
// parsed here and must be intra; `inter` gives each slice's inter parameters
// (null entries: I slices) and `poc` the picture's POC.
hpvcd_status parse_picture(const Sps& sps, const Pps& pps, const CodedPicture& cp, ParsedPicture& out,
                           const std::vector<SliceHeader>* headers = nullptr,
                           const std::vector<const InterSlice*>* inter = nullptr, int32_t poc = 0);

// NAL units of a length-prefixed sample (size_len 1, 2 or 4 bytes) or of an
// Annex B stream (size_len 0); each escaped, header included
hpvcd_status split_nals(const uint8_t* data, size_t size, int size_len, std::vector<std::pair<const uint8_t*, size_t>>& out);

// a record set from parsed pictures: pictures are placed by their own
// records; identical scaling sets are shared
hpvcd_status assemble_records(std::vector<ParsedPicture>& pics, const std::vector<hpvcd_output>& outputs,
                              hpvcd_records** out);

}  // namespace hpvcd
