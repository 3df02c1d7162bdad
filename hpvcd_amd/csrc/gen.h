// gen.h — the seeded record generator's picture structure, shared by the
// record generator and the HEVC / HEIC writer.
#pragma once

#include <memory>
#include <vector>

#include "../../include/hpvcd.h"
#include "walker.h"

namespace hpvcd {

// Uniformly spaced tile boundaries in CTBs and the tile scan (CtbAddrTsToRs).
struct TileLayout {
    std::vector<int> col_bd, row_bd;
    std::vector<uint32_t> ts_to_rs;
    std::vector<uint8_t> tile_of;  // per raster CTB
    TileLayout(int cols, int rows, int ntc, int ntr);
    int num_tiles() const { return (int)(col_bd.size() - 1) * (int)(row_bd.size() - 1); }
};

PicParams gen_params(const hpvcd_gen_config& c);
// slice index per raster CTB: equal runs of the tile scan
std::vector<uint8_t> gen_slices(const hpvcd_gen_config& c, const TileLayout& T);

// deblocking parameters of the slices of one picture
void gen_slice_params(const hpvcd_gen_config& c, uint32_t index, int nslices, hpvcd_slice_params* out);

// Turns the generator's decisions into coded syntax: wrap() returns the
// Source the walker reads, which forwards every decision to `inner` and
// codes it.
struct PicCoder {
    virtual ~PicCoder() = default;
    virtual std::unique_ptr<Source> wrap(Source& inner, const PicParams& P) = 0;
};

// Picture `index` of the configuration: its records, coded through `enc`
// when given.
void gen_picture(const hpvcd_gen_config& cfg, uint32_t index, PicRecords& out, PicCoder* enc);

// the configuration's checks (also run by the writer)
hpvcd_status gen_validate(const hpvcd_gen_config& c);
// one record scaling set (HPVCD_SCALING_SET_BYTES)
void make_scaling_set(const hpvcd_gen_config& c, uint8_t* set);
// the record set of already generated pictures
hpvcd_status gen_assemble(const hpvcd_gen_config& c, std::vector<PicRecords>& pics, hpvcd_records** out);

}  // namespace hpvcd
