// gen.h — the seeded record generator's picture structure, shared by the
// record generator and the HEVC / HEIC writer.
#pragma once

#include <functional>
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

// Display-order number (POC) of the picture at decode index `index`.
int32_t gen_poc(const hpvcd_gen_config& c, uint32_t index);
// Reference lists of an inter picture as batch indices (list 1 empty for P
// slices): the GOP's, or with gop <= 1 the pictures just before it (list 1
// the same pictures, farthest first).
void gen_lists(const hpvcd_gen_config& c, uint32_t index, std::vector<uint32_t>& l0, std::vector<uint32_t>& l1);
// The inter slice parameters of picture `index` (> 0) of an inter
// sequence: lists, weights, merge / temporal parameters; `col` is left null.
void gen_inter_slice(const hpvcd_gen_config& c, uint32_t index, InterSlice& s);
// The picture whose motion is the collocated motion of picture `index`, or -1.
int gen_col_picture(const hpvcd_gen_config& c, uint32_t index);
// Runs fn(i) for every picture on host threads, each after its collocated picture.
void gen_for_pictures(const hpvcd_gen_config& c, const std::function<void(uint32_t)>& fn);

// Picture `index` of the configuration: its records, coded through `enc`
// when given; `col` is its collocated picture's motion (temporal MVP).
void gen_picture(const hpvcd_gen_config& cfg, uint32_t index, PicRecords& out, PicCoder* enc,
                 const ColField* col = nullptr);

// the configuration's checks (also run by the writer)
hpvcd_status gen_validate(const hpvcd_gen_config& c);
// one record scaling set (HPVCD_SCALING_SET_BYTES)
void make_scaling_set(const hpvcd_gen_config& c, uint8_t* set);
// the record set of already generated pictures
hpvcd_status gen_assemble(const hpvcd_gen_config& c, std::vector<PicRecords>& pics, hpvcd_records** out);

}  // namespace hpvcd
