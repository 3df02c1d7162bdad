// heif.h — HEIF container parse and write (heif.cpp).
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/hpvcd.h"
#include "hevc_ps.h"
#include "parse.h"

namespace hpvcd {

// ParseLimits defaults (src/limits.rs:85-102)
struct HeifLimits {
    uint32_t max_dimension = 16384;
    uint64_t max_pixels = 64ull << 20;
    uint64_t max_box_size = 256ull << 20;
    uint64_t max_item_size = 256ull << 20;
    size_t max_hvcc_size = 64 << 10;
    uint32_t max_items = 4096;
    uint32_t max_extents_per_item = 64;
    uint32_t max_tiles = 16384;
};

struct HeifItem {
    uint32_t id = 0;
    uint64_t offset = 0, length = 0;   // absolute file range of the sample
    std::vector<uint8_t> hvcc;         // hvcC payload
    uint32_t display_w = 0, display_h = 0;   // ispe
    bool has_cicp = false;             // colr nclx
    uint8_t cicp_matrix = 2;
    bool cicp_full_range = false;
    uint8_t orientation = HPVCD_ORIENT_NORMAL;   // irot / imir
};

struct HeifFile {
    HeifItem primary;                  // the first tile of a grid
    bool is_grid = false;
    uint32_t rows = 1, cols = 1, out_w = 0, out_h = 0;
    std::vector<HeifItem> tiles;       // dimg order, failed items dropped
    uint8_t grid_orientation = HPVCD_ORIENT_NORMAL;
};

struct HeicInfo {
    uint8_t orientation = HPVCD_ORIENT_NORMAL;
    int tiles = 0, rows = 1, cols = 1, failed_tiles = 0;
};

// want_kind of heic_records: the kind decode_heic returns, the RGB8 kind of
// decode_heic_rgb8, or an explicit HPVCD_OUT_* kind
constexpr int HEIC_KIND_DEFAULT = -1;
constexpr int HEIC_KIND_RGB8 = -2;
constexpr int HEIC_KIND_YUV = -3;   // raw planes (decode_yuv)

uint8_t canonical_matrix(int m);
hpvcd_status check_image(const HeifLimits& L, uint64_t w, uint64_t h);
hpvcd_status heif_parse(const uint8_t* file, size_t n, const HeifLimits& L, HeifFile& out);
hpvcd_status parse_hvcc(const std::vector<uint8_t>& hvcc, Sps& sps, Pps& pps);
hpvcd_status item_picture(const uint8_t* file, size_t n, const HeifItem& it, const std::vector<uint8_t>& hvcc,
                          ParsedPicture& out, Sps& sps);
// the records of a HEIC file's primary image (one output image)
hpvcd_status heic_records(const uint8_t* file, size_t n, const HeifLimits& L, int threads, int want_kind,
                          hpvcd_records** out, HeicInfo* info);
hpvcd_status gen_heic(const hpvcd_gen_config& c, std::vector<uint8_t>& file);
HeifLimits limits_of(const hpvcd_settings* s);

}  // namespace hpvcd
