// motion.h — motion data of inter pictures and the derivations the parse
// runs for every prediction unit: merge candidates (src/motion.rs:155
// derive_merge, spatial :238, temporal :282, combined bi-predictive :333,
// zero :375), AMVP predictors (:417 derive_amvp, :515 amvp_spatial), the
// POC-distance scaling (:41 scale_mv), and the collocated (temporal) vector
// of a picture's stored motion (src/decode.rs:6180 DecoderNeighbors::temporal).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace hpvcd {

// Motion of one 4x4 luma cell.
struct MotionCell {
    uint8_t intra = 1;          // intra (or never coded): no motion
    uint8_t pred = 0;           // bit 0: list 0 used, bit 1: list 1 used
    uint8_t lt = 0;             // bit l: the list-l reference is a long-term picture
    uint8_t skip = 0;           // cu_skip_flag of the CU (context of the next CUs' flags)
    int8_t ref_idx[2] = {-1, -1};
    int16_t mv[2][2] = {{0, 0}, {0, 0}};  // quarter-sample (x, y) per list
    int32_t poc[2] = {0, 0};    // POC of the reference picture per list
    int32_t pic[2] = {-1, -1};  // the reference's batch picture index (deblocking compares pictures)
};

// A reference list entry.
struct RefPic {
    int32_t poc = 0;
    int32_t pic = -1;  // batch picture index
    bool lt = false;
};

// Stored motion of a decoded picture as collocated motion: one cell per
// 16x16 luma block, the cell at the block's top-left corner (8.5.3.2.8,
// xColPb = (x >> 4) << 4).
struct ColField {
    int cw = 0, ch = 0;     // blocks per row / column
    int32_t poc = 0;        // the picture's POC
    std::vector<MotionCell> cells;
    const MotionCell* at(int x, int y) const {
        const int bx = x >> 4, by = y >> 4;
        if (bx >= cw || by >= ch) return nullptr;
        return &cells[(size_t)by * cw + bx];
    }
};

// The current picture's decoded motion as the derivations see it.
class Neighbours {
public:
    virtual ~Neighbours() = default;
    // z-scan / slice / tile availability of a position outside the current
    // coding block (6.4.1): decoded, in the same slice and tile
    virtual bool available(int x, int y) const = 0;
    // the cell at an in-picture position (any state)
    virtual const MotionCell& motion_at(int x, int y) const = 0;
};

// Slice-level motion parameters.
struct MotionParams {
    int width = 0, height = 0;
    bool is_b = false;
    int32_t cur_poc = 0;
    int log2_ctb = 6;
    int par_mrg_level = 2;       // Log2ParMrgLevel
    int max_merge = 5;           // MaxNumMergeCand
    bool tmvp = false;           // slice_temporal_mvp_enabled_flag
    const RefPic* list[2] = {nullptr, nullptr};
    int num[2] = {0, 0};
    bool col_from_l0 = true;     // collocated_from_l0_flag
    const ColField* col = nullptr;  // the collocated picture's motion (null: none)
};

struct PuGeom {
    int x = 0, y = 0, w = 0, h = 0;
    int cu_x = 0, cu_y = 0, cu_w = 0, cu_h = 0;
    int part_idx = 0;
};

struct MergeCand {
    uint8_t pred = 0;
    int8_t ref_idx[2] = {-1, -1};
    int16_t mv[2][2] = {{0, 0}, {0, 0}};
};

// 8.5.3.2.9 POC-distance scaling (src/motion.rs:41)
void scale_mv(const int16_t in[2], int col_dist, int cur_dist, int16_t out[2]);

// The merge candidate list, up to MaxNumMergeCand entries (src/motion.rs:155
// before the merge_idx pick); returns its length.
int derive_merge_list(const Neighbours& nb, const MotionParams& mp, const PuGeom& pu, MergeCand* out);

// The merge motion of a PU for merge_idx (src/decode.rs:6626
// derive_pu_merge): an 8x8 CU with Log2ParMrgLevel > 2 derives one list for
// the whole CU; an index past the list gives the zero candidate; an 8x4 /
// 4x8 PU keeps list 0 of a bi-predictive candidate.
MergeCand derive_merge(const Neighbours& nb, const MotionParams& mp, const PuGeom& pu, int merge_idx);

// The two AMVP predictors of list `list` for the target reference POC
// (src/motion.rs:417).
void derive_amvp(const Neighbours& nb, const MotionParams& mp, const PuGeom& pu, int list, int32_t ref_poc,
                 int16_t out[2][2]);

// The collocated vector at (x, y) for the target list and reference POC,
// scaled (src/decode.rs:6180); false when unavailable.
bool temporal_mv(const MotionParams& mp, int x, int y, int list, int32_t ref_poc, int16_t out[2]);

// inter_motion_differs (src/decode.rs:620): deblocking bS 1 between two
// inter cells
bool motion_differs(const MotionCell& p, const MotionCell& q);

}  // namespace hpvcd
