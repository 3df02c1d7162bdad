// device.h — device-side descriptors shared by the HIP kernels and the host
// runtime that prepares them.  Plain structs, no device types.
#pragma once

#include <cstdint>

#include "../../include/hpvcd_records.h"

namespace hpvcd {

// Per-picture descriptor as the kernels see it (host-derived geometry plus
// offsets into the job's device pools).
struct DevPic {
    int32_t w, h, cw, ch;           // luma / chroma plane sizes
    int32_t gw, gh;                 // 4x4 luma cell grid
    int32_t sub_w_log2, sub_h_log2;
    int32_t ctb_log2, ctb_cols, ctb_rows;
    int32_t bd, bd_c;
    uint32_t flags;                 // HPVCD_PF_*
    int32_t num_slices;
    hpvcd_slice_params slices[HPVCD_MAX_SLICES];
    int32_t pps_cb_qp_offset, pps_cr_qp_offset;
    int32_t chroma_idc, matrix;
    int32_t crop_left, crop_top, win_w, win_h, dst_x, dst_y, copy_w, copy_h;
    int32_t out_kind, out_width;    // output image kind / row length in pixels
    int32_t out_height, orient;     // output canvas rows; HPVCD_ORIENT_* of the output image
    uint64_t out_off;               // byte offset of the output image in the output pool
    uint64_t plane_off[3];          // element offsets of Y/Cb/Cr in the sample pools
    uint64_t res_off[3];            // element offsets of Y/Cb/Cr in the residual pool
    uint32_t tb_begin;              // first TB of the picture in the job's TB array
    uint32_t coeff_begin;           // first coefficient
    uint32_t ctb_begin;             // first SAO record
    uint32_t cell_begin;            // first cell (info/order)
    int32_t scaling_index;          // -1 = flat
    uint32_t suppress_any;          // some cell is TQB or PCM-with-loop-filter-disabled
    uint32_t sao_across_slices;     // pps flag && every slice's flag (SAO neighbour gate)
    uint32_t deblock_on;            // some slice of the picture enables deblocking
    uint32_t inter;                 // the picture has prediction units (MC output in the plane)
};

// One prediction unit as the MC kernel sees it: references resolved to
// class-local pictures (lists whose reference is missing are dropped from
// flags) and the weight-table row of each list.
struct DevPu {
    int32_t x, y, w, h;             // luma geometry
    uint32_t flags;                 // usable lists: bit 0 list 0, bit 1 list 1
    uint32_t weighted;              // explicit weighted prediction
    int32_t log2_denom[2];          // luma, chroma
    int32_t mv[2][2];               // [list][x, y], quarter luma samples
    uint32_t ref[2];                // class-local reference picture per list
    int32_t weight[2][3], offset[2][3];
};

// MC work item: a block of at most 16 x 16 luma samples of one PU (and the
// co-located chroma), offset (ox, oy) from the PU origin.
struct DevMcTask {
    uint32_t pu;
    uint16_t pic;
    uint8_t ox, oy;
};

// One residual TB as K1 sees it, packed by the host in launch order: the
// kernel's only dependent load before the coefficient words.  Everything the
// TB's flags, its picture and the scaling lists decide is folded in here.
struct DevResJob {
    uint64_t out;            // residual-pool element of the TB's sample (x, y)
    uint32_t luma_delta;     // out minus the co-located luma residual element (CCP)
    uint32_t coeff;          // first coefficient word in the class's coefficient pool
    int32_t scaling;         // byte offset of the TB's scaling entry, -1 = flat
    uint16_t coeff_cnt;
    uint16_t stride;         // residual plane row length
    uint16_t flags;          // RJ_*
    uint8_t qp;              // Qp' clamped to the bit depth's range; PCM: the sample shift
    uint8_t nx;              // columns that hold nonzero levels
    int8_t res_scale;        // ResScaleVal (CCP)
    uint8_t bds;             // luma bit depth - 8 | (chroma bit depth - 8) << 4
    uint8_t vw, vh;          // TB samples inside the plane
};
static_assert(sizeof(DevResJob) == 32, "DevResJob is 32 bytes");

// DevResJob.flags
constexpr uint16_t RJ_CBF = 1u << 0;
constexpr uint16_t RJ_TS = 1u << 1;
constexpr uint16_t RJ_TQB = 1u << 2;
constexpr uint16_t RJ_PCM = 1u << 3;
constexpr uint16_t RJ_TRANSFORM = 1u << 4;  // CBF without TS, TQB or PCM
constexpr uint16_t RJ_DST = 1u << 5;        // 4x4 intra luma DST-VII
constexpr uint16_t RJ_I16 = 1u << 6;        // residual arithmetic wraps to 16 bits
constexpr uint16_t RJ_ROTATE = 1u << 7;     // TS rotation of a 4x4 intra TB
constexpr uint16_t RJ_RDPCM = 1u << 8;      // residual DPCM on ...
constexpr uint16_t RJ_RDPCM_V = 1u << 9;    // ... vertical (else horizontal)
constexpr uint16_t RJ_EP = 1u << 10;        // extended precision processing
constexpr uint16_t RJ_CHROMA = 1u << 11;    // component 1 or 2
constexpr uint16_t RJ_CCP = 1u << 12;       // cross-component prediction from luma (res_scale != 0)

// One (picture, plane, CTB) work unit of the intra reconstruction wavefront
// (its intra and I_PCM TBs; inter TBs arrive reconstructed).
struct DevUnit {
    uint32_t pic;
    uint16_t cx, cy;        // CTB coordinates
    uint8_t plane;          // 0 Y, 1 Cb, 2 Cr
    uint8_t nlevels;
    uint16_t ntb;
    uint32_t tb_list;       // offset into unit_tbs
    uint32_t desc_base;     // first descriptor (tb_list, or the Cb unit's when Cr shares its preparation)
    uint32_t level_list;    // offset into level_off (nlevels + 1 entries)
    uint32_t sbase;         // first per-sample reference word of the unit
    uint32_t nsamp;         // reference words of the unit (sum of its TB areas)
    uint32_t rbase;         // first reference map entry of the unit
    uint32_t nref;          // reference map entries (sum of 4n + 1 over its TBs)
    uint64_t inter_bottom;  // bottom-row samples of inter TBs (bit = column)
    uint64_t inter_right;   // right-column samples of inter TBs (bit = row)
};

// The fields of a TB record K2a reads, copied by the host in each unit's TB
// order (one coalesced load per TB).
struct PrepTb {
    uint16_t x, y;
    uint8_t log2n, mode;
    uint16_t flags;
    uint32_t order;
};

// TB flag bits added by the runtime (upper byte of hpvcd_tb.flags).
constexpr uint16_t TBX_HAS_RES = 1u << 8;  // residual samples exist in the residual pool
// bits 9-11: in the wavefront's TB order of its unit, the TB leads a group of
// this many consecutive 4 x 4 fast-path TBs of one dependency level (they do
// not read each other's samples and run in one step); 0 inside a group
constexpr int TBX_GROUP_SHIFT = 9;
// (step leaders only) some TB of the step writes the unit's bottom row or
// right column; some TB of the step is planar or horizontal / vertical; bits
// 14-15: size of the next step minus one
constexpr uint16_t TBX_STEP_PUB = 1u << 12;
constexpr uint16_t TBX_STEP_PLAIN = 1u << 13;
constexpr int TBX_NEXT_SHIFT = 14;

}  // namespace hpvcd
