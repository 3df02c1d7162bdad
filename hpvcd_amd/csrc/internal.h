// internal.h — host-side internals shared by the generator and the runtime.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/hpvcd.h"

struct hpvcd_records {
    std::vector<hpvcd_picture> pictures;
    std::vector<hpvcd_tb> tbs;
    std::vector<uint32_t> coeffs;
    std::vector<hpvcd_sao> sao;
    std::vector<uint32_t> cells;
    std::vector<uint8_t> scaling;
    std::vector<hpvcd_output> outputs;
    std::vector<hpvcd_pu> pus;
    std::vector<hpvcd_ref> refs;
};

namespace hpvcd {

// Records the message for hpvcd_last_error() on this thread; returns st.
hpvcd_status set_error(hpvcd_status st, const char* msg);
uint64_t tile_seed(uint64_t base, uint32_t config_id, uint32_t tile);

inline int sub_w_of(int idc) { return (idc == 1 || idc == 2) ? 2 : 1; }
inline int sub_h_of(int idc) { return idc == 1 ? 2 : 1; }
inline int chroma_w(const hpvcd_picture& p) { return p.chroma_idc ? (p.width + sub_w_of(p.chroma_idc) - 1) / sub_w_of(p.chroma_idc) : 0; }
inline int chroma_h(const hpvcd_picture& p) { return p.chroma_idc ? (p.height + sub_h_of(p.chroma_idc) - 1) / sub_h_of(p.chroma_idc) : 0; }

// Validates ranges, sizes and cross references of a batch (host side).
hpvcd_status validate_batch(const hpvcd_batch& b);
// hpvcd_job_create for a batch the caller has validated already
hpvcd_status job_create_validated(hpvcd_gpu* g, const hpvcd_batch* b, hpvcd_job** out);

}  // namespace hpvcd
