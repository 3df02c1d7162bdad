// runtime.hip — host runtime of the reconstruction back end: record
// validation and preparation (availability stamps, wavefront units and
// dependency levels, residual work lists), device pools, the per-stage
// launch chain (batched, plain per-picture, or replayed from a hipGraph) and
// the C ABI of include/hpvcd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>
#include <type_traits>
#include <cstring>
#include <mutex>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"
#include "kernels.hip"

namespace hpvcd {

static thread_local std::string g_last_error;

hpvcd_status set_error(hpvcd_status st, const char* msg) {
    g_last_error = msg ? msg : "";
    return st;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            std::string m_ = std::string(#expr) + ": " + hipGetErrorString(e_);       \
            return set_error(HPVCD_E_DEVICE, m_.c_str());                              \
        }                                                                              \
    } while (0)

// Launch-time HIP calls inside launch_class (stream forks and joins) record
// their first failure here; the entry points check it after launching.
thread_local hipError_t t_launch_err = hipSuccess;
#define HIP_TRY_V(expr)                                                                \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess && t_launch_err == hipSuccess) t_launch_err = e_;        \
    } while (0)

// ---------------------------------------------------------------------------
// validation: every index a kernel dereferences is checked here
// ---------------------------------------------------------------------------
// Prediction units, reference entries and the TBs of inter CUs of picture p.
hpvcd_status validate_inter(const hpvcd_batch& b, uint32_t p) {
    const hpvcd_picture& P = b.pictures[p];
    if ((uint64_t)P.pu_begin + P.pu_count > b.num_pus) return set_error(HPVCD_E_INVALID, "pu range");
    if ((uint64_t)P.ref_begin + P.ref_count > b.num_refs) return set_error(HPVCD_E_INVALID, "ref range");
    for (uint32_t k = 0; k < P.ref_count; k++) {
        const hpvcd_ref& E = b.refs[P.ref_begin + k];
        if (E.log2_denom[0] > 7 || E.log2_denom[1] > 7) return set_error(HPVCD_E_INVALID, "weight denominator");
        if (E.picture == HPVCD_REF_MISSING) continue;
        if (E.picture >= p) return set_error(HPVCD_E_INVALID, "reference picture must come earlier in the batch");
        const hpvcd_picture& Q = b.pictures[E.picture];
        if (Q.width != P.width || Q.height != P.height || Q.chroma_idc != P.chroma_idc ||
            Q.bit_depth_luma != P.bit_depth_luma || Q.bit_depth_chroma != P.bit_depth_chroma ||
            ((Q.flags ^ P.flags) & HPVCD_PF_EXTENDED_PRECISION))
            return set_error(HPVCD_E_INVALID, "reference picture of another size or format");
    }
    if (!P.pu_count) {
        for (uint32_t t = 0; t < P.tb_count; t++)
            if (b.tbs[P.tb_begin + t].flags & HPVCD_TB_INTER)
                return set_error(HPVCD_E_INVALID, "inter TB outside every prediction unit");
        return HPVCD_OK;
    }
    const int gw = (P.width + 3) / 4, gh = (P.height + 3) / 4;
    std::vector<uint8_t> cover((size_t)gw * gh, 0);
    for (uint32_t i = 0; i < P.pu_count; i++) {
        const hpvcd_pu& U = b.pus[P.pu_begin + i];
        if (U.w < 4 || U.h < 4 || U.w > 64 || U.h > 64 || (U.w & 3) || (U.h & 3) || (U.x & 3) || (U.y & 3) ||
            U.x + U.w > P.width || U.y + U.h > P.height)
            return set_error(HPVCD_E_INVALID, "prediction unit geometry");
        if (U.flags & ~(HPVCD_PU_L0 | HPVCD_PU_L1)) return set_error(HPVCD_E_INVALID, "prediction unit flags");
        for (int l = 0; l < 2; l++)
            if (((U.flags >> l) & 1) && U.ref[l] >= P.ref_count)
                return set_error(HPVCD_E_INVALID, "prediction unit reference entry");
        if ((U.flags & 3) == 3) {
            const hpvcd_ref& E0 = b.refs[P.ref_begin + U.ref[0]];
            const hpvcd_ref& E1 = b.refs[P.ref_begin + U.ref[1]];
            if (E0.weighted != E1.weighted || E0.log2_denom[0] != E1.log2_denom[0] || E0.log2_denom[1] != E1.log2_denom[1])
                return set_error(HPVCD_E_INVALID, "bi-prediction from entries of different weight tables");
        }
        for (int yy = U.y; yy < U.y + U.h; yy += 4)
            for (int xx = U.x; xx < U.x + U.w; xx += 4) cover[(size_t)(yy / 4) * gw + xx / 4] = 1;
    }
    const int sw = sub_w_of(P.chroma_idc), sh = sub_h_of(P.chroma_idc);
    for (uint32_t t = 0; t < P.tb_count; t++) {
        const hpvcd_tb& T = b.tbs[P.tb_begin + t];
        if (!(T.flags & HPVCD_TB_INTER)) continue;
        if (T.flags & HPVCD_TB_PCM) return set_error(HPVCD_E_INVALID, "PCM TB with residual-coding flags");
        int n = 1 << T.log2n, x0 = T.comp ? T.x * sw : T.x, y0 = T.comp ? T.y * sh : T.y;
        int nx = T.comp ? n * sw : n, ny = T.comp ? n * sh : n;
        for (int yy = y0; yy < std::min(y0 + ny, P.height); yy += 4)
            for (int xx = x0; xx < std::min(x0 + nx, P.width); xx += 4)
                if (!cover[(size_t)(yy / 4) * gw + xx / 4])
                    return set_error(HPVCD_E_INVALID, "inter TB outside every prediction unit");
    }
    return HPVCD_OK;
}

// The checks of one picture's records (ranges, cells, TBs, levels, inter
// records); independent of the other pictures.
hpvcd_status validate_picture(const hpvcd_batch& b, uint32_t p) {
    const hpvcd_picture& P = b.pictures[p];
    if (P.width <= 0 || P.height <= 0 || P.width > 16384 || P.height > 16384 || (P.width & 7) || (P.height & 7))
        return set_error(HPVCD_E_BAD_DIMENSIONS, "picture size must be a positive multiple of 8");
    if (P.chroma_idc > 3) return set_error(HPVCD_E_UNSUPPORTED_CHROMA, "chroma_idc");
    // 4:2:2 (two stacked chroma TBs per luma TB, the remapped chroma
    // modes, chroma CTBs of ctb/2 x ctb) has no record producer and no
    // parity evidence yet, so the back end refuses it
    if (P.chroma_idc == 2) return set_error(HPVCD_E_UNSUPPORTED_CHROMA, "4:2:2 pictures are not supported");
    if (P.bit_depth_luma < 8 || P.bit_depth_luma > 12 || P.bit_depth_chroma < 8 || P.bit_depth_chroma > 12)
        return set_error(HPVCD_E_UNSUPPORTED_BITDEPTH, "bit depth must be 8..12");
    if ((P.bit_depth_luma > 8) != (P.bit_depth_chroma > 8))
        return set_error(HPVCD_E_UNSUPPORTED_BITDEPTH, "luma and chroma storage classes differ");
    if (P.log2_ctb < 4 || P.log2_ctb > 6) return set_error(HPVCD_E_PARAMSET, "log2_ctb must be 4..6");
    if (P.num_slices > HPVCD_MAX_SLICES) return set_error(HPVCD_E_INVALID, "num_slices");
    if ((uint64_t)P.tb_begin + P.tb_count > b.num_tbs) return set_error(HPVCD_E_INVALID, "tb range");
    if ((uint64_t)P.coeff_begin + P.coeff_count > b.num_coeffs) return set_error(HPVCD_E_INVALID, "coeff range");
    int ctb = 1 << P.log2_ctb;
    uint64_t nctb = (uint64_t)((P.width + ctb - 1) / ctb) * ((P.height + ctb - 1) / ctb);
    if ((uint64_t)P.ctb_begin + nctb > b.num_ctbs) return set_error(HPVCD_E_INVALID, "sao range");
    uint64_t ncell = (uint64_t)((P.width + 3) / 4) * ((P.height + 3) / 4);
    if ((uint64_t)P.cell_begin + ncell > b.num_cells) return set_error(HPVCD_E_INVALID, "cell range");
    // the loop filters index slices[] by the cell's slice
    const uint32_t nsl = P.num_slices ? P.num_slices : 1u;
    for (uint64_t g = 0; g < ncell; g++)
        if (HPVCD_CELL_SLICE(b.cell_info[P.cell_begin + g]) >= nsl)
            return set_error(HPVCD_E_INVALID, "cell slice index beyond num_slices");
    if (P.scaling_index >= 0 && (uint32_t)P.scaling_index >= b.num_scaling)
        return set_error(HPVCD_E_INVALID, "scaling_index");
    if (P.out_index >= b.num_outputs) return set_error(HPVCD_E_INVALID, "out_index");
    const hpvcd_output& O = b.outputs[P.out_index];
    const bool yuv = HPVCD_OUT_IS_YUV(O.kind);
    // raw planes may copy past the window (clamped to the coded plane)
    if (P.copy_w < 0 || P.copy_h < 0 || (!yuv && (P.copy_w > P.win_w || P.copy_h > P.win_h)) || P.dst_x < 0 ||
        P.dst_y < 0 || (int64_t)P.dst_x + P.copy_w > O.width || (int64_t)P.dst_y + P.copy_h > O.height)
        return set_error(HPVCD_E_INVALID, "picture window placement outside its output image");
    if (yuv) {
        if (HPVCD_OUT_YUV_CHROMA(O.kind) != P.chroma_idc)
            return set_error(HPVCD_E_INVALID, "raw-plane output of another chroma format");
        if ((HPVCD_OUT_YUV_WIDE(O.kind) != 0) != (P.bit_depth_luma > 8) ||
            (P.chroma_idc && (P.bit_depth_chroma > 8) != (P.bit_depth_luma > 8)))
            return set_error(HPVCD_E_INVALID, "raw-plane output sample size");
        if (O.channels != (P.chroma_idc ? 3 : 1)) return set_error(HPVCD_E_INVALID, "output channels");
        if (P.win_w <= 0 || P.win_h <= 0) return set_error(HPVCD_E_INVALID, "picture window");
    }
    const bool luma_out = O.kind == HPVCD_OUT_LUMA8 || O.kind == HPVCD_OUT_LUMA16;
    if (luma_out && P.chroma_idc != 0) return set_error(HPVCD_E_INVALID, "luma output of a colour picture");
    if (!yuv && O.channels != (luma_out ? 1 : 3)) return set_error(HPVCD_E_INVALID, "output channels");
    bool wide = O.kind == HPVCD_OUT_RGB16 || O.kind == HPVCD_OUT_LUMA16;
    if (wide && P.bit_depth_luma == 8) return set_error(HPVCD_E_INVALID, "16-bit output of an 8-bit picture");
    int sw = sub_w_of(P.chroma_idc), sh = sub_h_of(P.chroma_idc);
    int cw = chroma_w(P), ch = chroma_h(P);
    int64_t last_ctb[3] = {-1, -1, -1};
    int ctb_cols = (P.width + ctb - 1) / ctb;
    for (uint32_t t = 0; t < P.tb_count; t++) {
        const hpvcd_tb& T = b.tbs[P.tb_begin + t];
        if (T.log2n < 2 || T.log2n > 5) return set_error(HPVCD_E_INVALID, "tb log2 size");
        if (T.comp <= 2) {
            int lx = T.comp ? T.x * sub_w_of(P.chroma_idc) : T.x, ly = T.comp ? T.y * sub_h_of(P.chroma_idc) : T.y;
            int64_t r = (int64_t)(ly >> P.log2_ctb) * ctb_cols + (lx >> P.log2_ctb);
            if (r < last_ctb[T.comp]) return set_error(HPVCD_E_INVALID, "tbs not in CTB raster order");
            last_ctb[T.comp] = r;
        }
        if (T.comp > 2 || (T.comp && !P.chroma_idc)) return set_error(HPVCD_E_INVALID, "tb component");
        int n = 1 << T.log2n;
        int pw = T.comp ? cw : P.width, ph = T.comp ? ch : P.height;
        int ctbw = T.comp ? ctb / sw : ctb, ctbh = T.comp ? ctb / sh : ctb;
        if (n > ctbw || n > ctbh || (T.x % ctbw) + n > ctbw || (T.y % ctbh) + n > ctbh)
            return set_error(HPVCD_E_INVALID, "tb crosses a CTB boundary");
        if (T.x >= pw || T.y >= ph) return set_error(HPVCD_E_INVALID, "tb outside its plane");
        if ((T.x & 3) || (T.y & 3)) return set_error(HPVCD_E_INVALID, "tb not 4-aligned");
        if (T.mode > 34) return set_error(HPVCD_E_INVALID, "intra mode");
        if (T.res_scale && (T.comp == 0 || P.chroma_idc != 3)) return set_error(HPVCD_E_INVALID, "res_scale");
        const bool pcm = (T.flags & HPVCD_TB_PCM) != 0;
        const int bd = T.comp ? P.bit_depth_chroma : P.bit_depth_luma;
        if (pcm) {
            if (T.flags & (HPVCD_TB_TS | HPVCD_TB_TQB | HPVCD_TB_INTER | HPVCD_TB_RDPCM_H | HPVCD_TB_RDPCM_V |
                           HPVCD_TB_ACT) || T.res_scale)
                return set_error(HPVCD_E_INVALID, "PCM TB with residual-coding flags");
            if (T.qp >= bd) return set_error(HPVCD_E_INVALID, "PCM sample shift must be below the bit depth");
        }
        if (T.flags & HPVCD_TB_ACT) {
            // a luma TB opens the triple; its Cb and Cr TBs follow at once
            if (P.chroma_idc != 3 || T.comp > 2 || t < T.comp || t - T.comp + 2 >= P.tb_count)
                return set_error(HPVCD_E_INVALID, "ACT TB outside a 4:4:4 luma/Cb/Cr triple");
            const uint32_t t0 = t - T.comp;
            for (int k = 0; k <= 2; k++) {
                const hpvcd_tb& U = b.tbs[P.tb_begin + t0 + k];
                if (!(U.flags & HPVCD_TB_ACT) || U.comp != k || U.x != T.x || U.y != T.y || U.log2n != T.log2n)
                    return set_error(HPVCD_E_INVALID, "ACT triple not co-located luma, Cb, Cr of one size");
            }
        }
        if (T.flags & (HPVCD_TB_CBF | HPVCD_TB_PCM)) {
            if ((uint64_t)T.coeff_off + T.coeff_cnt > P.coeff_count) return set_error(HPVCD_E_INVALID, "tb coeff range");
            if (T.coeff_cnt > (uint32_t)(n * n)) return set_error(HPVCD_E_INVALID, "more coefficient words than TB positions");
            for (uint32_t k = 0; k < T.coeff_cnt; k++) {
                uint32_t w = b.coeffs[P.coeff_begin + T.coeff_off + k];
                if (HPVCD_COEFF_POS(w) >= (uint32_t)(n * n)) return set_error(HPVCD_E_INVALID, "coefficient position");
                if (pcm && (HPVCD_COEFF_LEVEL(w) < 0 || HPVCD_COEFF_LEVEL(w) >= (1 << (bd - T.qp))))
                    return set_error(HPVCD_E_INVALID, "PCM sample outside PcmBitDepth");
            }
        }
    }
    hpvcd_status st = validate_inter(b, p);
    if (st != HPVCD_OK) return st;
    return HPVCD_OK;
}

hpvcd_status validate_batch(const hpvcd_batch& b) {
    if (b.num_pictures && !b.pictures) return set_error(HPVCD_E_INVALID, "pictures is null");
    if (b.num_tbs && !b.tbs) return set_error(HPVCD_E_INVALID, "tbs is null");
    if (b.num_coeffs && !b.coeffs) return set_error(HPVCD_E_INVALID, "coeffs is null");
    if (b.num_ctbs && !b.sao) return set_error(HPVCD_E_INVALID, "sao is null");
    if (b.num_cells && !b.cell_info) return set_error(HPVCD_E_INVALID, "cell_info is null");
    if (b.num_outputs && !b.outputs) return set_error(HPVCD_E_INVALID, "outputs is null");
    if (b.num_scaling && !b.scaling) return set_error(HPVCD_E_INVALID, "scaling is null");
    if (b.num_pus && !b.pus) return set_error(HPVCD_E_INVALID, "pus is null");
    if (b.num_refs && !b.refs) return set_error(HPVCD_E_INVALID, "refs is null");
    for (uint32_t o = 0; o < b.num_outputs; o++) {
        const hpvcd_output& O = b.outputs[o];
        if (O.width <= 0 || O.height <= 0 || O.width > 65536 || O.height > 65536)
            return set_error(HPVCD_E_BAD_DIMENSIONS, "output image size");
        if (O.kind > HPVCD_OUT_YUV400_16) return set_error(HPVCD_E_INVALID, "output kind");
        if (O.orientation > HPVCD_ORIENT_ROTATE_270) return set_error(HPVCD_E_INVALID, "output orientation");
    }
    // pictures are checked on the host threads; the first failing picture
    // (in batch order) reports
    const uint32_t np = b.num_pictures;
    std::vector<hpvcd_status> st(np, HPVCD_OK);
    std::vector<std::string> msg(np);
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        for (uint32_t p; (p = next.fetch_add(1)) < np;) {
            st[p] = validate_picture(b, p);
            if (st[p]) msg[p] = g_last_error;
        }
    };
    const int nth = std::max(1, std::min<int>((int)std::max(1u, std::min(16u, std::thread::hardware_concurrency())),
                                              (int)np));
    std::vector<std::thread> pool;
    for (int t = 1; t < nth; t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    for (uint32_t p = 0; p < np; p++)
        if (st[p]) return set_error(st[p], msg[p].c_str());
    return HPVCD_OK;
}

}  // namespace hpvcd

using namespace hpvcd;

// ---------------------------------------------------------------------------
// device context
// ---------------------------------------------------------------------------
struct hpvcd_gpu {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;                   // per-TB preparation, concurrent with the residual stage
    hipEvent_t fork = nullptr, join = nullptr;
    int num_cus = 0;
    int occ_intra[4] = {0, 0, 0, 0};   // blocks per CU for k_intra<T,R> (T16*2 + R32)
    // pinned staging of record uploads, reused by every job of this GPU
    // (grown on demand; job creation holds `mu` while it stages and copies)
    std::mutex mu;
    uint8_t* pin = nullptr;
    size_t pin_size = 0;
    // one launch chain at a time on the GPU's streams (the fork/join events
    // and graph captures are per GPU); decode calls from several threads
    // interleave whole chains
    std::mutex launch_mu;
    // this device's copy of the allocation-free entry point's scratch array
    uint8_t* scratch = nullptr;
    // launch shapes; `tuning_gen` counts changes, so that graphs captured
    // with older shapes are captured again
    hpvcd_tuning tuning;
    uint32_t tuning_gen = 0;
    char arch[64] = {0};
};

// Scratch of hpvcd_gpu_decode_batch_noalloc: every device array of one batch
// (uploaded records, output images, planes, residuals, sync and border
// words) is laid out inside it.  Each GPU the library runs on holds its own
// copy until the process ends.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_setup_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};
// calls of the allocation-free entry point on one device share its array:
// they run one at a time within the process
std::mutex g_scratch_mu[kMaxDevices];

// The device address of `scratch` on `device` (current), looked up once per
// device; the array's name is a host-side stand-in, never a device address.
hpvcd_status scratch_address(int device, uint8_t** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return set_error(HPVCD_E_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_setup_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            std::string m = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return set_error(HPVCD_E_DEVICE, m.c_str());
        }
        g_scratch_ptr[device] = static_cast<uint8_t*>(ptr);
    }
    *out = g_scratch_ptr[device];
    return HPVCD_OK;
}

}  // namespace

namespace {

struct DevBuf {
    char* base = nullptr;
    size_t size = 0, used = 0;
    // next 256-byte aligned block of count elements (null while sizing, when
    // base is not allocated yet)
    template <typename X>
    X* take(size_t count) {
        size_t off = (used + 255) & ~(size_t)255;
        used = off + count * sizeof(X);
        return base ? reinterpret_cast<X*>(base + off) : nullptr;
    }
};

// Allocator whose resize() leaves trivially constructible elements
// uninitialised: the merged arrays below are written element by element
// right after they are sized, so zero-filling them first would only cost a
// serial pass over tens of megabytes.
template <typename T>
struct UninitAlloc : std::allocator<T> {
    template <typename U> struct rebind { using other = UninitAlloc<U>; };
    UninitAlloc() = default;
    template <typename U> UninitAlloc(const UninitAlloc<U>&) noexcept {}
    template <typename U> void construct(U* p) noexcept { ::new (static_cast<void*>(p)) U; }
    template <typename U, typename... A> void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};
template <typename T> using PodVec = std::vector<T, UninitAlloc<T>>;

// rows of a picture's raw-plane copy (zero for the other output kinds)
inline uint32_t yuv_rows(const DevPic& D) {
    if (!HPVCD_OUT_IS_YUV(D.out_kind)) return 0;
    return (uint32_t)(D.copy_h + 2 * yuv_copy(D).cch);
}

struct ClassJob {
    bool samp16 = false, res32 = false;
    std::vector<uint32_t> batch_pics;     // batch picture index per class-local picture
    // host-prepared arrays
    std::vector<DevPic> pics;
    PodVec<hpvcd_tb> tbs;
    PodVec<uint32_t> tb_pic, coeffs, cells, order;
    PodVec<hpvcd_sao> sao;
    PodVec<DevUnit> units;
    std::vector<uint32_t> tickets, plain_tickets;
    PodVec<uint32_t> unit_tbs;
    PodVec<PrepTb> unit_recs;                      // K2a's fields of each unit_tbs entry
    std::vector<uint32_t> prep_list;               // units K2a prepares (Cr units sharing Cb's are not)
    std::vector<uint32_t> plain_prep, pic_prep_begin; // the same units per picture (plain mode) and segments
    PodVec<uint16_t> level_off;
    PodVec<uint32_t> unit_refoff;                  // per unit TB entry: map offset | sample offset / 16 << 13 | level << 21
    size_t smap_words = 0, rmap_words = 0;
    std::vector<uint32_t> res_list[2][4];            // [pass][log2-2], batched order
    PodVec<DevResJob> res_jobs[2][4];                // K1 job of each res_list entry
    std::vector<uint32_t> pic_units_begin;            // plain_tickets segment per picture (npics+1)
    std::vector<uint32_t> pic_res_begin[2][4];        // per picture list segments (npics+1)
    std::vector<uint32_t> act_list, pic_act_begin;    // luma TB of each ACT triple; per picture segments
    std::vector<DevPu> pus;
    std::vector<DevMcTask> mc_tasks;
    std::vector<uint32_t> pic_mc_begin;               // MC task segment per picture (npics+1)
    // dependency waves: pictures are sorted by reference depth; wave w is the
    // picture range [wave_begin[w], wave_begin[w+1]) and its wavefront units
    // are tickets [wave_tick_begin[w], wave_tick_begin[w+1])
    std::vector<uint32_t> wave_begin, wave_tick_begin;
    std::vector<uint8_t> pic_is_ref;                  // referenced by a later picture: always completes SAO
    size_t plane_elems = 0, res_elems = 0;
    uint32_t max_luma_v = 0, max_luma_h = 0, max_chroma_v = 0, max_chroma_h = 0, max_plane = 0, max_copy = 0;
    uint32_t max_ctbs = 0, max_rgb_qx = 0, max_rgb_rows = 0;  // K4 blocks; K5 quads per row, rows
    uint32_t max_yuv_rows = 0;                                  // rows of raw-plane copies (luma + 2 chroma)
    // device pointers
    DevPic* d_pics = nullptr;
    hpvcd_tb* d_tbs = nullptr;
    uint32_t *d_tb_pic = nullptr, *d_coeffs = nullptr, *d_cells = nullptr, *d_order = nullptr;
    hpvcd_sao* d_sao = nullptr;
    DevUnit* d_units = nullptr;
    uint32_t *d_tickets = nullptr, *d_plain_tickets = nullptr;
    PrepTb* d_unit_recs = nullptr;
    uint32_t* d_prep_list = nullptr;
    uint32_t* d_plain_prep = nullptr;
    uint16_t* d_level_off = nullptr;
    uint32_t* d_unit_refoff = nullptr;
    uint32_t* d_smap = nullptr;   // per-sample reference words
    uint16_t* d_rmap = nullptr;   // substituted reference maps
    uint4* d_desc = nullptr;      // wavefront TB descriptors (per unit TB entry)
    DevResJob* d_res_jobs[2][4] = {};
    uint32_t* d_act_list = nullptr;
    DevPu* d_pus = nullptr;
    DevMcTask* d_mc_tasks = nullptr;
    void *d_planes_a = nullptr, *d_planes_b = nullptr, *d_res = nullptr;
    uint32_t* d_sync = nullptr;   // [0] err, [1] border tag, [2..] ticket counters (batched + one per picture)
    uint32_t sync_words = 0;
    uint32_t* d_border = nullptr; // kBorderWords per unit
    hipGraphExec_t graph = nullptr;
    int graph_stop = -1;
    uint32_t graph_gen = 0;    // the GPU's tuning generation the graph was captured with
};

}  // namespace

struct hpvcd_job {
    hpvcd_gpu* g = nullptr;
    std::vector<ClassJob> classes;
    std::vector<hpvcd_output> outputs;
    std::vector<size_t> out_off;
    DevBuf mem;
    uint8_t* d_out = nullptr;
    uint8_t* d_scaling = nullptr;
    size_t out_bytes = 0;
    int last_stop = HPVCD_STAGE_RGB;
    std::vector<uint32_t> pic_class, pic_local;   // batch picture -> (class, local index)
    double prep_ms = 0.0, upload_ms = 0.0;        // host preparation; staging + H2D + pool clears
    bool mem_external = false;                    // the pool is the scratch array (not freed)
};

namespace {


// K1 job of one residual TB: the flag logic of src/decode.rs:4356
// (rext_post_ops), :3660 (ACT keeps 32-bit residuals) and
// src/transform.rs:764-910 (scaling-list choice) decided once on the host.
DevResJob make_res_job(const hpvcd_batch& b, const DevPic& D, const hpvcd_tb& T, uint32_t coeff_begin) {
    DevResJob J;
    std::memset(&J, 0, sizeof(J));
    const int comp = T.comp, n = 1 << T.log2n;
    const int bd = comp ? D.bd_c : D.bd;
    const bool ep = (D.flags & HPVCD_PF_EXTENDED_PRECISION) != 0;
    const bool inter = (T.flags & HPVCD_TB_INTER) != 0;
    const bool ts = (T.flags & HPVCD_TB_TS) != 0;
    const bool tqb = (T.flags & HPVCD_TB_TQB) != 0;
    const bool cbf = (T.flags & HPVCD_TB_CBF) != 0;
    const bool pcm = (T.flags & HPVCD_TB_PCM) != 0;
    const bool i16 = bd <= 8 && !ep && (comp == 0 || T.res_scale == 0) && !(T.flags & HPVCD_TB_ACT);
    const bool transform = cbf && !ts && !tqb && !pcm;
    uint16_t f = 0;
    if (cbf) f |= RJ_CBF;
    if (ts) f |= RJ_TS;
    if (tqb) f |= RJ_TQB;
    if (pcm) f |= RJ_PCM;
    if (transform) f |= RJ_TRANSFORM;
    if (transform && comp == 0 && n == 4 && !inter) f |= RJ_DST;
    if (i16) f |= RJ_I16;
    if (ep) f |= RJ_EP;
    if (comp) f |= RJ_CHROMA;
    if (T.res_scale != 0) f |= RJ_CCP;
    if (cbf && (ts || tqb)) {
        if (ts && !tqb && (D.flags & HPVCD_PF_TS_ROTATION) && n == 4 && !inter) f |= RJ_ROTATE;
        if (inter) {
            if (T.flags & HPVCD_TB_RDPCM_V) f |= RJ_RDPCM | RJ_RDPCM_V;
            else if (T.flags & HPVCD_TB_RDPCM_H) f |= RJ_RDPCM;
        } else if ((D.flags & HPVCD_PF_IMPLICIT_RDPCM) && (T.mode == 10 || T.mode == 26)) {
            f |= RJ_RDPCM | (T.mode == 26 ? RJ_RDPCM_V : 0);
        }
    }
    J.flags = f;
    const int max_qp = 51 + 6 * (bd - 8);
    J.qp = (uint8_t)std::min<int>(T.qp, max_qp);
    J.nx = T.nx;
    J.res_scale = T.res_scale;
    J.bds = (uint8_t)((D.bd - 8) | ((D.bd_c - 8) << 4));
    const int pw = comp ? D.cw : D.w, ph = comp ? D.ch : D.h;
    J.stride = (uint16_t)pw;
    J.vw = (uint8_t)std::max(0, std::min(n, pw - (int)T.x));
    J.vh = (uint8_t)std::max(0, std::min(n, ph - (int)T.y));
    J.out = D.res_off[comp] + (uint64_t)T.y * pw + T.x;
    J.luma_delta = (uint32_t)(D.res_off[comp] - D.res_off[0]);
    J.coeff = coeff_begin + T.coeff_off;
    J.coeff_cnt = (uint16_t)T.coeff_cnt;
    J.scaling = -1;
    if (!tqb && D.scaling_index >= 0 && !(ts && n != 4)) {
        const int size_id = T.log2n - 2;
        const int matrix_id = (comp < 2 ? comp : 2) + (inter ? 3 : 0);
        const size_t off = (size_t)D.scaling_index * HPVCD_SCALING_SET_BYTES +
                           (size_t)(size_id * 6 + matrix_id) * HPVCD_SCALING_ENTRY_BYTES;
        if (!b.scaling[off + 65]) J.scaling = (int32_t)off;  // a flat-16 list is the flat path
    }
    return J;
}

struct UnitKey { uint32_t t, pic, plane, idx; };

// One picture's share of a class's prepared arrays.  TB, unit, descriptor,
// level and map indices are local to the picture until prepare_class merges
// the pictures (in order) into the class arrays.
struct PicPrep {
    DevPic D;
    std::vector<hpvcd_tb> tbs;
    std::vector<uint32_t> order;
    std::vector<DevUnit> units;
    std::vector<uint32_t> unit_tbs, unit_refoff, prep_units;
    std::vector<PrepTb> unit_recs;
    std::vector<uint16_t> level_off;
    std::vector<UnitKey> keys;
    std::vector<uint32_t> res[2][4], act;
    size_t smap_words = 0, rmap_words = 0;
    hpvcd_status st = HPVCD_OK;
    std::string msg;
};

// Availability stamps, wavefront units with their dependency levels and
// reference-map layout, and residual lists of class-local picture lp.
hpvcd_status prepare_picture(const hpvcd_batch& b, uint32_t lp, const hpvcd_picture& P,
                             const std::vector<size_t>& out_off, PicPrep& Q) {
    DevPic& D = Q.D;
    std::memset(&D, 0, sizeof(D));
    D.w = P.width; D.h = P.height;
    D.cw = chroma_w(P); D.ch = chroma_h(P);
    D.gw = (P.width + 3) / 4; D.gh = (P.height + 3) / 4;
    D.sub_w_log2 = sub_w_of(P.chroma_idc) == 2 ? 1 : 0;
    D.sub_h_log2 = sub_h_of(P.chroma_idc) == 2 ? 1 : 0;
    D.ctb_log2 = P.log2_ctb;
    const int ctb = 1 << P.log2_ctb;
    D.ctb_cols = (P.width + ctb - 1) / ctb; D.ctb_rows = (P.height + ctb - 1) / ctb;
    D.bd = P.bit_depth_luma; D.bd_c = P.bit_depth_chroma;
    D.flags = P.flags;
    D.num_slices = P.num_slices;
    std::memcpy(D.slices, P.slices, sizeof(D.slices));
    D.pps_cb_qp_offset = P.pps_cb_qp_offset; D.pps_cr_qp_offset = P.pps_cr_qp_offset;
    D.chroma_idc = P.chroma_idc; D.matrix = P.matrix;
    D.crop_left = P.crop_left; D.crop_top = P.crop_top; D.win_w = P.win_w; D.win_h = P.win_h;
    D.dst_x = P.dst_x; D.dst_y = P.dst_y; D.copy_w = P.copy_w; D.copy_h = P.copy_h;
    D.out_kind = b.outputs[P.out_index].kind;
    D.out_width = b.outputs[P.out_index].width;
    D.out_height = b.outputs[P.out_index].height;
    D.orient = b.outputs[P.out_index].orientation;
    D.out_off = out_off[P.out_index];
    const size_t ny = (size_t)D.w * D.h, nc = (size_t)D.cw * D.ch;
    // plane offsets relative to the picture (the merge adds the class base)
    D.plane_off[0] = 0;
    D.plane_off[1] = ny;
    D.plane_off[2] = ny + nc;
    D.scaling_index = P.scaling_index;
    bool all_across = true, deb = false;
    for (int s = 0; s < std::max<int>(1, P.num_slices); s++) {
        if (P.num_slices) { all_across &= P.slices[s].lf_across != 0; deb |= !P.slices[s].disabled; }
        else deb = !(P.flags & HPVCD_PF_DEBLOCK_DISABLED);
    }
    D.sao_across_slices = (P.flags & HPVCD_PF_LF_ACROSS_SLICES) && all_across;
    D.deblock_on = deb;
    D.inter = P.pu_count > 0;

    const hpvcd_tb* src_tb = b.tbs + P.tb_begin;
    const uint32_t ncell = (uint32_t)(D.gw * D.gh);
    uint32_t suppress = 0;
    for (uint32_t g = 0; g < ncell; g++) {
        uint32_t inf = b.cell_info[P.cell_begin + g];
        if ((inf & HPVCD_CELL_TQB) || ((P.flags & HPVCD_PF_PCM_LF_DISABLED) && (inf & HPVCD_CELL_PCM))) suppress = 1;
    }
    D.suppress_any = suppress;

    // availability stamps: luma TB k gets stamp k (>=1) painted over its
    // cells; a chroma TB sees every luma TB decoded before it.
    Q.order.assign(ncell, 0xffffffffu);
    uint32_t stamp = 0;
    Q.tbs.reserve(P.tb_count);
    for (uint32_t t = 0; t < P.tb_count; t++) {
        hpvcd_tb T = src_tb[t];
        T.flags &= 0xffu;
        if (T.comp == 0) {
            stamp++;
            T.order = stamp;
            int n = 1 << T.log2n;
            for (int yy = T.y; yy < std::min(T.y + n, D.h); yy += 4)
                for (int xx = T.x; xx < std::min(T.x + n, D.w); xx += 4)
                    Q.order[(size_t)(yy / 4) * D.gw + xx / 4] = stamp;
        } else {
            T.order = stamp + 1;
        }
        if (T.flags & HPVCD_TB_PCM) T.flags |= HPVCD_TB_CBF;  // the samples travel as levels
        // ACT mixes all three residuals: each TB of a triple gets one
        if ((T.flags & (HPVCD_TB_CBF | HPVCD_TB_ACT)) || T.res_scale != 0) T.flags |= TBX_HAS_RES;
        Q.tbs.push_back(T);
    }

    // wavefront units: (plane, CTB) in decode order of their TBs
    const int nplanes = P.chroma_idc ? 3 : 1;
    const uint32_t nunits = (uint32_t)(nplanes * D.ctb_cols * D.ctb_rows);
    std::vector<std::vector<uint32_t>> utbs(nunits);
    // inter TBs leave K1c reconstructed: the wavefront only publishes
    // their samples on the unit's bottom row and right column
    std::vector<uint64_t> inter_bottom(nunits, 0), inter_right(nunits, 0);
    for (uint32_t t = 0; t < P.tb_count; t++) {
        const hpvcd_tb& T = Q.tbs[t];
        int sw = T.comp ? D.sub_w_log2 : 0, sh = T.comp ? D.sub_h_log2 : 0;
        int cx = (T.x << sw) >> P.log2_ctb, cy = (T.y << sh) >> P.log2_ctb;
        uint32_t u = (uint32_t)((T.comp * D.ctb_rows + cy) * D.ctb_cols + cx);
        if (T.flags & HPVCD_TB_INTER) {
            const int ctbw = (1 << P.log2_ctb) >> sw, ctbh = (1 << P.log2_ctb) >> sh, n = 1 << T.log2n;
            const int x0 = T.x - cx * ctbw, y0 = T.y - cy * ctbh;
            const uint64_t run = (n >= 64 ? ~0ull : ((1ull << n) - 1ull));
            if (y0 + n == ctbh) inter_bottom[u] |= run << x0;
            if (x0 + n == ctbw) inter_right[u] |= run << y0;
        } else {
            utbs[u].push_back(t);
        }
        if (T.flags & TBX_HAS_RES) {
            int pass = T.res_scale != 0 ? 1 : 0;
            Q.res[pass][T.log2n - 2].push_back(t);
        }
        if ((T.flags & HPVCD_TB_ACT) && T.comp == 0) Q.act.push_back(t);
    }
    for (int pl = 0; pl < nplanes; pl++)
        for (int cy = 0; cy < D.ctb_rows; cy++)
            for (int cx = 0; cx < D.ctb_cols; cx++) {
                uint32_t u = (uint32_t)((pl * D.ctb_rows + cy) * D.ctb_cols + cx);
                DevUnit U;
                std::memset(&U, 0, sizeof(U));
                U.pic = lp; U.cx = (uint16_t)cx; U.cy = (uint16_t)cy; U.plane = (uint8_t)pl;
                U.inter_bottom = inter_bottom[u];
                U.inter_right = inter_right[u];
                // dependency levels inside the unit
                int sw = pl ? D.sub_w_log2 : 0, sh = pl ? D.sub_h_log2 : 0;
                int ctbw = ctb >> sw, ctbh = ctb >> sh, cgw = ctbw / 4, cgh = ctbh / 4;
                int ux0 = cx * ctbw, uy0 = cy * ctbh;
                std::vector<uint16_t> lmap((size_t)cgw * cgh, 0);
                std::vector<std::pair<uint16_t, uint32_t>> lt;  // (level, tb)
                for (uint32_t tbi : utbs[u]) {
                    const hpvcd_tb& T = Q.tbs[tbi];
                    int n = 1 << T.log2n, x0 = T.x - ux0, y0 = T.y - uy0;
                    int lvl = 0;
                    auto probe = [&](int rx, int ry) {
                        if (rx < 0 || ry < 0 || rx >= ctbw || ry >= ctbh) return;
                        lvl = std::max<int>(lvl, lmap[(ry / 4) * cgw + rx / 4]);
                    };
                    for (int k = -1; k < 2 * n; k += 4) { probe(x0 - 1, y0 + k); probe(x0 + k, y0 - 1); }
                    probe(x0 - 1, y0 + 2 * n - 1);
                    probe(x0 + 2 * n - 1, y0 - 1);
                    for (int yy = y0; yy < y0 + n; yy += 4)
                        for (int xx = x0; xx < x0 + n; xx += 4) lmap[(yy / 4) * cgw + xx / 4] = (uint16_t)(lvl + 1);
                    lt.push_back({(uint16_t)lvl, tbi});
                }
                std::stable_sort(lt.begin(), lt.end(),
                                 [](const std::pair<uint16_t, uint32_t>& a, const std::pair<uint16_t, uint32_t>& b) {
                                     return a.first < b.first;
                                 });
                int nlev = lt.empty() ? 0 : lt.back().first + 1;
                if (nlev > 255) return set_error(HPVCD_E_INVALID, "too many dependency levels in one CTB");
                U.nlevels = (uint8_t)nlev;
                if (lt.size() > (size_t)kMaxUnitTbs) return set_error(HPVCD_E_INVALID, "too many TBs in one CTB");
                U.ntb = (uint16_t)lt.size();
                // groups of up to four consecutive independent (same
                // level) 4x4 TBs that take the fast path: not DC, or I_PCM
                auto groupable = [&](uint32_t tbi) {
                    const hpvcd_tb& T = Q.tbs[tbi];
                    return T.log2n == 2 && (T.mode != 1 || (T.flags & HPVCD_TB_PCM));
                };
                // per step leader also: whether a TB of the step writes
                // the unit's bottom row or right column, whether one is
                // planar or H/V (not I_PCM), and the next step's size
                std::vector<size_t> leaders;
                for (size_t a = 0; a < lt.size();) {
                    size_t e = a + 1;
                    if (groupable(lt[a].second))
                        while (e < lt.size() && e - a < 4 && lt[e].first == lt[a].first && groupable(lt[e].second)) e++;
                    uint16_t f = (uint16_t)((e - a) << TBX_GROUP_SHIFT);
                    for (size_t m = a; m < e; m++) {
                        const hpvcd_tb& M = Q.tbs[lt[m].second];
                        const int nm = 1 << M.log2n;
                        if ((M.y - uy0) + nm == ctbh || (M.x - ux0) + nm == ctbw) f |= TBX_STEP_PUB;
                        if (!(M.flags & HPVCD_TB_PCM) && (M.mode == 0 || M.mode == 10 || M.mode == 26))
                            f |= TBX_STEP_PLAIN;
                    }
                    Q.tbs[lt[a].second].flags |= f;
                    leaders.push_back(a);
                    a = e;
                }
                for (size_t k = 0; k + 1 < leaders.size(); k++) {
                    const size_t gnext = (k + 2 < leaders.size() ? leaders[k + 2] : lt.size()) - leaders[k + 1];
                    Q.tbs[lt[leaders[k]].second].flags |= (uint16_t)((gnext - 1) << TBX_NEXT_SHIFT);
                }
                U.tb_list = (uint32_t)Q.unit_tbs.size();
                U.level_list = (uint32_t)Q.level_off.size();
                size_t k = 0;
                for (int l = 0; l <= nlev; l++) {
                    while (k < lt.size() && lt[k].first < l) k++;
                    Q.level_off.push_back((uint16_t)k);
                }
                // per TB: reference map offset, sample word offset (a
                // multiple of 16) and level
                uint32_t roff = 0, soff = 0;
                // (map offsets of a unit are multiples of 8 entries: the
                // wavefront stages them in 16-byte vectors)
                Q.rmap_words = (Q.rmap_words + 7) & ~(size_t)7;
                U.sbase = (uint32_t)Q.smap_words;
                U.rbase = (uint32_t)Q.rmap_words;
                for (auto& e : lt) {
                    const uint32_t n = 1u << Q.tbs[e.second].log2n;
                    if (roff + 4 * n + 1 > (uint32_t)kMaxUnitRefs || soff + n * n > (uint32_t)(kMaxCtb * kMaxCtb))
                        return set_error(HPVCD_E_INVALID, "too many reference samples in one CTB");
                    Q.unit_tbs.push_back(e.second);
                    const hpvcd_tb& R = Q.tbs[e.second];
                    Q.unit_recs.push_back(PrepTb{R.x, R.y, R.log2n, R.mode, R.flags, R.order});
                    Q.unit_refoff.push_back(roff | ((soff >> 4) << 13) | ((uint32_t)e.first << 21));
                    roff += 4 * n + 1;
                    soff += n * n;
                }
                U.nsamp = soff;
                U.nref = roff;
                U.desc_base = U.tb_list;
                // Cr repeats Cb's TB geometry, modes and availability in
                // every conforming stream: when the two units' TBs agree
                // in everything K2a reads, Cr takes Cb's descriptors,
                // reference maps and words (residuals are read whole from
                // the pool, so residual presence does not matter)
                bool shared = false;
                if (pl == 2) {
                    const DevUnit& B = Q.units[Q.units.size() - (size_t)D.ctb_rows * D.ctb_cols];
                    shared = B.ntb == U.ntb;
                    const uint16_t keep = HPVCD_TB_PCM | 0xfe00u;  // PCM and the step bits
                    for (uint32_t e = 0; shared && e < U.ntb; e++) {
                        const PrepTb& a = Q.unit_recs[B.tb_list + e];
                        const PrepTb& c = Q.unit_recs[U.tb_list + e];
                        shared = a.x == c.x && a.y == c.y && a.log2n == c.log2n && a.mode == c.mode &&
                                 a.order == c.order && (a.flags & keep) == (c.flags & keep);
                    }
                    if (shared) {
                        U.sbase = B.sbase;
                        U.rbase = B.rbase;
                        U.desc_base = B.desc_base;
                    }
                }
                if (!shared) {
                    Q.smap_words += soff;
                    Q.rmap_words += roff;
                    Q.prep_units.push_back(u);
                }
                Q.units.push_back(U);
                Q.keys.push_back({(uint32_t)(cx + 2 * cy), lp, (uint32_t)pl, u});
            }
    return HPVCD_OK;
}

// Availability stamps, units, levels and residual lists for one class: the
// pictures are prepared on `threads` host threads, then merged in order.
hpvcd_status prepare_class(const hpvcd_batch& b, ClassJob& C, const std::vector<size_t>& out_off,
                           const std::vector<uint32_t>& depth, const std::vector<uint32_t>& pic_local,
                           const std::vector<uint8_t>& is_ref, int threads) {
    const uint32_t np = (uint32_t)C.batch_pics.size();
    std::vector<PicPrep> Q(np);
    {
        std::atomic<uint32_t> next{0};
        auto work = [&]() {
            for (uint32_t lp; (lp = next.fetch_add(1)) < np;) {
                Q[lp].st = prepare_picture(b, lp, b.pictures[C.batch_pics[lp]], out_off, Q[lp]);
                if (Q[lp].st) Q[lp].msg = g_last_error;
            }
        };
        const int nth = std::max(1, std::min<int>(threads, (int)np));
        std::vector<std::thread> pool;
        for (int t = 1; t < nth; t++) pool.emplace_back(work);
        work();
        for (auto& t : pool) t.join();
    }
    for (uint32_t lp = 0; lp < np; lp++)
        if (Q[lp].st) return set_error(Q[lp].st, Q[lp].msg.c_str());
    // class offsets of every picture's arrays
    struct Base { size_t tb, coeff, ctb, cell, unit, utb, lev, smap, rmap, plane; };
    std::vector<Base> B(np + 1);
    Base run{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t lp = 0; lp < np; lp++) {
        const hpvcd_picture& P = b.pictures[C.batch_pics[lp]];
        const DevPic& D = Q[lp].D;
        run.rmap = (run.rmap + 7) & ~(size_t)7;
        B[lp] = run;
        run.tb += Q[lp].tbs.size();
        run.coeff += P.coeff_count;
        run.ctb += (size_t)D.ctb_cols * D.ctb_rows;
        run.cell += (size_t)D.gw * D.gh;
        run.unit += Q[lp].units.size();
        run.utb += Q[lp].unit_tbs.size();
        run.lev += Q[lp].level_off.size();
        run.smap += Q[lp].smap_words;
        run.rmap += Q[lp].rmap_words;
        run.plane += (size_t)D.w * D.h + 2 * (size_t)D.cw * D.ch;
    }
    B[np] = run;
    C.pics.resize(np);
    C.tbs.resize(run.tb);
    C.tb_pic.resize(run.tb);
    C.coeffs.resize(run.coeff);
    C.sao.resize(run.ctb);
    C.cells.resize(run.cell);
    C.order.resize(run.cell);
    C.units.resize(run.unit);
    C.unit_tbs.resize(run.utb);
    C.unit_recs.resize(run.utb);
    C.unit_refoff.resize(run.utb);
    C.level_off.resize(run.lev);
    {
        std::atomic<uint32_t> next{0};
        auto work = [&]() {
            for (uint32_t lp; (lp = next.fetch_add(1)) < np;) {
                const hpvcd_picture& P = b.pictures[C.batch_pics[lp]];
                PicPrep& R = Q[lp];
                const Base& o = B[lp];
                DevPic D = R.D;
                D.tb_begin = (uint32_t)o.tb;
                D.coeff_begin = (uint32_t)o.coeff;
                D.ctb_begin = (uint32_t)o.ctb;
                D.cell_begin = (uint32_t)o.cell;
                for (int k = 0; k < 3; k++) {
                    D.plane_off[k] += o.plane;
                    D.res_off[k] = D.plane_off[k];
                }
                C.pics[lp] = D;
                std::copy(R.tbs.begin(), R.tbs.end(), C.tbs.begin() + o.tb);
                std::fill(C.tb_pic.begin() + o.tb, C.tb_pic.begin() + o.tb + R.tbs.size(), lp);
                std::copy(b.coeffs + P.coeff_begin, b.coeffs + P.coeff_begin + P.coeff_count, C.coeffs.begin() + o.coeff);
                const size_t nctb = (size_t)D.ctb_cols * D.ctb_rows, ncell = (size_t)D.gw * D.gh;
                std::copy(b.sao + P.ctb_begin, b.sao + P.ctb_begin + nctb, C.sao.begin() + o.ctb);
                std::copy(b.cell_info + P.cell_begin, b.cell_info + P.cell_begin + ncell, C.cells.begin() + o.cell);
                std::copy(R.order.begin(), R.order.end(), C.order.begin() + o.cell);
                for (size_t u = 0; u < R.units.size(); u++) {
                    DevUnit U = R.units[u];
                    U.tb_list += (uint32_t)o.utb;
                    U.desc_base += (uint32_t)o.utb;
                    U.level_list += (uint32_t)o.lev;
                    U.sbase += (uint32_t)o.smap;
                    U.rbase += (uint32_t)o.rmap;
                    C.units[o.unit + u] = U;
                }
                for (size_t e = 0; e < R.unit_tbs.size(); e++) C.unit_tbs[o.utb + e] = R.unit_tbs[e] + (uint32_t)o.tb;
                std::copy(R.unit_recs.begin(), R.unit_recs.end(), C.unit_recs.begin() + o.utb);
                std::copy(R.unit_refoff.begin(), R.unit_refoff.end(), C.unit_refoff.begin() + o.utb);
                std::copy(R.level_off.begin(), R.level_off.end(), C.level_off.begin() + o.lev);
            }
        };
        const int nth = std::max(1, std::min<int>(threads, (int)np));
        std::vector<std::thread> pool;
        for (int t = 1; t < nth; t++) pool.emplace_back(work);
        work();
        for (auto& t : pool) t.join();
    }
    C.smap_words = run.smap;
    C.rmap_words = run.rmap;
    const size_t plane_elems = run.plane;
    const uint32_t unit_base = (uint32_t)run.unit;
    C.pic_units_begin.assign(np + 1, 0);
    for (int pass = 0; pass < 2; pass++)
        for (int s = 0; s < 4; s++) C.pic_res_begin[pass][s].assign(np + 1, 0);
    std::vector<std::vector<uint32_t>> per_pic_res[2][4];
    for (int pass = 0; pass < 2; pass++)
        for (int s = 0; s < 4; s++) per_pic_res[pass][s].resize(np);
    std::vector<std::vector<uint32_t>> per_pic_act(np);
    std::vector<UnitKey> keys;
    keys.reserve(run.unit);
    C.prep_list.clear();
    for (uint32_t lp = 0; lp < np; lp++) {
        const PicPrep& R = Q[lp];
        const Base& o = B[lp];
        const DevPic& D = C.pics[lp];
        const hpvcd_picture& P = b.pictures[C.batch_pics[lp]];
        C.pic_units_begin[lp] = (uint32_t)o.unit;
        for (uint32_t u : R.prep_units) C.prep_list.push_back((uint32_t)o.unit + u);
        for (const UnitKey& k : R.keys) keys.push_back({k.t, k.pic, k.plane, (uint32_t)o.unit + k.idx});
        for (int pass = 0; pass < 2; pass++)
            for (int s = 0; s < 4; s++) {
                auto& v = per_pic_res[pass][s][lp];
                v.resize(R.res[pass][s].size());
                for (size_t e = 0; e < v.size(); e++) v[e] = R.res[pass][s][e] + (uint32_t)o.tb;
            }
        per_pic_act[lp].resize(R.act.size());
        for (size_t e = 0; e < R.act.size(); e++) per_pic_act[lp][e] = R.act[e] + (uint32_t)o.tb;
        const size_t ny = (size_t)D.w * D.h;
        C.max_luma_v = std::max<uint32_t>(C.max_luma_v, (D.w >= 12 ? (D.w - 4) / 8 : 0) * (D.h / 4));
        C.max_luma_h = std::max<uint32_t>(C.max_luma_h, (D.h >= 12 ? (D.h - 4) / 8 : 0) * (D.w / 4));
        C.max_chroma_v = std::max<uint32_t>(C.max_chroma_v, (D.cw >= 10 ? (D.cw - 2) / 8 : 0) * (D.ch / 4));
        C.max_chroma_h = std::max<uint32_t>(C.max_chroma_h, (D.ch >= 10 ? (D.ch - 2) / 8 : 0) * (D.cw / 4));
        C.max_plane = std::max<uint32_t>(C.max_plane, (uint32_t)ny);
        C.max_copy = std::max<uint32_t>(C.max_copy, (uint32_t)(P.copy_w * P.copy_h));
        C.max_ctbs = std::max<uint32_t>(C.max_ctbs, (uint32_t)(D.ctb_cols * D.ctb_rows));
        C.max_rgb_qx = std::max<uint32_t>(C.max_rgb_qx, (uint32_t)((P.copy_w + 3) / 4));
        C.max_rgb_rows = std::max<uint32_t>(C.max_rgb_rows, (uint32_t)P.copy_h);
        C.max_yuv_rows = std::max<uint32_t>(C.max_yuv_rows, yuv_rows(D));
    }
    std::vector<PicPrep>().swap(Q);
    C.pic_units_begin[np] = unit_base;
    // K2a takes the units with the most predicted samples first, so that the
    // launch's tail holds the small ones; plain mode prepares each picture's
    // units in its own launch
    auto larger = [&](uint32_t a, uint32_t b) { return C.units[a].nsamp > C.units[b].nsamp; };
    C.plain_prep = C.prep_list;  // in picture order
    C.pic_prep_begin.assign(np + 1, 0);
    {
        size_t k = 0;
        for (uint32_t lp = 0; lp < np; lp++) {
            C.pic_prep_begin[lp] = (uint32_t)k;
            const size_t k0 = k;
            while (k < C.plain_prep.size() && C.units[C.plain_prep[k]].pic == lp) k++;
            std::stable_sort(C.plain_prep.begin() + k0, C.plain_prep.begin() + k, larger);
        }
        C.pic_prep_begin[np] = (uint32_t)k;
    }
    std::stable_sort(C.prep_list.begin(), C.prep_list.end(), larger);
    // batched tickets: topological (t = cx + 2 cy) across all pictures
    std::vector<UnitKey> sorted = keys;
    std::stable_sort(sorted.begin(), sorted.end(), [](const UnitKey& a, const UnitKey& b) {
        if (a.t != b.t) return a.t < b.t;
        if (a.pic != b.pic) return a.pic < b.pic;
        return a.plane < b.plane;
    });
    // plain tickets: per picture, same order restricted to the picture
    C.plain_tickets.assign(sorted.size(), 0);
    {
        std::vector<uint32_t> fill(np + 1, 0);
        for (auto& k : sorted) fill[k.pic + 1]++;
        for (uint32_t lp = 0; lp < np; lp++) fill[lp + 1] += fill[lp];
        for (auto& k : sorted) C.plain_tickets[fill[k.pic]++] = k.idx;
    }
    // residual lists (the K1 jobs of each picture built on the host threads)
    for (int pass = 0; pass < 2; pass++)
        for (int s = 0; s < 4; s++) {
            C.res_list[pass][s].clear();
            for (uint32_t lp = 0; lp < np; lp++) {
                C.pic_res_begin[pass][s][lp] = (uint32_t)C.res_list[pass][s].size();
                auto& v = per_pic_res[pass][s][lp];
                C.res_list[pass][s].insert(C.res_list[pass][s].end(), v.begin(), v.end());
            }
            C.pic_res_begin[pass][s][np] = (uint32_t)C.res_list[pass][s].size();
            C.res_jobs[pass][s].resize(C.res_list[pass][s].size());
        }
    {
        std::atomic<uint32_t> next{0};
        auto work = [&]() {
            for (uint32_t lp; (lp = next.fetch_add(1)) < np;)
                for (int pass = 0; pass < 2; pass++)
                    for (int s = 0; s < 4; s++)
                        for (uint32_t e = C.pic_res_begin[pass][s][lp]; e < C.pic_res_begin[pass][s][lp + 1]; e++)
                            C.res_jobs[pass][s][e] =
                                make_res_job(b, C.pics[lp], C.tbs[C.res_list[pass][s][e]], C.pics[lp].coeff_begin);
        };
        const int nth = std::max(1, std::min<int>(threads, (int)np));
        std::vector<std::thread> pool;
        for (int t = 1; t < nth; t++) pool.emplace_back(work);
        work();
        for (auto& t : pool) t.join();
    }
    C.act_list.clear();
    C.pic_act_begin.assign(np + 1, 0);
    for (uint32_t lp = 0; lp < np; lp++) {
        C.pic_act_begin[lp] = (uint32_t)C.act_list.size();
        C.act_list.insert(C.act_list.end(), per_pic_act[lp].begin(), per_pic_act[lp].end());
    }
    C.pic_act_begin[np] = (uint32_t)C.act_list.size();
    // prediction units and MC blocks (16 x 16 luma at most)
    C.pus.clear();
    C.mc_tasks.clear();
    C.pic_mc_begin.assign(np + 1, 0);
    C.pic_is_ref.assign(np, 0);
    for (uint32_t lp = 0; lp < np; lp++) {
        const uint32_t bp = C.batch_pics[lp];
        const hpvcd_picture& P = b.pictures[bp];
        C.pic_is_ref[lp] = is_ref[bp];
        C.pic_mc_begin[lp] = (uint32_t)C.mc_tasks.size();
        for (uint32_t i = 0; i < P.pu_count; i++) {
            const hpvcd_pu& S = b.pus[P.pu_begin + i];
            DevPu U;
            std::memset(&U, 0, sizeof(U));
            U.x = S.x; U.y = S.y; U.w = S.w; U.h = S.h;
            const hpvcd_ref* e[2] = {nullptr, nullptr};
            for (int l = 0; l < 2; l++) {
                if (!((S.flags >> l) & 1)) continue;
                e[l] = &b.refs[P.ref_begin + S.ref[l]];
                if (e[l]->picture == HPVCD_REF_MISSING) continue;  // that list reads nothing
                U.flags |= 1u << l;
                U.ref[l] = pic_local[e[l]->picture];
                U.mv[l][0] = S.mv[l][0];
                U.mv[l][1] = S.mv[l][1];
                for (int c = 0; c < 3; c++) { U.weight[l][c] = e[l]->weight[c]; U.offset[l][c] = e[l]->offset[c]; }
            }
            const hpvcd_ref* w = (U.flags & 1) ? e[0] : ((U.flags & 2) ? e[1] : nullptr);
            if (w) {
                U.weighted = w->weighted;
                U.log2_denom[0] = w->log2_denom[0];
                U.log2_denom[1] = w->log2_denom[1];
            }
            const uint32_t ui = (uint32_t)C.pus.size();
            C.pus.push_back(U);
            for (int oy = 0; oy < S.h; oy += kMcBlock)
                for (int ox = 0; ox < S.w; ox += kMcBlock) C.mc_tasks.push_back({ui, (uint16_t)lp, (uint8_t)ox, (uint8_t)oy});
        }
    }
    C.pic_mc_begin[np] = (uint32_t)C.mc_tasks.size();
    if (np > 0xffffu && !C.mc_tasks.empty()) return set_error(HPVCD_E_LIMIT, "too many pictures in one format class");
    // dependency waves (pictures arrive sorted by depth) and their tickets
    C.wave_begin.clear();
    std::vector<uint32_t> wave_of(np, 0);
    for (uint32_t lp = 0; lp < np; lp++) {
        if (lp == 0 || depth[C.batch_pics[lp]] != depth[C.batch_pics[lp - 1]]) C.wave_begin.push_back(lp);
        wave_of[lp] = (uint32_t)C.wave_begin.size() - 1;
    }
    C.wave_begin.push_back(np);
    const uint32_t nwaves = (uint32_t)C.wave_begin.size() - 1;
    C.tickets.clear();
    C.wave_tick_begin.assign(nwaves + 1, 0);
    for (uint32_t w = 0; w < nwaves; w++) {
        C.wave_tick_begin[w] = (uint32_t)C.tickets.size();
        for (auto& k : sorted)
            if (wave_of[k.pic] == w) C.tickets.push_back(k.idx);
    }
    C.wave_tick_begin[nwaves] = (uint32_t)C.tickets.size();
    C.plane_elems = plane_elems;
    C.res_elems = plane_elems;
    // [0] error, [1] border tag, then one ticket counter per wave and per picture
    C.sync_words = 2 + nwaves + np;
    return HPVCD_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------
namespace {

struct Timer {
    hipEvent_t ev[2 * HPVCD_KERNEL_COUNT * 4];
    int used = 0;
    bool on = false;
    int kinds[2 * HPVCD_KERNEL_COUNT * 4];
};

inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

template <typename R, int LOG2N>
void launch_residual_n(hipStream_t s, ClassJob& C, const uint8_t* sc, int block, const DevResJob* jobs, uint32_t count) {
    R* res = reinterpret_cast<R*>(C.d_res);
    // block / n TBs per workgroup (n lanes each)
    const uint32_t grid = cdiv(count, (uint32_t)block >> LOG2N);
    switch (block) {
        case 64: k_residual<R, LOG2N, 64><<<grid, 64, 0, s>>>(jobs, count, C.d_coeffs, sc, res); break;
        case 128: k_residual<R, LOG2N, 128><<<grid, 128, 0, s>>>(jobs, count, C.d_coeffs, sc, res); break;
        default: k_residual<R, LOG2N, 256><<<grid, 256, 0, s>>>(jobs, count, C.d_coeffs, sc, res); break;
    }
}

template <typename R>
void launch_residual(hipStream_t s, ClassJob& C, const uint8_t* sc, int log2, int block, const DevResJob* jobs,
                     uint32_t count) {
    if (!count) return;
    switch (log2) {
        case 2: launch_residual_n<R, 2>(s, C, sc, block, jobs, count); break;
        case 3: launch_residual_n<R, 3>(s, C, sc, block, jobs, count); break;
        case 4: launch_residual_n<R, 4>(s, C, sc, block, jobs, count); break;
        default: launch_residual_n<R, 5>(s, C, sc, block, jobs, count); break;
    }
}

template <typename T, typename R>
void launch_class(hpvcd_gpu* g, ClassJob& C, const uint8_t* d_scaling, uint8_t* d_out, int mode, int stop,
                  std::vector<std::pair<int, hipEvent_t>>* marks) {
    hipStream_t s = g->stream;
    const uint32_t np = (uint32_t)C.pics.size();
    const uint32_t nwaves = (uint32_t)C.wave_begin.size() - 1;
    T* pa = reinterpret_cast<T*>(C.d_planes_a);
    T* pb = reinterpret_cast<T*>(C.d_planes_b);
    const R* res = reinterpret_cast<const R*>(C.d_res);
    auto mark = [&](int kind) {
        if (!marks) return;
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, s);
        marks->push_back({kind, e});
    };
    const hpvcd_tuning& tn = g->tuning;
    int occ = g->occ_intra[(sizeof(T) == 2 ? 2 : 0) + (sizeof(R) == 4 ? 1 : 0)];
    if (tn.intra_per_cu > 0 && tn.intra_per_cu < occ) occ = tn.intra_per_cu;
    if (occ < 1) occ = 1;
    const uint32_t max_blocks = (uint32_t)(g->num_cus * occ);
    const uint32_t ncounters = nwaves + np;  // one ticket counter per wave, then one per picture
    k_begin_run<<<cdiv(ncounters, 256), 256, 0, s>>>(C.d_sync, ncounters);
    uint32_t* err = C.d_sync;
    const uint32_t* gen = C.d_sync + 1;
    uint32_t* counters = C.d_sync + 2;
    const uint32_t nunits = (uint32_t)C.units.size();
    const bool plain = mode == HPVCD_MODE_PLAIN;
    // batched: one step per dependency wave (all-intra batches have one);
    // plain: one step per picture, references first
    const int nsteps = plain ? (int)np : (int)nwaves;
    for (int step = 0; step < nsteps; step++) {
        const uint32_t p0 = plain ? (uint32_t)step : C.wave_begin[step];
        const uint32_t p1 = plain ? (uint32_t)step + 1 : C.wave_begin[step + 1];
        const uint32_t pn = p1 - p0;
        // pictures that later pictures predict from (a prefix of the step)
        // complete the loop filters whatever the stop stage: MC reads final
        // samples; the others stop at the stop stage
        uint32_t pref = p0;
        while (pref < p1 && C.pic_is_ref[pref]) pref++;
        // reference availability, reference maps, per-sample reference
        // words and descriptors of the TBs (records only) on the side stream,
        // concurrently with the residual stage: every picture at once in the
        // first step (batched), the step's picture (plain)
        const uint32_t* prep_units = plain ? C.d_plain_prep + C.pic_prep_begin[p0] : C.d_prep_list;
        const uint32_t nprep = plain ? C.pic_prep_begin[p0 + 1] - C.pic_prep_begin[p0]
                                     : (step == 0 ? (uint32_t)C.prep_list.size() : 0u);
        const bool prep = nprep > 0;
        if (prep) {
            HIP_TRY_V(hipEventRecord(g->fork, s));
            HIP_TRY_V(hipStreamWaitEvent(g->side, g->fork, 0));
            k_prep<<<nprep, kPrepThreads, 0, g->side>>>(C.d_pics, C.d_units, prep_units, C.d_unit_recs,
                                                        C.d_unit_refoff, C.d_order, C.d_cells, C.d_desc, C.d_rmap,
                                                        C.d_smap);
            HIP_TRY_V(hipEventRecord(g->join, g->side));
        }
        if (plain || step == 0) {
            // residuals depend on the records only: the step's picture in
            // plain mode, every picture of the class at once otherwise
            mark(HPVCD_KERNEL_RESIDUAL);
            for (int pass = 0; pass < 2; pass++)
                for (int sz = 0; sz < 4; sz++) {
                    uint32_t b0 = plain ? C.pic_res_begin[pass][sz][p0] : 0;
                    uint32_t b1 = plain ? C.pic_res_begin[pass][sz][p0 + 1] : (uint32_t)C.res_list[pass][sz].size();
                    launch_residual<R>(s, C, d_scaling, sz + 2, tn.residual_block[sz], C.d_res_jobs[pass][sz] + b0,
                                       b1 - b0);
                }
            uint32_t a0 = plain ? C.pic_act_begin[p0] : 0;
            uint32_t a1 = plain ? C.pic_act_begin[p0 + 1] : (uint32_t)C.act_list.size();
            if (a1 > a0)
                k_act<R><<<cdiv(a1 - a0, 4), 256, 0, s>>>(C.d_pics, C.d_tbs, C.d_tb_pic, C.d_act_list + a0, a1 - a0,
                                                          reinterpret_cast<R*>(C.d_res));
        }
        if (prep) HIP_TRY_V(hipStreamWaitEvent(s, g->join, 0));
        {
            // motion compensation of the step's pictures from the final
            // planes of earlier steps
            const uint32_t t0 = C.pic_mc_begin[p0], t1 = C.pic_mc_begin[p1];
            if (t1 > t0) {
                mark(HPVCD_KERNEL_MC);
                k_mc<T, R><<<cdiv(t1 - t0, kMcWaves), 64 * kMcWaves, 0, s>>>(C.d_pics, C.d_pus, C.d_mc_tasks + t0,
                                                                            t1 - t0, pa, pb, res);
            }
        }
        mark(HPVCD_KERNEL_INTRA);
        {
            const uint32_t* tick = plain ? C.d_plain_tickets + C.pic_units_begin[p0] : C.d_tickets + C.wave_tick_begin[step];
            uint32_t nu = plain ? C.pic_units_begin[p0 + 1] - C.pic_units_begin[p0]
                                : C.wave_tick_begin[step + 1] - C.wave_tick_begin[step];
            uint32_t* ctr = counters + (plain ? nwaves + p0 : (uint32_t)step);
            uint32_t blocks = std::min<uint32_t>(nu, max_blocks);
            if (blocks)
                k_intra<T, R><<<blocks, kIntraThreads, 0, s>>>(C.d_pics, C.d_units, tick, nu, C.d_desc, C.d_rmap, C.d_smap, pa,
                                                     res, C.d_border, ctr, gen, err);
        }
        mark(-1);
        (void)nunits;
        // the loop filters over pictures [q0, q1) of the step
        auto filters = [&](uint32_t q0, uint32_t q1, int upto) {
            const uint32_t qn = q1 - q0;
            if (!qn || upto < HPVCD_STAGE_DEBLOCK) return;
            mark(HPVCD_KERNEL_DEBLOCK);
            const DevPic& D0 = C.pics[q0];
            uint32_t lv = plain ? (uint32_t)((D0.w >= 12 ? (D0.w - 4) / 8 : 0) * (D0.h / 4)) : C.max_luma_v;
            uint32_t lh = plain ? (uint32_t)((D0.h >= 12 ? (D0.h - 4) / 8 : 0) * (D0.w / 4)) : C.max_luma_h;
            uint32_t cv = plain ? (uint32_t)((D0.cw >= 10 ? (D0.cw - 2) / 8 : 0) * (D0.ch / 4)) : C.max_chroma_v;
            uint32_t ch = plain ? (uint32_t)((D0.ch >= 10 ? (D0.ch - 2) / 8 : 0) * (D0.cw / 4)) : C.max_chroma_h;
            const uint32_t db = tn.deblock_block;
            if (lv + cv)
                k_deblock<T><<<dim3(cdiv(lv, db) + cdiv(cv, db), qn), db, 0, s>>>(C.d_pics, C.d_cells, pa, 1, q0,
                                                                              cdiv(lv, db));
            if (lh + ch)
                k_deblock<T><<<dim3(cdiv(lh, db) + cdiv(ch, db), qn), db, 0, s>>>(C.d_pics, C.d_cells, pa, 0, q0,
                                                                              cdiv(lh, db));
            mark(-1);
            if (upto < HPVCD_STAGE_SAO) return;
            mark(HPVCD_KERNEL_SAO);
            uint32_t nctb = plain ? (uint32_t)(D0.ctb_cols * D0.ctb_rows) : C.max_ctbs;
            // a launch of few CTBs splits each luma CTB's rows over several workgroups
            const int parts = nctb * qn < 4u * (uint32_t)g->num_cus ? (int)tn.sao_parts : 1;
            k_sao<T><<<dim3(nctb, qn, parts + 2), 256, 0, s>>>(C.d_pics, C.d_cells, C.d_sao, pa, pb, q0, parts);
            mark(-1);
        };
        if (stop >= HPVCD_STAGE_SAO || pref == p0) {
            filters(p0, p1, stop);
        } else {
            filters(p0, pref, HPVCD_STAGE_SAO);
            filters(pref, p1, stop);
        }
        if (stop >= HPVCD_STAGE_RGB) {
            mark(HPVCD_KERNEL_RGB);
            // quads of 4 pixels: 64 of a row by rgb_rows rows per workgroup
            const uint32_t qx = plain ? (uint32_t)((C.pics[p0].copy_w + 3) / 4) : C.max_rgb_qx;
            const uint32_t rows = plain ? (uint32_t)C.pics[p0].copy_h : C.max_rgb_rows;
            const uint32_t rr = tn.rgb_rows;
            bool oriented = false, rgb = false, raw = false;
            for (uint32_t p = p0; p < p1; p++) {
                const bool y = HPVCD_OUT_IS_YUV(C.pics[p].out_kind);
                raw = raw || y;
                rgb = rgb || !y;
                oriented = oriented || (!y && orient_moves(C.pics[p].orient));
            }
            const dim3 grid(cdiv(qx, 64), cdiv(rows, rr), pn);
            if (rgb && qx && rows && oriented) k_rgb<T, true><<<grid, 64 * rr, 0, s>>>(C.d_pics, pb, d_out, p0);
            else if (rgb && qx && rows) k_rgb<T, false><<<grid, 64 * rr, 0, s>>>(C.d_pics, pb, d_out, p0);
            // raw planes (decode_yuv outputs): luma rows, then Cb and Cr rows
            const uint32_t yrows = plain ? yuv_rows(C.pics[p0]) : C.max_yuv_rows;
            if (raw && qx && yrows)
                k_yuv<T><<<dim3(cdiv(qx, 64), cdiv(yrows, rr), pn), 64 * rr, 0, s>>>(C.d_pics, pb, d_out, p0);
            mark(-1);
        }
    }
}

void launch_job(hpvcd_job* j, int mode, int stop, std::vector<std::pair<int, hipEvent_t>>* marks) {
    for (auto& C : j->classes) {
        if (C.samp16) {
            if (C.res32) launch_class<uint16_t, int32_t>(j->g, C, j->d_scaling, j->d_out, mode, stop, marks);
            else launch_class<uint16_t, int16_t>(j->g, C, j->d_scaling, j->d_out, mode, stop, marks);
        } else {
            if (C.res32) launch_class<uint8_t, int32_t>(j->g, C, j->d_scaling, j->d_out, mode, stop, marks);
            else launch_class<uint8_t, int16_t>(j->g, C, j->d_scaling, j->d_out, mode, stop, marks);
        }
    }
}

hpvcd_status check_err(hpvcd_job* j) {
    for (auto& C : j->classes) {
        uint32_t e = 0;
        HIP_TRY(hipMemcpy(&e, C.d_sync, sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (e) return set_error(HPVCD_E_DEVICE, "intra wavefront: a dependency wait exceeded its spin limit");
    }
    return HPVCD_OK;
}


}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {

// Page-locked blocks handed out by hpvcd_host_alloc, and the freed ones kept
// for reuse (decoded images come and go at the same sizes; hipHostMalloc of
// tens of MiB costs milliseconds): at most kPinnedKeep blocks and
// kPinnedKeepBytes, released when the process ends.
constexpr size_t kPinnedKeep = 8;
constexpr size_t kPinnedKeepBytes = (size_t)1 << 30;
std::mutex g_pinned_mu;
std::vector<std::pair<void*, size_t>> g_pinned_live, g_pinned_free;

void pinned_note(void* p, size_t n) {
    std::lock_guard<std::mutex> lock(g_pinned_mu);
    g_pinned_live.push_back({p, n});
}

// a kept block of at least n bytes and at most twice that, smallest first
void* pinned_take(size_t n) {
    std::lock_guard<std::mutex> lock(g_pinned_mu);
    size_t best = SIZE_MAX;
    for (size_t i = 0; i < g_pinned_free.size(); i++)
        if (g_pinned_free[i].second >= n && g_pinned_free[i].second <= 2 * n &&
            (best == SIZE_MAX || g_pinned_free[i].second < g_pinned_free[best].second))
            best = i;
    if (best == SIZE_MAX) return nullptr;
    auto blk = g_pinned_free[best];
    g_pinned_free.erase(g_pinned_free.begin() + best);
    g_pinned_live.push_back(blk);
    return blk.first;
}

// true when p is kept for reuse (false: the caller frees it)
bool pinned_keep(void* p) {
    std::lock_guard<std::mutex> lock(g_pinned_mu);
    for (size_t i = 0; i < g_pinned_live.size(); i++)
        if (g_pinned_live[i].first == p) {
            auto blk = g_pinned_live[i];
            g_pinned_live.erase(g_pinned_live.begin() + i);
            size_t kept = blk.second;
            for (auto& f : g_pinned_free) kept += f.second;
            if (g_pinned_free.size() < kPinnedKeep && kept <= kPinnedKeepBytes) {
                g_pinned_free.push_back(blk);
                return true;
            }
            return false;
        }
    return false;
}

}  // namespace

extern "C" {

const char* hpvcd_last_error(void) { return g_last_error.c_str(); }
const char* hpvcd_version(void) { return "hpvcd-mi355x 0.1.0 (gfx950)"; }


hpvcd_status hpvcd_validate_batch(const hpvcd_batch* b) {
    if (!b) return set_error(HPVCD_E_INVALID, "null argument");
    return validate_batch(*b);
}

size_t hpvcd_output_bytes(const hpvcd_output* o) {
    if (!o) return 0;
    if (HPVCD_OUT_IS_YUV(o->kind)) {
        const int c = HPVCD_OUT_YUV_CHROMA(o->kind);
        const size_t bps = HPVCD_OUT_YUV_WIDE(o->kind) ? 2 : 1;
        const size_t sw = c == 1 ? 2 : 1, sh = c == 1 ? 2 : 1;
        const size_t cw = c ? ((size_t)o->width + sw - 1) / sw : 0, ch = c ? ((size_t)o->height + sh - 1) / sh : 0;
        return ((size_t)o->width * o->height + 2 * cw * ch) * bps;
    }
    size_t bps = (o->kind == HPVCD_OUT_RGB16 || o->kind == HPVCD_OUT_LUMA16) ? 2 : 1;
    size_t ch = (o->kind == HPVCD_OUT_LUMA8 || o->kind == HPVCD_OUT_LUMA16) ? 1 : 3;
    return (size_t)o->width * o->height * ch * bps;
}

void hpvcd_records_view(const hpvcd_records* r, hpvcd_batch* out) {
    std::memset(out, 0, sizeof(*out));
    if (!r) return;
    out->num_pictures = (uint32_t)r->pictures.size(); out->pictures = r->pictures.data();
    out->num_tbs = (uint32_t)r->tbs.size(); out->tbs = r->tbs.data();
    out->num_coeffs = (uint32_t)r->coeffs.size(); out->coeffs = r->coeffs.data();
    out->num_ctbs = (uint32_t)r->sao.size(); out->sao = r->sao.data();
    out->num_cells = (uint32_t)r->cells.size(); out->cell_info = r->cells.data();
    out->num_scaling = (uint32_t)(r->scaling.size() / HPVCD_SCALING_SET_BYTES);
    out->scaling = r->scaling.empty() ? nullptr : r->scaling.data();
    out->num_outputs = (uint32_t)r->outputs.size(); out->outputs = r->outputs.data();
    out->num_pus = (uint32_t)r->pus.size(); out->pus = r->pus.data();
    out->num_refs = (uint32_t)r->refs.size(); out->refs = r->refs.data();
}

void hpvcd_records_free(hpvcd_records* r) { delete r; }

hpvcd_status hpvcd_records_concat(hpvcd_records* const* sets, uint32_t count, hpvcd_records** out) {
    if (!out || (count && !sets)) return set_error(HPVCD_E_INVALID, "null argument");
    hpvcd_records* r = new (std::nothrow) hpvcd_records();
    if (!r) return set_error(HPVCD_E_ALLOC, "out of memory");
    for (uint32_t i = 0; i < count; i++) {
        const hpvcd_records* s = sets[i];
        if (!s) { delete r; return set_error(HPVCD_E_INVALID, "null record set"); }
        uint32_t tb0 = (uint32_t)r->tbs.size(), co0 = (uint32_t)r->coeffs.size(), ctb0 = (uint32_t)r->sao.size();
        uint32_t cell0 = (uint32_t)r->cells.size(), out0 = (uint32_t)r->outputs.size();
        int32_t sc0 = (int32_t)(r->scaling.size() / HPVCD_SCALING_SET_BYTES);
        uint32_t pic0 = (uint32_t)r->pictures.size(), pu0 = (uint32_t)r->pus.size(), ref0 = (uint32_t)r->refs.size();
        for (hpvcd_picture p : s->pictures) {
            p.tb_begin += tb0; p.coeff_begin += co0; p.ctb_begin += ctb0; p.cell_begin += cell0;
            p.pu_begin += pu0; p.ref_begin += ref0;
            p.out_index += out0;
            if (p.scaling_index >= 0) p.scaling_index += sc0;
            r->pictures.push_back(p);
        }
        r->tbs.insert(r->tbs.end(), s->tbs.begin(), s->tbs.end());
        r->coeffs.insert(r->coeffs.end(), s->coeffs.begin(), s->coeffs.end());
        r->sao.insert(r->sao.end(), s->sao.begin(), s->sao.end());
        r->cells.insert(r->cells.end(), s->cells.begin(), s->cells.end());
        r->scaling.insert(r->scaling.end(), s->scaling.begin(), s->scaling.end());
        r->outputs.insert(r->outputs.end(), s->outputs.begin(), s->outputs.end());
        r->pus.insert(r->pus.end(), s->pus.begin(), s->pus.end());
        for (hpvcd_ref e : s->refs) {
            if (e.picture != HPVCD_REF_MISSING) e.picture += pic0;
            r->refs.push_back(e);
        }
    }
    *out = r;
    return HPVCD_OK;
}

hpvcd_status hpvcd_gpu_open(int device, hpvcd_gpu** out) {
    if (!out) return set_error(HPVCD_E_INVALID, "null argument");
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_error(HPVCD_E_DEVICE, "no such GPU");
    HIP_TRY(hipSetDevice(device));
    hpvcd_gpu* g = new (std::nothrow) hpvcd_gpu();
    if (!g) return set_error(HPVCD_E_ALLOC, "out of memory");
    g->device = device;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    g->num_cus = prop.multiProcessorCount;
    {
        // the architecture name up to its feature suffix ("gfx950:sramecc+:xnack-")
        size_t n = 0;
        while (n + 1 < sizeof(g->arch) && prop.gcnArchName[n] && prop.gcnArchName[n] != ':') {
            g->arch[n] = prop.gcnArchName[n];
            n++;
        }
        g->arch[n] = 0;
    }
    hpvcd_tuning_default(&g->tuning);
    if (const char* tf = std::getenv("HPVCD_TUNING_FILE")) {
        char key[64];
        hpvcd_tuning t;
        if (*tf && !hpvcd_gpu_tuning_key(g, key, sizeof(key)) && !hpvcd_tuning_load(tf, key, &t)) g->tuning = t;
    }
    HIP_TRY(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&g->side, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&g->join, hipEventDisableTiming));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&g->occ_intra[0], k_intra<uint8_t, int16_t>, kIntraThreads, 0));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&g->occ_intra[1], k_intra<uint8_t, int32_t>, kIntraThreads, 0));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&g->occ_intra[2], k_intra<uint16_t, int16_t>, kIntraThreads, 0));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&g->occ_intra[3], k_intra<uint16_t, int32_t>, kIntraThreads, 0));
    if (hpvcd_status st = scratch_address(device, &g->scratch)) {
        hpvcd_gpu_close(g);
        return st;
    }
    *out = g;
    return HPVCD_OK;
}

void hpvcd_gpu_close(hpvcd_gpu* g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    if (g->side) (void)hipStreamDestroy(g->side);
    if (g->fork) (void)hipEventDestroy(g->fork);
    if (g->join) (void)hipEventDestroy(g->join);
    if (g->pin) (void)hipHostFree(g->pin);
    delete g;
}

hpvcd_status hpvcd_gpu_mem_info(hpvcd_gpu* g, size_t* free_b, size_t* total_b) {
    if (!g || !free_b || !total_b) return set_error(HPVCD_E_INVALID, "null argument");
    HIP_TRY(hipSetDevice(g->device));
    HIP_TRY(hipMemGetInfo(free_b, total_b));
    return HPVCD_OK;
}

void hpvcd_tuning_default(hpvcd_tuning* t) {
    if (!t) return;
    std::memset(t, 0, sizeof(*t));
    for (int k = 0; k < 4; k++) t->residual_block[k] = 256;
    t->deblock_block = 256;
    t->rgb_rows = 4;
    t->intra_per_cu = 0;
    t->sao_parts = 4;
}

hpvcd_status hpvcd_tuning_check(const hpvcd_tuning* t) {
    if (!t) return set_error(HPVCD_E_INVALID, "null argument");
    auto in = [](unsigned v, std::initializer_list<unsigned> ok) {
        for (unsigned o : ok)
            if (v == o) return true;
        return false;
    };
    for (int k = 0; k < 4; k++)
        if (!in(t->residual_block[k], {64, 128, 256})) return set_error(HPVCD_E_INVALID, "tuning: residual_block");
    if (!in(t->deblock_block, {64, 128, 256})) return set_error(HPVCD_E_INVALID, "tuning: deblock_block");
    if (!in(t->rgb_rows, {1, 2, 4})) return set_error(HPVCD_E_INVALID, "tuning: rgb_rows");
    if (t->intra_per_cu > 8) return set_error(HPVCD_E_INVALID, "tuning: intra_per_cu");
    if (!in(t->sao_parts, {1, 2, 4})) return set_error(HPVCD_E_INVALID, "tuning: sao_parts");
    return HPVCD_OK;
}

hpvcd_status hpvcd_gpu_set_tuning(hpvcd_gpu* g, const hpvcd_tuning* t) {
    if (!g || !t) return set_error(HPVCD_E_INVALID, "null argument");
    if (hpvcd_status st = hpvcd_tuning_check(t)) return st;
    std::lock_guard<std::mutex> launch_lock(g->launch_mu);
    g->tuning = *t;
    std::memset(g->tuning._pad, 0, sizeof(g->tuning._pad));
    g->tuning_gen++;
    return HPVCD_OK;
}

hpvcd_status hpvcd_gpu_get_tuning(hpvcd_gpu* g, hpvcd_tuning* t) {
    if (!g || !t) return set_error(HPVCD_E_INVALID, "null argument");
    std::lock_guard<std::mutex> launch_lock(g->launch_mu);
    *t = g->tuning;
    return HPVCD_OK;
}

hpvcd_status hpvcd_gpu_tuning_key(hpvcd_gpu* g, char* key, size_t key_size) {
    if (!g || !key || !key_size) return set_error(HPVCD_E_INVALID, "null argument");
    const int n = std::snprintf(key, key_size, "%s:%dcu", g->arch[0] ? g->arch : "gpu", g->num_cus);
    if (n < 0 || (size_t)n >= key_size) return set_error(HPVCD_E_INVALID, "key buffer too small");
    return HPVCD_OK;
}

hpvcd_status hpvcd_tune(hpvcd_job* j, int iters, hpvcd_tuning* best, double* best_ms, double* default_ms) {
    if (!j || iters < 1) return set_error(HPVCD_E_INVALID, "bad arguments");
    hpvcd_gpu* g = j->g;
    hpvcd_tuning cur;
    if (hpvcd_status st = hpvcd_gpu_get_tuning(g, &cur)) return st;
    auto timed = [&](const hpvcd_tuning& t, double* ms) -> hpvcd_status {
        if (hpvcd_status st = hpvcd_gpu_set_tuning(g, &t)) return st;
        return hpvcd_job_time(j, HPVCD_MODE_BATCHED, 1, iters, ms, nullptr);
    };
    double cur_ms = 0.0;
    if (hpvcd_status st = timed(cur, &cur_ms)) return st;
    if (default_ms) *default_ms = cur_ms;
    // knobs in launch order: K1 per TU size, K2, K3, K4, K5
    struct Knob { int field; std::vector<unsigned> values; };
    const std::vector<Knob> knobs = {
        {0, {64, 128, 256}}, {1, {64, 128, 256}}, {2, {64, 128, 256}}, {3, {64, 128, 256}},
        {4, {0, 1, 2, 3, 4}}, {5, {64, 128, 256}}, {6, {1, 2, 4}}, {7, {1, 2, 4}}};
    auto with = [](hpvcd_tuning t, int field, unsigned v) {
        if (field < 4) t.residual_block[field] = (uint16_t)v;
        else if (field == 4) t.intra_per_cu = (uint8_t)v;
        else if (field == 5) t.deblock_block = (uint16_t)v;
        else if (field == 6) t.sao_parts = (uint8_t)v;
        else t.rgb_rows = (uint16_t)v;
        return t;
    };
    for (const Knob& k : knobs)
        for (unsigned v : k.values) {
            const hpvcd_tuning t = with(cur, k.field, v);
            if (std::memcmp(&t, &cur, sizeof(t)) == 0) continue;
            double ms = 0.0;
            if (hpvcd_status st = timed(t, &ms)) {
                (void)hpvcd_gpu_set_tuning(g, &cur);
                return st;
            }
            if (ms < cur_ms) {
                cur = t;
                cur_ms = ms;
            }
        }
    if (hpvcd_status st = hpvcd_gpu_set_tuning(g, &cur)) return st;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return HPVCD_OK;
}

hpvcd_status hpvcd_tuning_load(const char* path, const char* key, hpvcd_tuning* t) {
    if (!path || !key || !t) return set_error(HPVCD_E_INVALID, "null argument");
    FILE* f = std::fopen(path, "r");
    if (!f) return set_error(HPVCD_E_INVALID, "tuning file cannot be read");
    char line[256];
    bool found = false;
    while (std::fgets(line, sizeof(line), f)) {
        char k[96];
        unsigned v[8];
        if (std::sscanf(line, "%95s %u %u %u %u %u %u %u %u", k, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6],
                        &v[7]) != 9 || std::strcmp(k, key) != 0)
            continue;
        hpvcd_tuning c;
        hpvcd_tuning_default(&c);
        bool fits = true;
        for (int i = 0; i < 8; i++) fits = fits && v[i] <= 65535u;
        if (!fits) continue;
        for (int i = 0; i < 4; i++) c.residual_block[i] = (uint16_t)v[i];
        c.deblock_block = (uint16_t)v[4];
        c.rgb_rows = (uint16_t)v[5];
        c.intra_per_cu = (uint8_t)std::min(v[6], 255u);
        c.sao_parts = (uint8_t)std::min(v[7], 255u);
        if (hpvcd_tuning_check(&c) == HPVCD_OK) {
            *t = c;  // the last valid line of the key wins
            found = true;
        }
    }
    std::fclose(f);
    if (!found) return set_error(HPVCD_E_INVALID, "tuning file has no valid line for the key");
    return HPVCD_OK;
}

hpvcd_status hpvcd_tuning_save(const char* path, const char* key, const hpvcd_tuning* t) {
    if (!path || !key || !t) return set_error(HPVCD_E_INVALID, "null argument");
    if (hpvcd_status st = hpvcd_tuning_check(t)) return st;
    for (const char* c = key; *c; c++)
        if (*c == ' ' || *c == '\t' || *c == '\n' || *c == '\r') return set_error(HPVCD_E_INVALID, "key has blanks");
    if (!*key || std::strlen(key) > 95) return set_error(HPVCD_E_INVALID, "key length");
    // keep every other key's lines, replace this key's
    std::vector<std::string> keep;
    if (FILE* f = std::fopen(path, "r")) {
        char line[256];
        while (std::fgets(line, sizeof(line), f)) {
            char k[96];
            if (std::sscanf(line, "%95s", k) == 1 && std::strcmp(k, key) == 0) continue;
            keep.emplace_back(line);
        }
        std::fclose(f);
    }
    char mine[256];
    std::snprintf(mine, sizeof(mine), "%s %u %u %u %u %u %u %u %u\n", key, t->residual_block[0], t->residual_block[1],
                  t->residual_block[2], t->residual_block[3], t->deblock_block, t->rgb_rows, t->intra_per_cu,
                  t->sao_parts);
    keep.emplace_back(mine);
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "w");
    if (!f) return set_error(HPVCD_E_INVALID, "tuning file cannot be written");
    bool ok = true;
    for (const std::string& l : keep) {
        ok = ok && std::fputs(l.c_str(), f) >= 0;
        if (!l.empty() && l.back() != '\n') ok = ok && std::fputc('\n', f) != EOF;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return set_error(HPVCD_E_INVALID, "tuning file cannot be written");
    }
    return HPVCD_OK;
}

hpvcd_status hpvcd_gpu_sync(hpvcd_gpu* g) {
    if (!g) return set_error(HPVCD_E_INVALID, "null argument");
    HIP_TRY(hipStreamSynchronize(g->stream));
    return HPVCD_OK;
}

void hpvcd_job_free(hpvcd_job* j) {
    if (!j) return;
    (void)hipSetDevice(j->g->device);
    for (auto& C : j->classes)
        if (C.graph) (void)hipGraphExecDestroy(C.graph);
    if (j->mem.base && !j->mem_external) (void)hipFree(j->mem.base);
    delete j;
}

}  // extern "C"

namespace {

// Everything job creation does on the host before touching the GPU:
// output offsets, format classes, reference depths and the per-class
// preparation (availability stamps, wavefront units, work lists).
hpvcd_status job_prepare(const hpvcd_batch& b, hpvcd_job* j, int threads) {
    j->outputs.assign(b.outputs, b.outputs + b.num_outputs);
    size_t ob = 0;
    for (auto& o : j->outputs) {
        j->out_off.push_back(ob);
        ob += (hpvcd_output_bytes(&o) + 255) & ~(size_t)255;
    }
    j->out_bytes = ob;
    // format classes: (16-bit samples, 32-bit residuals)
    j->pic_class.resize(b.num_pictures);
    j->pic_local.resize(b.num_pictures);
    for (uint32_t p = 0; p < b.num_pictures; p++) {
        const hpvcd_picture& P = b.pictures[p];
        bool s16 = P.bit_depth_luma > 8;
        bool r32 = s16 || (P.flags & HPVCD_PF_EXTENDED_PRECISION) || P.chroma_idc == 3;
        int ci = -1;
        for (size_t c = 0; c < j->classes.size(); c++)
            if (j->classes[c].samp16 == s16 && j->classes[c].res32 == r32) ci = (int)c;
        if (ci < 0) {
            j->classes.emplace_back();
            j->classes.back().samp16 = s16;
            j->classes.back().res32 = r32;
            ci = (int)j->classes.size() - 1;
        }
        j->pic_class[p] = (uint32_t)ci;
        j->classes[ci].batch_pics.push_back(p);
    }
    // reference depth of every picture (references come earlier in the
    // batch and share its format class); each class runs its pictures in
    // order of depth, one wave per depth
    std::vector<uint32_t> depth(b.num_pictures, 0);
    std::vector<uint8_t> is_ref(b.num_pictures, 0);
    for (uint32_t p = 0; p < b.num_pictures; p++) {
        const hpvcd_picture& P = b.pictures[p];
        if (!P.pu_count) continue;
        for (uint32_t k = 0; k < P.ref_count; k++) {
            uint32_t r = b.refs[P.ref_begin + k].picture;
            if (r == HPVCD_REF_MISSING) continue;
            is_ref[r] = 1;
            depth[p] = std::max(depth[p], depth[r] + 1);
        }
    }
    // within a depth, pictures that later pictures reference come first: the
    // launch chain filters that prefix to SAO whatever the stop stage, and the
    // rest only to the stop stage
    for (auto& C : j->classes) {
        std::stable_sort(C.batch_pics.begin(), C.batch_pics.end(), [&](uint32_t a, uint32_t c) {
            if (depth[a] != depth[c]) return depth[a] < depth[c];
            return is_ref[a] > is_ref[c];
        });
        for (uint32_t i = 0; i < C.batch_pics.size(); i++) j->pic_local[C.batch_pics[i]] = i;
    }
    hpvcd_status st = HPVCD_OK;
    for (auto& C : j->classes) {
        st = prepare_class(b, C, j->out_off, depth, j->pic_local, is_ref, threads);
        if (st != HPVCD_OK) return st;
    }
    return HPVCD_OK;
}

// host worker threads for batch preparation
int host_threads() {
    static const int n = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    return n;
}

// Copies the uploaded arrays to their offsets in the staging buffer on
// `threads` host threads, in pieces of at most 1 MiB.
void stage_copy(uint8_t* stage, const std::vector<std::pair<size_t, std::pair<const void*, size_t>>>& ups,
                int threads) {
    constexpr size_t kPiece = (size_t)1 << 20;
    struct Piece { uint8_t* dst; const uint8_t* src; size_t n; };
    std::vector<Piece> pieces;
    size_t total = 0;
    for (const auto& u : ups)
        for (size_t o = 0; o < u.second.second; o += kPiece) {
            const size_t n = std::min(kPiece, u.second.second - o);
            pieces.push_back({stage + u.first + o, static_cast<const uint8_t*>(u.second.first) + o, n});
            total += n;
        }
    const int nth = (int)std::max<size_t>(1, std::min<size_t>((size_t)threads, total / (4 * kPiece)));
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t k; (k = next.fetch_add(1)) < pieces.size();) std::memcpy(pieces[k].dst, pieces[k].src, pieces[k].n);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nth; t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
}

// Job creation into a pool of the library's own (ext_base null: hipMalloc,
// pinned staging) or into an existing device region of ext_size bytes
// (staged from ordinary host memory, nothing allocated on the device).
hpvcd_status job_create_impl(hpvcd_gpu* g, const hpvcd_batch* b, hpvcd_job** out, uint8_t* ext_base,
                             size_t ext_size, bool validate) {
    if (!g || !b || !out) return set_error(HPVCD_E_INVALID, "null argument");
    *out = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    hpvcd_status st = validate ? validate_batch(*b) : HPVCD_OK;
    if (st != HPVCD_OK) return st;
    HIP_TRY(hipSetDevice(g->device));
    hpvcd_job* j = new (std::nothrow) hpvcd_job();
    if (!j) return set_error(HPVCD_E_ALLOC, "out of memory");
    j->g = g;
    st = job_prepare(*b, j, host_threads());
    if (st != HPVCD_OK) {
        delete j;
        return st;
    }
    const size_t ob = j->out_bytes;
    size_t sc_bytes = (size_t)b->num_scaling * HPVCD_SCALING_SET_BYTES;
    // Device layout: first every array uploaded from the host (one
    // contiguous region, copied once from pinned staging), then the pools
    // the kernels write (outputs, sync words, border words, prep output,
    // planes, residuals).  Two passes of the same take() sequence: the
    // first sizes the regions, the second assigns pointers.
    using Upload = std::pair<size_t, std::pair<const void*, size_t>>;
    auto lay_out = [&](DevBuf& m, bool assign, std::vector<Upload>* ups) {
        auto up = [&](auto*& dst, const auto& v) {
            using X = typename std::remove_reference<decltype(v)>::type::value_type;
            X* d = m.take<X>(std::max<size_t>(v.size(), 1));
            if (assign) dst = d;
            if (ups && !v.empty()) ups->push_back({(size_t)((char*)d - m.base), {v.data(), v.size() * sizeof(X)}});
        };
        uint8_t* dsc = m.take<uint8_t>(sc_bytes + 1);
        if (assign) j->d_scaling = dsc;
        if (ups && sc_bytes) ups->push_back({(size_t)((char*)dsc - m.base), {b->scaling, sc_bytes}});
        for (auto& C : j->classes) {
            up(C.d_pics, C.pics); up(C.d_tbs, C.tbs); up(C.d_tb_pic, C.tb_pic); up(C.d_coeffs, C.coeffs);
            up(C.d_cells, C.cells); up(C.d_order, C.order); up(C.d_sao, C.sao); up(C.d_units, C.units);
            up(C.d_tickets, C.tickets); up(C.d_plain_tickets, C.plain_tickets); up(C.d_unit_recs, C.unit_recs);
            up(C.d_level_off, C.level_off); up(C.d_prep_list, C.prep_list); up(C.d_plain_prep, C.plain_prep);
            up(C.d_unit_refoff, C.unit_refoff);
            for (int pass = 0; pass < 2; pass++)
                for (int sz = 0; sz < 4; sz++) up(C.d_res_jobs[pass][sz], C.res_jobs[pass][sz]);
            up(C.d_act_list, C.act_list); up(C.d_pus, C.pus); up(C.d_mc_tasks, C.mc_tasks);
        }
        const size_t upload_end = m.used;
        uint8_t* dout = m.take<uint8_t>(ob + 1);
        if (assign) j->d_out = dout;
        for (auto& C : j->classes) {
            size_t sb = C.samp16 ? 2 : 1, rb = C.res32 ? 4 : 2;
            uint32_t* sync = m.take<uint32_t>(C.sync_words);
            uint32_t* border = m.take<uint32_t>(C.units.size() * kBorderWords + 1);
            uint32_t* smap = m.take<uint32_t>(C.smap_words + 1);
            uint16_t* rmap = m.take<uint16_t>(C.rmap_words + 8);
            uint4* desc = m.take<uint4>(C.unit_tbs.size() + 1);
            char* pa = m.take<char>(C.plane_elems * sb);
            char* pb = m.take<char>(C.plane_elems * sb);
            char* res = m.take<char>(C.res_elems * rb);
            if (assign) {
                C.d_sync = sync; C.d_border = border; C.d_smap = smap; C.d_rmap = rmap; C.d_desc = desc;
                C.d_planes_a = pa; C.d_planes_b = pb; C.d_res = res;
            }
        }
        return upload_end;
    };
    DevBuf sizing;
    const size_t upload_bytes = lay_out(sizing, false, nullptr);
    const size_t total = sizing.used + 256;
    if (ext_base) {
        if (total > ext_size) {
            delete j;
            std::string m = "batch needs " + std::to_string(total) + " bytes of device scratch; the allocation-free "
                            "entry point holds " + std::to_string(ext_size);
            return set_error(HPVCD_E_LIMIT, m.c_str());
        }
        j->mem.base = reinterpret_cast<char*>(ext_base);
        j->mem.size = ext_size;
        j->mem_external = true;
    } else {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total > free_b) {
            delete j;
            return set_error(HPVCD_E_DEVICE, "not enough free HBM for the batch");
        }
        if (hipMalloc(&j->mem.base, total) != hipSuccess) {
            delete j;
            return set_error(HPVCD_E_DEVICE, "hipMalloc of the job pool failed");
        }
        j->mem.size = total;
    }
    std::vector<Upload> ups;
    lay_out(j->mem, true, &ups);
    if (j->mem.used > j->mem.size) { hpvcd_job_free(j); return set_error(HPVCD_E_DEVICE, "pool overflow"); }
    const auto t1 = std::chrono::steady_clock::now();
    j->prep_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (ext_base) {
        // one transfer from ordinary host staging, then the same pool clears
        // (the scratch holds the previous call's contents)
        std::vector<uint8_t> stage(upload_bytes);
        stage_copy(stage.data(), ups, host_threads());
        hipError_t e = hipSuccess;
        if (upload_bytes) e = hipMemcpyAsync(j->mem.base, stage.data(), upload_bytes, hipMemcpyHostToDevice, g->stream);
        if (e == hipSuccess) e = hipMemsetAsync(j->d_out, 0, ob + 1, g->stream);
        for (auto& C : j->classes) {
            size_t rb = C.res32 ? 4 : 2;
            if (e == hipSuccess) e = hipMemsetAsync(C.d_sync, 0, C.sync_words * sizeof(uint32_t), g->stream);
            if (e == hipSuccess)
                e = hipMemsetAsync(C.d_border, 0, (C.units.size() * kBorderWords + 1) * sizeof(uint32_t), g->stream);
            if (e == hipSuccess) e = hipMemsetAsync(C.d_res, 0, C.res_elems * rb, g->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
        if (e != hipSuccess) {
            hpvcd_job_free(j);
            return set_error(HPVCD_E_DEVICE, hipGetErrorString(e));
        }
    } else {
        // stage the uploaded arrays at their device offsets in the GPU's
        // pinned buffer, copy them in one transfer, clear the written pools
        std::lock_guard<std::mutex> lock(g->mu);
        if (g->pin_size < upload_bytes) {
            if (g->pin) (void)hipHostFree(g->pin);
            g->pin = nullptr;
            g->pin_size = 0;
            if (hipHostMalloc((void**)&g->pin, upload_bytes, hipHostMallocDefault) != hipSuccess) {
                g->pin = nullptr;
                hpvcd_job_free(j);
                return set_error(HPVCD_E_DEVICE, "hipHostMalloc of the upload staging buffer failed");
            }
            g->pin_size = upload_bytes;
        }
        stage_copy(g->pin, ups, host_threads());
        hipError_t e = hipMemcpyAsync(j->mem.base, g->pin, upload_bytes, hipMemcpyHostToDevice, g->stream);
        if (e == hipSuccess) e = hipMemsetAsync(j->d_out, 0, ob + 1, g->stream);
        for (auto& C : j->classes) {
            size_t rb = C.res32 ? 4 : 2;
            // samples of TBs without a residual are never written: they stay zero
            if (e == hipSuccess) e = hipMemsetAsync(C.d_sync, 0, C.sync_words * sizeof(uint32_t), g->stream);
            if (e == hipSuccess)
                e = hipMemsetAsync(C.d_border, 0, (C.units.size() * kBorderWords + 1) * sizeof(uint32_t), g->stream);
            if (e == hipSuccess) e = hipMemsetAsync(C.d_res, 0, C.res_elems * rb, g->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
        if (e != hipSuccess) {
            hpvcd_job_free(j);
            return set_error(HPVCD_E_DEVICE, hipGetErrorString(e));
        }
    }
    j->upload_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    *out = j;
    return HPVCD_OK;
}

}  // namespace

extern "C" {

hpvcd_status hpvcd_job_create(hpvcd_gpu* g, const hpvcd_batch* b, hpvcd_job** out) {
    return job_create_impl(g, b, out, nullptr, 0, true);
}

}  // extern "C"

namespace hpvcd {
hpvcd_status job_create_validated(hpvcd_gpu* g, const hpvcd_batch* b, hpvcd_job** out) {
    return job_create_impl(g, b, out, nullptr, 0, false);
}
}  // namespace hpvcd

extern "C" {

size_t hpvcd_job_device_bytes(const hpvcd_job* j) { return j ? j->mem.size : 0; }

hpvcd_status hpvcd_job_host_ms(const hpvcd_job* j, double* prep_ms, double* upload_ms) {
    if (!j || !prep_ms || !upload_ms) return set_error(HPVCD_E_INVALID, "null argument");
    *prep_ms = j->prep_ms;
    *upload_ms = j->upload_ms;
    return HPVCD_OK;
}

hpvcd_status hpvcd_host_alloc(size_t bytes, void** out) {
    if (!out) return set_error(HPVCD_E_INVALID, "null argument");
    *out = nullptr;
    if ((*out = pinned_take(bytes ? bytes : 1))) return HPVCD_OK;
    if (hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return set_error(HPVCD_E_ALLOC, "hipHostMalloc failed");
    }
    pinned_note(*out, bytes ? bytes : 1);
    return HPVCD_OK;
}

void hpvcd_host_free(void* p) {
    if (p && !pinned_keep(p)) (void)hipHostFree(p);
}

hpvcd_status hpvcd_job_run(hpvcd_job* j, int mode, int stop) {
    if (!j) return set_error(HPVCD_E_INVALID, "null argument");
    if (stop < HPVCD_STAGE_RECON || stop > HPVCD_STAGE_RGB) return set_error(HPVCD_E_INVALID, "stop stage");
    HIP_TRY(hipSetDevice(j->g->device));
    std::lock_guard<std::mutex> launch_lock(j->g->launch_mu);
    j->last_stop = stop;
    t_launch_err = hipSuccess;
    if (mode == HPVCD_MODE_GRAPH) {
        for (auto& C : j->classes) {
            if (!C.graph || C.graph_stop != stop || C.graph_gen != j->g->tuning_gen) {
                if (C.graph) { (void)hipGraphExecDestroy(C.graph); C.graph = nullptr; }
                hipGraph_t gr;
                HIP_TRY(hipStreamBeginCapture(j->g->stream, hipStreamCaptureModeThreadLocal));
                if (C.samp16) {
                    if (C.res32) launch_class<uint16_t, int32_t>(j->g, C, j->d_scaling, j->d_out, HPVCD_MODE_BATCHED, stop, nullptr);
                    else launch_class<uint16_t, int16_t>(j->g, C, j->d_scaling, j->d_out, HPVCD_MODE_BATCHED, stop, nullptr);
                } else {
                    if (C.res32) launch_class<uint8_t, int32_t>(j->g, C, j->d_scaling, j->d_out, HPVCD_MODE_BATCHED, stop, nullptr);
                    else launch_class<uint8_t, int16_t>(j->g, C, j->d_scaling, j->d_out, HPVCD_MODE_BATCHED, stop, nullptr);
                }
                HIP_TRY(hipStreamEndCapture(j->g->stream, &gr));
                if (t_launch_err != hipSuccess) {
                    (void)hipGraphDestroy(gr);
                    return set_error(HPVCD_E_DEVICE, hipGetErrorString(t_launch_err));
                }
                hipError_t e = hipGraphInstantiate(&C.graph, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (e != hipSuccess) return set_error(HPVCD_E_DEVICE, "hipGraphInstantiate failed");
                C.graph_stop = stop;
                C.graph_gen = j->g->tuning_gen;
            }
            HIP_TRY(hipGraphLaunch(C.graph, j->g->stream));
        }
    } else if (mode == HPVCD_MODE_BATCHED || mode == HPVCD_MODE_PLAIN) {
        launch_job(j, mode, stop, nullptr);
        HIP_TRY(t_launch_err);
        HIP_TRY(hipGetLastError());
    } else {
        return set_error(HPVCD_E_INVALID, "mode");
    }
    return HPVCD_OK;
}

hpvcd_status hpvcd_job_download(hpvcd_job* j, void** outputs) {
    if (!j || !outputs) return set_error(HPVCD_E_INVALID, "null argument");
    HIP_TRY(hipSetDevice(j->g->device));
    HIP_TRY(hipStreamSynchronize(j->g->stream));
    hpvcd_status st = check_err(j);
    if (st) return st;
    for (size_t o = 0; o < j->outputs.size(); o++) {
        if (!outputs[o]) continue;
        HIP_TRY(hipMemcpy(outputs[o], j->d_out + j->out_off[o], hpvcd_output_bytes(&j->outputs[o]), hipMemcpyDeviceToHost));
    }
    return HPVCD_OK;
}

hpvcd_status hpvcd_job_download_planes(hpvcd_job* j, uint32_t picture, uint16_t* y, uint16_t* cb, uint16_t* cr) {
    if (!j || !y) return set_error(HPVCD_E_INVALID, "null argument");
    if (picture >= j->pic_class.size()) return set_error(HPVCD_E_INVALID, "picture index");
    HIP_TRY(hipSetDevice(j->g->device));
    HIP_TRY(hipStreamSynchronize(j->g->stream));
    hpvcd_status st = check_err(j);
    if (st) return st;
    ClassJob& C = j->classes[j->pic_class[picture]];
    const DevPic& D = C.pics[j->pic_local[picture]];
    const uint32_t lp = j->pic_local[picture];
    const bool final_planes = j->last_stop >= HPVCD_STAGE_SAO || C.pic_is_ref[lp];
    const char* pool = (const char*)(final_planes ? C.d_planes_b : C.d_planes_a);
    size_t sb = C.samp16 ? 2 : 1;
    uint16_t* dst[3] = {y, cb, cr};
    size_t cnt[3] = {(size_t)D.w * D.h, (size_t)D.cw * D.ch, (size_t)D.cw * D.ch};
    std::vector<uint8_t> tmp;
    for (int c = 0; c < (D.chroma_idc ? 3 : 1); c++) {
        if (!dst[c]) continue;
        tmp.resize(cnt[c] * sb);
        HIP_TRY(hipMemcpy(tmp.data(), pool + D.plane_off[c] * sb, cnt[c] * sb, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < cnt[c]; i++)
            dst[c][i] = sb == 2 ? ((const uint16_t*)tmp.data())[i] : tmp[i];
    }
    return HPVCD_OK;
}

hpvcd_status hpvcd_job_time(hpvcd_job* j, int mode, int warmup, int iters, double* ms_per_run, double* kernel_ms) {
    if (!j || !ms_per_run || iters < 1) return set_error(HPVCD_E_INVALID, "bad arguments");
    HIP_TRY(hipSetDevice(j->g->device));
    hpvcd_status st;
    for (int i = 0; i < warmup; i++)
        if ((st = hpvcd_job_run(j, mode, HPVCD_STAGE_RGB))) return st;
    HIP_TRY(hipStreamSynchronize(j->g->stream));
    if ((st = check_err(j))) return st;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, j->g->stream));
    for (int i = 0; i < iters; i++)
        if ((st = hpvcd_job_run(j, mode, HPVCD_STAGE_RGB))) return st;
    HIP_TRY(hipEventRecord(e1, j->g->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_run = ms / iters;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if ((st = check_err(j))) return st;
    if (kernel_ms) {
        for (int k = 0; k < HPVCD_KERNEL_COUNT; k++) kernel_ms[k] = 0.0;
        // per-kernel breakdown from event marks around each stage (batched
        // or plain launch sequence; a graph replays the batched sequence)
        int kmode = mode == HPVCD_MODE_GRAPH ? HPVCD_MODE_BATCHED : mode;
        for (int i = 0; i < iters; i++) {
            std::vector<std::pair<int, hipEvent_t>> marks;
            j->last_stop = HPVCD_STAGE_RGB;
            t_launch_err = hipSuccess;
            {
                std::lock_guard<std::mutex> launch_lock(j->g->launch_mu);
                launch_job(j, kmode, HPVCD_STAGE_RGB, &marks);
            }
            HIP_TRY(t_launch_err);
            HIP_TRY(hipStreamSynchronize(j->g->stream));
            for (size_t m = 0; m + 1 < marks.size(); m++) {
                if (marks[m].first < 0) continue;
                float t = 0.f;
                (void)hipEventElapsedTime(&t, marks[m].second, marks[m + 1].second);
                kernel_ms[marks[m].first] += t / iters;
            }
            for (auto& m : marks) (void)hipEventDestroy(m.second);
        }
        if ((st = check_err(j))) return st;
    }
    return HPVCD_OK;
}

hpvcd_status hpvcd_gpu_decode_batch(hpvcd_gpu* g, const hpvcd_batch* b, void** outputs) {
    hpvcd_job* j = nullptr;
    hpvcd_status st = hpvcd_job_create(g, b, &j);
    if (st) return st;
    st = hpvcd_job_run(j, HPVCD_MODE_BATCHED, HPVCD_STAGE_RGB);
    if (!st) st = hpvcd_job_download(j, outputs);
    hpvcd_job_free(j);
    return st;
}

hpvcd_status hpvcd_gpu_decode_batch_noalloc(hpvcd_gpu* g, const hpvcd_batch* b, void** outputs) {
    if (!g || !b || !outputs) return set_error(HPVCD_E_INVALID, "null argument");
    if (!g->scratch) return set_error(HPVCD_E_DEVICE, "scratch array address unavailable");
    HIP_TRY(hipSetDevice(g->device));
    std::lock_guard<std::mutex> lock(g_scratch_mu[g->device]);
    hpvcd_job* j = nullptr;
    hpvcd_status st = job_create_impl(g, b, &j, g->scratch, kScratchBytes, true);
    if (st) return st;
    st = hpvcd_job_run(j, HPVCD_MODE_BATCHED, HPVCD_STAGE_RGB);
    if (!st) st = hpvcd_job_download(j, outputs);
    hpvcd_job_free(j);
    return st;
}

size_t hpvcd_noalloc_scratch_bytes(void) { return kScratchBytes; }

hpvcd_status hpvcd_prep_digest(const hpvcd_batch* b, int32_t threads, uint64_t* digest, double* ms) {
    if (!b || !digest) return set_error(HPVCD_E_INVALID, "null argument");
    hpvcd_status st = validate_batch(*b);
    if (st) return st;
    hpvcd_job* j = new (std::nothrow) hpvcd_job();
    if (!j) return set_error(HPVCD_E_ALLOC, "out of memory");
    const auto t0 = std::chrono::steady_clock::now();
    st = job_prepare(*b, j, threads > 0 ? threads : host_threads());
    if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    uint64_t h = 1469598103934665603ull;  // FNV-1a over every prepared array
    auto mix = [&](const void* p, size_t n) {
        const uint8_t* c = static_cast<const uint8_t*>(p);
        for (size_t i = 0; i < n; i++) h = (h ^ c[i]) * 1099511628211ull;
    };
    auto vec = [&](const auto& v) {
        const uint64_t n = v.size();
        mix(&n, sizeof(n));
        if (n) mix(v.data(), n * sizeof(v[0]));
    };
    if (!st)
        for (const ClassJob& C : j->classes) {
            vec(C.batch_pics); vec(C.pics); vec(C.tbs); vec(C.tb_pic); vec(C.coeffs); vec(C.cells); vec(C.order);
            vec(C.sao); vec(C.units); vec(C.tickets); vec(C.plain_tickets); vec(C.unit_tbs); vec(C.unit_recs);
            vec(C.prep_list); vec(C.plain_prep); vec(C.pic_prep_begin); vec(C.level_off); vec(C.unit_refoff);
            for (int pass = 0; pass < 2; pass++)
                for (int sz = 0; sz < 4; sz++) {
                    vec(C.res_list[pass][sz]);
                    vec(C.res_jobs[pass][sz]);
                    vec(C.pic_res_begin[pass][sz]);
                }
            vec(C.pic_units_begin); vec(C.act_list); vec(C.pic_act_begin); vec(C.pus); vec(C.mc_tasks);
            vec(C.pic_mc_begin); vec(C.wave_begin); vec(C.wave_tick_begin); vec(C.pic_is_ref);
            const uint64_t sc[15] = {C.smap_words, C.rmap_words, C.plane_elems, C.res_elems, C.max_luma_v,
                                     C.max_luma_h, C.max_chroma_v, C.max_chroma_h, C.max_plane, C.max_copy,
                                     C.max_ctbs, C.max_rgb_qx, C.max_rgb_rows, C.max_yuv_rows, C.sync_words};
            mix(sc, sizeof(sc));
        }
    delete j;
    *digest = h;
    return st;
}

}  // extern "C"
