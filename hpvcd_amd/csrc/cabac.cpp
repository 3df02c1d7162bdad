// cabac.cpp — context initialisation, the two arithmetic engines' decision
// bins, and residual_coding (see cabac.h for the references).
#include "cabac.h"

#include <algorithm>

namespace hpvcd {

// H.265 Table 9-52 (rangeTabLps) and Table 9-53 (transIdxLps / transIdxMps)
const uint8_t kCabacRangeLps[64][4] = {
    {128, 176, 208, 240}, {128, 167, 197, 227}, {128, 158, 187, 216}, {123, 150, 178, 205}, {116, 142, 169, 195},
    {111, 135, 160, 185}, {105, 128, 152, 175}, {100, 122, 144, 166}, {95, 116, 137, 158},  {90, 110, 130, 150},
    {85, 104, 123, 142},  {81, 99, 117, 135},   {77, 94, 111, 128},   {73, 89, 105, 122},   {69, 85, 100, 116},
    {66, 80, 95, 110},    {62, 76, 90, 104},    {59, 72, 86, 99},     {56, 69, 81, 94},     {53, 65, 77, 89},
    {51, 62, 73, 85},     {48, 59, 69, 80},     {46, 56, 66, 76},     {43, 53, 63, 72},     {41, 50, 59, 69},
    {39, 48, 56, 65},     {37, 45, 54, 62},     {35, 43, 51, 59},     {33, 41, 48, 56},     {32, 39, 46, 53},
    {30, 37, 43, 50},     {29, 35, 41, 48},     {27, 33, 39, 45},     {26, 31, 37, 43},     {24, 30, 35, 41},
    {23, 28, 33, 39},     {22, 27, 32, 37},     {21, 26, 30, 35},     {20, 24, 29, 33},     {19, 23, 27, 31},
    {18, 22, 26, 30},     {17, 21, 25, 28},     {16, 20, 23, 27},     {15, 19, 22, 25},     {14, 18, 21, 24},
    {14, 17, 20, 23},     {13, 16, 19, 22},     {12, 15, 18, 21},     {12, 14, 17, 20},     {11, 14, 16, 19},
    {11, 13, 15, 18},     {10, 12, 15, 17},     {10, 12, 14, 16},     {9, 11, 13, 15},      {9, 11, 12, 14},
    {8, 10, 12, 14},      {8, 9, 11, 13},       {7, 9, 11, 12},       {7, 9, 10, 12},       {7, 8, 10, 11},
    {6, 8, 9, 11},        {6, 7, 9, 10},        {6, 7, 8, 9},         {2, 2, 2, 2}};
const uint8_t kCabacTransLps[64] = {0,  0,  1,  2,  2,  4,  4,  5,  6,  7,  8,  9,  9,  11, 11, 12,
                                      13, 13, 15, 15, 16, 16, 18, 18, 19, 19, 21, 21, 22, 22, 23, 24,
                                      24, 25, 26, 26, 27, 27, 28, 29, 29, 30, 30, 30, 31, 32, 32, 33,
                                      33, 33, 34, 34, 35, 35, 35, 36, 36, 36, 37, 37, 37, 38, 38, 63};

uint32_t kCabacDecision[128][4];

namespace {
struct DecisionInit {
    DecisionInit() {
        for (int packed = 0; packed < 128; packed++) {
            const int state = packed & 63, mps = packed >> 6;
            const uint32_t next_mps = (uint32_t)((state < 62 ? state + 1 : state) | (mps << 6));
            const uint32_t next_lps = (uint32_t)(kCabacTransLps[state] | ((state == 0 ? mps ^ 1 : mps) << 6));
            for (int q = 0; q < 4; q++) kCabacDecision[packed][q] = kCabacRangeLps[state][q] | next_mps << 8 | next_lps << 16;
        }
    }
} decision_init;
}  // namespace

const uint8_t* cabac_range_lps_row(int s) { return kCabacRangeLps[s]; }
uint8_t cabac_trans_lps(int s) { return kCabacTransLps[s]; }
uint8_t cabac_trans_mps(int s) { return (uint8_t)(s < 62 ? s + 1 : s); }

// ---------------------------------------------------------------------------
// initValues per initType (the reference's tables, src/cabac/contexts.rs)
// ---------------------------------------------------------------------------
const CtxElement kCtxElements[] = {
    {"sao_merge_flag", CTX_SAO_MERGE, 1},
    {"sao_type_idx", CTX_SAO_TYPE, 1},
    {"split_cu_flag", CTX_SPLIT_CU, 3},
    {"cu_transquant_bypass_flag", CTX_TQB, 1},
    {"cu_skip_flag", CTX_SKIP, 3},
    {"pred_mode_flag", CTX_PRED_MODE, 1},
    {"part_mode", CTX_PART_MODE, 4},
    {"prev_intra_luma_pred_flag", CTX_PREV_INTRA, 1},
    {"intra_chroma_pred_mode", CTX_CHROMA_MODE, 1},
    {"rqt_root_cbf", CTX_RQT_ROOT_CBF, 1},
    {"merge_flag", CTX_MERGE_FLAG, 1},
    {"merge_idx", CTX_MERGE_IDX, 1},
    {"inter_pred_idc", CTX_INTER_PRED, 5},
    {"ref_idx", CTX_REF_IDX, 2},
    {"mvp_flag", CTX_MVP_FLAG, 1},
    {"abs_mvd_greater01", CTX_MVD, 2},
    {"split_transform_flag", CTX_SPLIT_TF, 3},
    {"cbf_luma", CTX_CBF_LUMA, 2},
    {"cbf_chroma", CTX_CBF_CHROMA, 5},
    {"cu_qp_delta_abs", CTX_QP_DELTA, 2},
    {"transform_skip_flag", CTX_TS, 2},
    {"last_sig_coeff_x_prefix", CTX_LAST_X, 18},
    {"last_sig_coeff_y_prefix", CTX_LAST_Y, 18},
    {"coded_sub_block_flag", CTX_CSBF, 4},
    {"sig_coeff_flag", CTX_SIG, 44},
    {"coeff_abs_level_greater1", CTX_G1, 24},
    {"coeff_abs_level_greater2", CTX_G2, 6},
    {"explicit_rdpcm_flag", CTX_RDPCM_FLAG, 2},
    {"explicit_rdpcm_dir", CTX_RDPCM_DIR, 2},
    {"log2_res_scale_abs_plus1", CTX_RES_SCALE, 8},
    {"res_scale_sign_flag", CTX_RES_SIGN, 2},
    {"cu_residual_act_flag", CTX_ACT, 1},
};
const int kNumCtxElements = (int)(sizeof(kCtxElements) / sizeof(kCtxElements[0]));

namespace {

struct CtxTable {
    uint8_t v[3][CTX_COUNT];
    void set(int off, const uint8_t* t0, const uint8_t* t1, const uint8_t* t2, int n) {
        std::memcpy(v[0] + off, t0, (size_t)n);
        std::memcpy(v[1] + off, t1, (size_t)n);
        std::memcpy(v[2] + off, t2, (size_t)n);
    }
    void all(int off, uint8_t x, int n) {
        for (int t = 0; t < 3; t++)
            for (int i = 0; i < n; i++) v[t][off + i] = x;
    }
    CtxTable() {
        std::memset(v, 154, sizeof(v));
        all(CTX_SAO_MERGE, 153, 1);
        const uint8_t sao_t[3] = {200, 185, 160};
        set(CTX_SAO_TYPE, &sao_t[0], &sao_t[1], &sao_t[2], 1);
        const uint8_t scu[3][3] = {{139, 141, 157}, {107, 139, 126}, {107, 139, 126}};
        set(CTX_SPLIT_CU, scu[0], scu[1], scu[2], 3);
        all(CTX_TQB, 154, 1);
        const uint8_t skip[3] = {197, 185, 201};
        set(CTX_SKIP, skip, skip, skip, 3);
        const uint8_t pm[3] = {149, 149, 134};
        set(CTX_PRED_MODE, &pm[0], &pm[1], &pm[2], 1);
        const uint8_t part0[4] = {184, 154, 139, 154}, part1[4] = {154, 139, 154, 154};
        set(CTX_PART_MODE, part0, part1, part1, 4);
        const uint8_t prev[3] = {184, 154, 183};
        set(CTX_PREV_INTRA, &prev[0], &prev[1], &prev[2], 1);
        const uint8_t chm[3] = {63, 152, 152};
        set(CTX_CHROMA_MODE, &chm[0], &chm[1], &chm[2], 1);
        all(CTX_RQT_ROOT_CBF, 79, 1);
        const uint8_t mf[3] = {110, 110, 154}, mi[3] = {122, 122, 137};
        set(CTX_MERGE_FLAG, &mf[0], &mf[1], &mf[2], 1);
        set(CTX_MERGE_IDX, &mi[0], &mi[1], &mi[2], 1);
        const uint8_t ipi[5] = {95, 79, 63, 31, 31};
        set(CTX_INTER_PRED, ipi, ipi, ipi, 5);
        all(CTX_REF_IDX, 153, 2);
        all(CTX_MVP_FLAG, 168, 1);
        const uint8_t mvd0[2] = {140, 198}, mvd2[2] = {169, 198};
        set(CTX_MVD, mvd0, mvd0, mvd2, 2);
        const uint8_t stf[3][3] = {{153, 138, 138}, {124, 138, 94}, {224, 167, 122}};
        set(CTX_SPLIT_TF, stf[0], stf[1], stf[2], 3);
        const uint8_t cbl[3][2] = {{111, 141}, {153, 111}, {153, 111}};
        set(CTX_CBF_LUMA, cbl[0], cbl[1], cbl[2], 2);
        const uint8_t cbc[3][5] = {{94, 138, 182, 154, 154}, {149, 107, 167, 154, 154}, {149, 92, 167, 154, 154}};
        set(CTX_CBF_CHROMA, cbc[0], cbc[1], cbc[2], 5);
        all(CTX_QP_DELTA, 154, 2);
        all(CTX_TS, 139, 2);
        static const uint8_t last[3][18] = {
            {110, 110, 124, 125, 140, 153, 125, 127, 140, 109, 111, 143, 127, 111, 79, 108, 123, 63},
            {125, 110, 94, 110, 95, 79, 125, 111, 110, 78, 110, 111, 111, 95, 94, 108, 123, 108},
            {125, 110, 124, 110, 95, 94, 125, 111, 111, 79, 125, 126, 111, 111, 79, 108, 123, 93}};
        set(CTX_LAST_X, last[0], last[1], last[2], 18);
        set(CTX_LAST_Y, last[0], last[1], last[2], 18);
        const uint8_t csbf[3][4] = {{91, 171, 134, 141}, {121, 140, 61, 154}, {121, 140, 61, 154}};
        set(CTX_CSBF, csbf[0], csbf[1], csbf[2], 4);
        static const uint8_t sig[3][44] = {
            {111, 111, 125, 110, 110, 94,  124, 108, 124, 107, 125, 141, 179, 153, 125, 107, 125, 141, 179, 153, 125, 107,
             125, 141, 179, 153, 125, 140, 139, 182, 182, 152, 136, 152, 136, 153, 136, 139, 111, 136, 139, 111, 141, 111},
            {155, 154, 139, 153, 139, 123, 123, 63,  153, 166, 183, 140, 136, 153, 154, 166, 183, 140, 136, 153, 154, 166,
             183, 140, 136, 153, 154, 170, 153, 123, 123, 107, 121, 107, 121, 167, 151, 183, 140, 151, 183, 140, 140, 140},
            {170, 154, 139, 153, 139, 123, 123, 63,  124, 166, 183, 140, 136, 153, 154, 166, 183, 140, 136, 153, 154, 166,
             183, 140, 136, 153, 154, 170, 153, 138, 138, 122, 121, 122, 121, 167, 151, 183, 140, 151, 183, 140, 140, 140}};
        set(CTX_SIG, sig[0], sig[1], sig[2], 44);
        static const uint8_t g1[3][24] = {
            {140, 92, 137, 138, 140, 152, 138, 139, 153, 74, 149, 92, 139, 107, 122, 152, 140, 179, 166, 182, 140, 227, 122,
             197},
            {154, 196, 196, 167, 154, 152, 167, 182, 182, 134, 149, 136, 153, 121, 136, 137, 169, 194, 166, 167, 154, 167,
             137, 182},
            {154, 196, 167, 167, 154, 152, 167, 182, 182, 134, 149, 136, 153, 121, 136, 122, 169, 208, 166, 167, 154, 152,
             167, 182}};
        set(CTX_G1, g1[0], g1[1], g1[2], 24);
        const uint8_t g2[3][6] = {{138, 153, 136, 167, 152, 152}, {107, 167, 91, 122, 107, 167}, {107, 167, 91, 107, 107, 167}};
        set(CTX_G2, g2[0], g2[1], g2[2], 6);
        all(CTX_RDPCM_FLAG, 139, 2);
        all(CTX_RDPCM_DIR, 139, 2);
        all(CTX_RES_SCALE, 154, 8);
        all(CTX_RES_SIGN, 154, 2);
        all(CTX_ACT, 154, 1);
    }
};

const CtxTable& table() {
    static const CtxTable t;
    return t;
}

}  // namespace

const uint8_t (&ctx_init_table())[3][CTX_COUNT] { return table().v; }

// 9.3.2.2: m = slopeIdx*5 - 45, n = (offsetIdx << 3) - 16,
// preCtxState = Clip3(1, 126, ((m * Clip3(0, 51, SliceQpY)) >> 4) + n)
void Contexts::init(int init_type, int slice_qp) {
    const uint8_t* iv = table().v[init_type];
    const int qp = std::clamp(slice_qp, 0, 51);
    for (int i = 0; i < CTX_COUNT; i++) {
        const int m = (iv[i] >> 4) * 5 - 45, n = ((iv[i] & 15) << 3) - 16;
        const int pre = std::clamp(((m * qp) >> 4) + n, 1, 126);
        s[i] = pre >= 64 ? (uint8_t)((pre - 64) | 64) : (uint8_t)(63 - pre);
    }
    std::memset(stat_coeff, 0, sizeof(stat_coeff));
}

void CabacEnc::encode(uint8_t& ctx, int bin) {
    const int state = ctx & 63, mps = ctx >> 6;
    const uint32_t lps = kCabacRangeLps[state][(range >> 6) & 3];
    range -= lps;
    if (bin != mps) {
        low += range;
        range = lps;
        ctx = (uint8_t)(kCabacTransLps[state] | ((state == 0 ? mps ^ 1 : mps) << 6));
    } else {
        ctx = (uint8_t)(cabac_trans_mps(state) | (mps << 6));
    }
    renorm();
}

// ---------------------------------------------------------------------------
// scans and contexts of residual_coding
// ---------------------------------------------------------------------------
namespace {

struct Scans {
    uint8_t order[4][3][64];  // [log2w][scan] positions x + y * w
    uint8_t index[4][3][64];  // position -> scan index
    Scans() {
        for (int lg = 0; lg < 4; lg++) {
            const int w = 1 << lg;
            for (int sc = 0; sc < 3; sc++) {
                int k = 0;
                if (sc == SCAN_HOR) {
                    for (int y = 0; y < w; y++)
                        for (int x = 0; x < w; x++) order[lg][sc][k++] = (uint8_t)(x + y * w);
                } else if (sc == SCAN_VER) {
                    for (int x = 0; x < w; x++)
                        for (int y = 0; y < w; y++) order[lg][sc][k++] = (uint8_t)(x + y * w);
                } else {  // up-right diagonal (6.5.3)
                    int x = 0, y = 0;
                    while (k < w * w) {
                        while (y >= 0) {
                            if (x < w && y < w) order[lg][sc][k++] = (uint8_t)(x + y * w);
                            y--;
                            x++;
                        }
                        y = x;
                        x = 0;
                    }
                }
                for (int i = 0; i < w * w; i++) index[lg][sc][order[lg][sc][i]] = (uint8_t)i;
            }
        }
    }
};

const Scans kScans;
inline const Scans& scans() { return kScans; }

// sig_coeff_flag context (9.3.4.2.5; src/cabac/residual_full.rs:189)
int sig_ctx(int xc, int yc, int prev_csbf, int log2n, int scan, bool luma);

// sig_ctx of the 16 positions (raster in the 4x4 sub-block) of a sub-block,
// per TB size, scan, csbf pattern of the right / below sub-blocks, component
// class and whether it is the DC sub-block
struct SigTable {
    uint8_t v[4][3][4][2][2][16];
    SigTable() {
        for (int lg = 0; lg < 4; lg++)
            for (int sc = 0; sc < 3; sc++)
                for (int prev = 0; prev < 4; prev++)
                    for (int luma = 0; luma < 2; luma++)
                        for (int dc = 0; dc < 2; dc++)
                            for (int pp = 0; pp < 16; pp++) {
                                const int sbx = dc ? 0 : 1, xc = (sbx << 2) + (pp & 3), yc = pp >> 2;
                                v[lg][sc][prev][luma][dc][pp] =
                                    (uint8_t)std::min(sig_ctx(xc, yc, prev, lg + 2, sc, luma != 0), 43);
                            }
    }
};

const SigTable kSigTable;
inline const SigTable& sig_table() { return kSigTable; }

int sig_ctx(int xc, int yc, int prev_csbf, int log2n, int scan, bool luma) {
    static const uint8_t map4[16] = {0, 1, 4, 5, 2, 3, 4, 5, 6, 6, 8, 8, 7, 7, 8, 8};
    int s;
    if (log2n == 2) s = map4[(yc << 2) + xc];
    else if (xc + yc == 0) s = 0;
    else {
        const int xp = xc & 3, yp = yc & 3;
        if (prev_csbf == 0) s = xp + yp >= 3 ? 0 : (xp + yp > 0 ? 1 : 2);
        else if (prev_csbf == 1) s = yp == 0 ? 2 : (yp == 1 ? 1 : 0);
        else if (prev_csbf == 2) s = xp == 0 ? 2 : (xp == 1 ? 1 : 0);
        else s = 2;
        if (luma) {
            if ((xc >> 2) + (yc >> 2) > 0) s += 3;
            s += log2n == 3 ? (scan == SCAN_DIAG ? 9 : 15) : 21;
        } else {
            s += log2n == 3 ? 9 : 12;
        }
    }
    return luma ? s : 27 + s;
}

const uint8_t kGroupIdx[32] = {0, 1, 2, 3, 4, 4, 5, 5, 6, 6, 6, 6, 7, 7, 7, 7,
                               8, 8, 8, 8, 8, 8, 8, 8, 9, 9, 9, 9, 9, 9, 9, 9};
const uint8_t kMinGroup[10] = {0, 1, 2, 3, 4, 6, 8, 12, 16, 24};

void last_ctx(int log2n, bool luma, int* off, int* shift) {
    if (luma) {
        *off = 3 * (log2n - 2) + ((log2n - 1) >> 2);
        *shift = (log2n + 1) >> 2;
    } else {
        *off = 15;
        *shift = log2n - 2;
    }
}

// one sub-block's significant coefficients in decreasing scan order
struct SbEntries {
    int n = 0;
    uint16_t coeff[16];  // raster index in the TB
    uint8_t k[16];       // scan position inside the sub-block
};

}  // namespace

const uint8_t* scan_order(int log2w, int scan) { return scans().order[log2w][scan]; }

int scan_for(int mode, int log2n, bool luma, bool chroma444) {
    const bool dep = log2n == 2 || (log2n == 3 && (luma || chroma444));
    if (!dep) return SCAN_DIAG;
    if (mode >= 6 && mode <= 14) return SCAN_VER;
    if (mode >= 22 && mode <= 30) return SCAN_HOR;
    return SCAN_DIAG;
}

// --- binarisations of coeff_abs_level_remaining (9.3.3.11 / 9.3.3.12) ------
static inline __attribute__((always_inline)) uint32_t dec_remaining(CabacDec& d, int rice, bool ep, int bd) {
    int q = 0;
    while (q < 4 && d.bypass()) q++;
    if (q < 4) return ((uint32_t)q << rice) + d.bypass_bits(rice);
    const int k = rice + 1;
    uint32_t suffix;
    if (!ep) {
        int prefix = 0;
        while (prefix < 31 && d.bypass()) prefix++;
        if (prefix == 31 || prefix + k >= 32) return 0;
        suffix = (((1u << prefix) - 1) << k) + d.bypass_bits(prefix + k);
    } else {
        const int range = std::max(15, bd + 6);
        const int max_ext = std::max(0, 28 - range);
        int ext = 0;
        while (ext < max_ext && d.bypass()) ext++;
        const int esc = ext == max_ext ? range : ext + k;
        if (ext >= 32 || esc > 32) return 0;
        suffix = (((1u << ext) - 1) << k) + d.bypass_bits(esc);
    }
    return (4u << rice) + suffix;
}

static void enc_remaining(CabacEnc& e, uint32_t v, int rice, bool ep, int bd) {
    const uint32_t q = v >> rice;
    if (q < 4) {
        for (uint32_t i = 0; i < q; i++) e.bypass(1);
        e.bypass(0);
        e.bypass_bits(v & ((1u << rice) - 1), rice);
        return;
    }
    for (int i = 0; i < 4; i++) e.bypass(1);
    uint32_t s = v - (4u << rice);
    const int k = rice + 1;
    if (!ep) {
        int prefix = 0;
        while (s >= (((2u << prefix) - 1) << k)) prefix++;
        for (int i = 0; i < prefix; i++) e.bypass(1);
        e.bypass(0);
        e.bypass_bits(s - (((1u << prefix) - 1) << k), prefix + k);
    } else {
        const int range = std::max(15, bd + 6);
        const int max_ext = std::max(0, 28 - range);
        int ext = 0;
        while (ext < max_ext && s >= (((2u << ext) - 1) << k)) ext++;
        for (int i = 0; i < ext; i++) e.bypass(1);
        if (ext < max_ext) {
            e.bypass(0);
            e.bypass_bits(s - (((1u << ext) - 1) << k), ext + k);
        } else {
            e.bypass_bits(s - (((1u << ext) - 1) << k), range);
        }
    }
}

// --- residual_coding --------------------------------------------------------
// The decoder follows src/cabac/residual_full.rs:392; the encoder writes the
// same bins in the same order.
namespace {

struct RcState {
    int n, sbw, log2sb;
    const uint8_t *sb_order, *sb_index, *pos_order, *pos_index;
};

RcState rc_state(const ResidualParams& p) {
    RcState s;
    s.n = 1 << p.log2n;
    s.log2sb = p.log2n - 2;
    s.sbw = 1 << s.log2sb;
    s.sb_order = scans().order[s.log2sb][p.scan];
    s.sb_index = scans().index[s.log2sb][p.scan];
    s.pos_order = scans().order[2][p.scan];
    s.pos_index = scans().index[2][p.scan];
    return s;
}

bool rdpcm_active(const ResidualParams& p, const ResidualOut& r) {
    return r.rdpcm >= 0 || (r.ts && p.implicit_rdpcm);
}

}  // namespace

static void residual_decode_local(CabacDec& d, Contexts& c, const ResidualParams& p, int32_t* levels, ResidualOut& r);

// The engine runs on a local copy: the context stores (bytes, which may alias
// anything) then cannot force its state out of registers after every bin.
void residual_decode(CabacDec& dref, Contexts& c, const ResidualParams& p, int32_t* levels, ResidualOut& r) {
    CabacDec d = dref;
    residual_decode_local(d, c, p, levels, r);
    dref = d;
}

static inline __attribute__((always_inline)) void residual_decode_local(CabacDec& d, Contexts& c,
                                                                        const ResidualParams& p, int32_t* levels,
                                                                        ResidualOut& r) {
    const RcState S = rc_state(p);
    const int n = S.n;
    if (r.nz) r.nnz = 0;
    else std::fill(levels, levels + n * n, 0);
    r.ts = false;
    r.rdpcm = -1;
    const int ci = p.luma ? 0 : 1;
    if (p.ts_allowed && !p.tqb) r.ts = d.decode(c.s[CTX_TS + ci]) != 0;
    if (p.explicit_rdpcm && (r.ts || p.tqb)) {
        if (d.decode(c.s[CTX_RDPCM_FLAG + ci])) r.rdpcm = d.decode(c.s[CTX_RDPCM_DIR + ci]);
    }
    const int ts_ctx = (p.ts_context && (r.ts || p.tqb)) ? (p.luma ? 42 : 43) : -1;
    int off, shift;
    last_ctx(p.log2n, p.luma, &off, &shift);
    const int max_group = kGroupIdx[n - 1];
    int gx = 0, gy = 0;
    while (gx < max_group && d.decode(c.s[CTX_LAST_X + off + (gx >> shift)])) gx++;
    while (gy < max_group && d.decode(c.s[CTX_LAST_Y + off + (gy >> shift)])) gy++;
    int lx = kMinGroup[gx], ly = kMinGroup[gy];
    if (gx > 3) lx += (int)d.bypass_bits((gx >> 1) - 1);
    if (gy > 3) ly += (int)d.bypass_bits((gy >> 1) - 1);
    if (p.scan == SCAN_VER) std::swap(lx, ly);
    if (lx >= n || ly >= n) {  // a malformed position: decode as DC
        lx = ly = 0;
        d.overrun = true;
    }
    const int last_sb = S.sb_index[(lx >> 2) + (ly >> 2) * S.sbw];
    const int last_in_sb = S.pos_index[(lx & 3) + (ly & 3) * 4];
    uint64_t csbf = 0;
    int c1_carry = 1;
    bool first_sb = true;
    for (int i = last_sb; i >= 0; i--) {
        const int sbp = S.sb_order[i], sbx = sbp % S.sbw, sby = sbp / S.sbw;
        const int right = sbx + 1 < S.sbw ? (int)((csbf >> (sbp + 1)) & 1) : 0;
        const int below = sby + 1 < S.sbw ? (int)((csbf >> (sbp + S.sbw)) & 1) : 0;
        const int prev = right | (below << 1);
        const bool infer_dc = i < last_sb && i > 0;
        const uint8_t* sig_row = sig_table().v[p.log2n - 2][p.scan][prev][p.luma ? 1 : 0][sbp == 0 ? 1 : 0];
        bool coded;
        if (infer_dc) coded = d.decode(c.s[CTX_CSBF + (p.luma ? 0 : 2) + ((right | below) & 1)]) != 0;
        else coded = true;
        if (coded) csbf |= 1ull << sbp;
        if (!coded) continue;
        SbEntries E;
        if (i == last_sb) {
            E.coeff[0] = (uint16_t)(ly * n + lx);
            E.k[0] = (uint8_t)last_in_sb;
            E.n = 1;
        }
        const int top = i == last_sb ? last_in_sb - 1 : 15;
        // every position is written and counted only when significant (no
        // branch on the decoded flag)
        const int sb_base = (sby << 2) * n + (sbx << 2);
        for (int k = top; k > 0; k--) {
            const int pp = S.pos_order[k];
            const int cx = ts_ctx >= 0 ? ts_ctx : sig_row[pp];
            const int sig = d.decode(c.s[CTX_SIG + cx]);
            E.coeff[E.n] = (uint16_t)(sb_base + (pp >> 2) * n + (pp & 3));
            E.k[E.n] = (uint8_t)k;
            E.n += sig;
        }
        if (top >= 0) {  // position 0 of the sub-block, inferred when nothing else is
            const int pp = S.pos_order[0];
            int sig;
            if (infer_dc && E.n == 0) sig = 1;
            else sig = d.decode(c.s[CTX_SIG + (ts_ctx >= 0 ? ts_ctx : sig_row[pp])]);
            E.coeff[E.n] = (uint16_t)(sb_base + (pp >> 2) * n + (pp & 3));
            E.k[E.n] = 0;
            E.n += sig;
        }
        if (!E.n) continue;
        int ctx_set = (i == 0 || !p.luma) ? 0 : 2;
        if (!first_sb && c1_carry == 0) ctx_set++;
        first_sb = false;
        int c1 = 1;
        uint32_t g1mask = 0;
        const int ng1 = std::min(E.n, 8);
        const int g1_base = CTX_G1 + ctx_set * 4 + (p.luma ? 0 : 16);
        for (int j = 0; j < ng1; j++) {
            const int g = d.decode(c.s[g1_base + std::min(c1, 3)]);
            g1mask |= (uint32_t)g << j;
            // greater1Ctx: 0 after a level above one (it stays 0), else one
            // more up to 3
            c1 = g ? 0 : c1 + (c1 > 0 && c1 < 3);
        }
        const int first_g1 = g1mask ? __builtin_ctz(g1mask) : 16;
        c1_carry = c1;
        const bool g2 = first_g1 < 16 ? d.decode(c.s[CTX_G2 + ctx_set + (p.luma ? 0 : 4)]) != 0 : false;
        const bool escape = E.n > 8 || __builtin_popcount(g1mask) > 1 || g2;
        if (p.bypass_alignment && escape) d.align_bypass();
        const bool hidden = p.sign_hiding && (E.k[0] - E.k[E.n - 1]) > 3 && !p.tqb && !rdpcm_active(p, r);
        const int nsign = E.n - (hidden ? 1 : 0);
        const uint32_t signs = d.bypass_bits(nsign);
        const int sb_type = 2 * (p.luma ? 1 : 0) + ((r.ts || p.tqb) ? 1 : 0);
        int rice = p.persistent_rice ? (c.stat_coeff[sb_type] >> 2) : 0;
        bool first_rem = true;
        int parity = 0;
        for (int j = 0; j < E.n; j++) {
            const int g1 = (g1mask >> j) & 1;
            const bool has_g2 = j == first_g1;
            const int base = 1 + g1 + ((has_g2 && g2) ? 1 : 0);
            const int thr = j < 8 ? (has_g2 ? 3 : 2) : 1;
            uint32_t rem = 0;
            if (base == thr) {
                rem = dec_remaining(d, rice, p.extended_precision, p.bit_depth);
                if (p.persistent_rice && first_rem) {
                    first_rem = false;
                    uint8_t& st = c.stat_coeff[sb_type];
                    if (rem >= (3u << (st >> 2))) st++;
                    else if (2 * rem < (1u << (st >> 2)) && st > 0) st--;
                }
                if ((int64_t)base + rem > (3ll << rice)) rice = p.persistent_rice ? rice + 1 : std::min(rice + 1, 4);
            }
            const int level = (int)std::min<int64_t>((int64_t)base + rem, 0x7fffffff);
            const bool neg = j < nsign && ((signs >> (nsign - 1 - j)) & 1);
            levels[E.coeff[j]] = neg ? -level : level;
            if (r.nz) r.nz[r.nnz++] = E.coeff[j];
            parity ^= level & 1;
        }
        if (hidden) {
            int32_t& v = levels[E.coeff[E.n - 1]];
            v = parity ? -std::abs(v) : std::abs(v);
        }
    }
}

void residual_hide_signs(const ResidualParams& p, const ResidualOut& r, int32_t* levels) {
    if (!p.sign_hiding || p.tqb || rdpcm_active(p, r)) return;
    const RcState S = rc_state(p);
    for (int i = 0; i < S.sbw * S.sbw; i++) {
        const int sbp = S.sb_order[i], sbx = sbp % S.sbw, sby = sbp / S.sbw;
        int first = -1, last = -1, parity = 0;
        for (int k = 0; k < 16; k++) {
            const int pp = S.pos_order[k];
            const int32_t v = levels[((sby << 2) + (pp >> 2)) * S.n + (sbx << 2) + (pp & 3)];
            if (!v) continue;
            if (first < 0) first = k;
            last = k;
            parity ^= std::abs(v) & 1;
        }
        if (first < 0 || last - first <= 3) continue;
        const int pp = S.pos_order[first];
        int32_t& v = levels[((sby << 2) + (pp >> 2)) * S.n + (sbx << 2) + (pp & 3)];
        v = parity ? -std::abs(v) : std::abs(v);
    }
}

void residual_encode(CabacEnc& e, Contexts& c, const ResidualParams& p, const int32_t* levels, const ResidualOut& r) {
    const RcState S = rc_state(p);
    const int n = S.n;
    const int ci = p.luma ? 0 : 1;
    if (p.ts_allowed && !p.tqb) e.encode(c.s[CTX_TS + ci], r.ts ? 1 : 0);
    if (p.explicit_rdpcm && (r.ts || p.tqb)) {
        e.encode(c.s[CTX_RDPCM_FLAG + ci], r.rdpcm >= 0 ? 1 : 0);
        if (r.rdpcm >= 0) e.encode(c.s[CTX_RDPCM_DIR + ci], r.rdpcm);
    }
    const int ts_ctx = (p.ts_context && (r.ts || p.tqb)) ? (p.luma ? 42 : 43) : -1;
    // last significant coefficient in scan order
    int last_sb = -1, last_in_sb = -1;
    for (int i = S.sbw * S.sbw - 1; i >= 0 && last_sb < 0; i--) {
        const int sbp = S.sb_order[i], sbx = sbp % S.sbw, sby = sbp / S.sbw;
        for (int k = 15; k >= 0; k--) {
            const int pp = S.pos_order[k];
            if (levels[((sby << 2) + (pp >> 2)) * n + (sbx << 2) + (pp & 3)]) {
                last_sb = i;
                last_in_sb = k;
                break;
            }
        }
    }
    const int sbp_l = S.sb_order[last_sb], pp_l = S.pos_order[last_in_sb];
    int lx = ((sbp_l % S.sbw) << 2) + (pp_l & 3), ly = ((sbp_l / S.sbw) << 2) + (pp_l >> 2);
    if (p.scan == SCAN_VER) std::swap(lx, ly);
    int off, shift;
    last_ctx(p.log2n, p.luma, &off, &shift);
    const int max_group = kGroupIdx[n - 1];
    const int gx = kGroupIdx[lx], gy = kGroupIdx[ly];
    for (int g = 0; g < gx; g++) e.encode(c.s[CTX_LAST_X + off + (g >> shift)], 1);
    if (gx < max_group) e.encode(c.s[CTX_LAST_X + off + (gx >> shift)], 0);
    for (int g = 0; g < gy; g++) e.encode(c.s[CTX_LAST_Y + off + (g >> shift)], 1);
    if (gy < max_group) e.encode(c.s[CTX_LAST_Y + off + (gy >> shift)], 0);
    if (gx > 3) e.bypass_bits((uint32_t)(lx - kMinGroup[gx]), (gx >> 1) - 1);
    if (gy > 3) e.bypass_bits((uint32_t)(ly - kMinGroup[gy]), (gy >> 1) - 1);
    uint64_t csbf = 0;
    int c1_carry = 1;
    bool first_sb = true;
    for (int i = last_sb; i >= 0; i--) {
        const int sbp = S.sb_order[i], sbx = sbp % S.sbw, sby = sbp / S.sbw;
        const int right = sbx + 1 < S.sbw ? (int)((csbf >> (sbp + 1)) & 1) : 0;
        const int below = sby + 1 < S.sbw ? (int)((csbf >> (sbp + S.sbw)) & 1) : 0;
        const int prev = right | (below << 1);
        const bool infer_dc = i < last_sb && i > 0;
        const uint8_t* sig_row = sig_table().v[p.log2n - 2][p.scan][prev][p.luma ? 1 : 0][sbp == 0 ? 1 : 0];
        auto lvl = [&](int k) {
            const int pp = S.pos_order[k];
            return levels[((sby << 2) + (pp >> 2)) * n + (sbx << 2) + (pp & 3)];
        };
        bool any = false;
        for (int k = 0; k < 16; k++) any |= lvl(k) != 0;
        if (infer_dc) e.encode(c.s[CTX_CSBF + (p.luma ? 0 : 2) + ((right | below) & 1)], any ? 1 : 0);
        const bool coded = infer_dc ? any : true;
        if (coded) csbf |= 1ull << sbp;
        if (!coded) continue;
        SbEntries E;
        if (i == last_sb) {
            E.k[0] = (uint8_t)last_in_sb;
            E.n = 1;
        }
        const int top = i == last_sb ? last_in_sb - 1 : 15;
        for (int k = top; k >= 0; k--) {
            const int pp = S.pos_order[k];
            const bool sig = lvl(k) != 0;
            if (!(k == 0 && infer_dc && E.n == 0)) {
                const int cx = ts_ctx >= 0 ? ts_ctx : sig_row[pp];
                e.encode(c.s[CTX_SIG + cx], sig ? 1 : 0);
            }
            if (sig) {
                E.k[E.n] = (uint8_t)k;
                E.n++;
            }
        }
        if (!E.n) continue;
        int ctx_set = (i == 0 || !p.luma) ? 0 : 2;
        if (!first_sb && c1_carry == 0) ctx_set++;
        first_sb = false;
        int c1 = 1, first_g1 = 16;
        uint32_t g1mask = 0;
        const int ng1 = std::min(E.n, 8);
        for (int j = 0; j < ng1; j++) {
            const int cx = ctx_set * 4 + std::min(c1, 3) + (p.luma ? 0 : 16);
            const int g = std::abs(lvl(E.k[j])) > 1 ? 1 : 0;
            e.encode(c.s[CTX_G1 + cx], g);
            g1mask |= (uint32_t)g << j;
            if (g) {
                c1 = 0;
                if (first_g1 == 16) first_g1 = j;
            } else if (c1 > 0 && c1 < 3) c1++;
        }
        c1_carry = c1;
        bool g2 = false;
        if (first_g1 < 16) {
            g2 = std::abs(lvl(E.k[first_g1])) > 2;
            e.encode(c.s[CTX_G2 + ctx_set + (p.luma ? 0 : 4)], g2 ? 1 : 0);
        }
        const bool escape = E.n > 8 || __builtin_popcount(g1mask) > 1 || g2;
        if (p.bypass_alignment && escape) e.align_bypass();
        const bool hidden = p.sign_hiding && (E.k[0] - E.k[E.n - 1]) > 3 && !p.tqb && !rdpcm_active(p, r);
        const int nsign = E.n - (hidden ? 1 : 0);
        for (int j = 0; j < nsign; j++) e.bypass(lvl(E.k[j]) < 0 ? 1 : 0);
        const int sb_type = 2 * (p.luma ? 1 : 0) + ((r.ts || p.tqb) ? 1 : 0);
        int rice = p.persistent_rice ? (c.stat_coeff[sb_type] >> 2) : 0;
        bool first_rem = true;
        for (int j = 0; j < E.n; j++) {
            const int g1 = (g1mask >> j) & 1;
            const bool has_g2 = j == first_g1;
            const int base = 1 + g1 + ((has_g2 && g2) ? 1 : 0);
            const int thr = j < 8 ? (has_g2 ? 3 : 2) : 1;
            if (base != thr) continue;
            const uint32_t rem = (uint32_t)(std::abs(lvl(E.k[j])) - base);
            enc_remaining(e, rem, rice, p.extended_precision, p.bit_depth);
            if (p.persistent_rice && first_rem) {
                first_rem = false;
                uint8_t& st = c.stat_coeff[sb_type];
                if (rem >= (3u << (st >> 2))) st++;
                else if (2 * rem < (1u << (st >> 2)) && st > 0) st--;
            }
            if ((int64_t)base + rem > (3ll << rice)) rice = p.persistent_rice ? rice + 1 : std::min(rice + 1, 4);
        }
    }
}

}  // namespace hpvcd

// ---------------------------------------------------------------------------
// table introspection for the C API (pins against the reference's tables)
// ---------------------------------------------------------------------------
extern "C" int32_t hpvcd_cabac_num_elements(void) { return hpvcd::kNumCtxElements; }

extern "C" int32_t hpvcd_cabac_element(int32_t i, const char** name, int32_t* offset, int32_t* count) {
    if (i < 0 || i >= hpvcd::kNumCtxElements || !name || !offset || !count) return -1;
    *name = hpvcd::kCtxElements[i].name;
    *offset = hpvcd::kCtxElements[i].off;
    *count = hpvcd::kCtxElements[i].count;
    return 0;
}

// initValues (3 x CTX_COUNT), rangeTabLps (64 x 4), transIdxLps / Mps (64 each);
// returns CTX_COUNT
extern "C" int32_t hpvcd_cabac_tables(uint8_t* init, uint8_t* range_lps, uint8_t* trans_lps, uint8_t* trans_mps) {
    if (init) std::memcpy(init, hpvcd::ctx_init_table(), sizeof(uint8_t) * 3 * hpvcd::CTX_COUNT);
    for (int s = 0; s < 64; s++) {
        if (range_lps) std::memcpy(range_lps + 4 * s, hpvcd::cabac_range_lps_row(s), 4);
        if (trans_lps) trans_lps[s] = hpvcd::cabac_trans_lps(s);
        if (trans_mps) trans_mps[s] = hpvcd::cabac_trans_mps(s);
    }
    return hpvcd::CTX_COUNT;
}

// packed context states (bit 6 valMps, bits 0-5 pStateIdx) after 9.3.2.2
// initialisation at SliceQpY `qp`
extern "C" int32_t hpvcd_cabac_init_states(int32_t init_type, int32_t qp, uint8_t* out) {
    if (init_type < 0 || init_type > 2 || !out) return -1;
    hpvcd::Contexts c;
    c.init(init_type, qp);
    std::memcpy(out, c.s, hpvcd::CTX_COUNT);
    return hpvcd::CTX_COUNT;
}
