// cabac.h — CABAC for the HEVC syntax of the record producer.
//
// Context models and their initialisation (H.265 9.3.2.2; the values the
// reference decoder uses, src/cabac/contexts.rs:108-430), the arithmetic
// decoding engine (9.3.4.3; src/cabac/engine.rs:219 decode_bin, bypass,
// terminate, the PCM byte alignment of :313), the arithmetic encoding engine
// of 9.3.5 (the reference has none: it pairs with the decoder above), and
// residual_coding in both directions (7.3.8.11, 9.3.4.2.5-7;
// src/cabac/residual_full.rs:392).
#pragma once

#include <cstdint>
#include <cstring>
#include <vector>

#include "bitio.h"

namespace hpvcd {

// ---------------------------------------------------------------------------
// context index layout (one flat array per slice)
// ---------------------------------------------------------------------------
enum CtxOff : int {
    CTX_SAO_MERGE = 0,
    CTX_SAO_TYPE = CTX_SAO_MERGE + 1,
    CTX_SPLIT_CU = CTX_SAO_TYPE + 1,
    CTX_TQB = CTX_SPLIT_CU + 3,
    CTX_SKIP = CTX_TQB + 1,
    CTX_PRED_MODE = CTX_SKIP + 3,
    CTX_PART_MODE = CTX_PRED_MODE + 1,
    CTX_PREV_INTRA = CTX_PART_MODE + 4,
    CTX_CHROMA_MODE = CTX_PREV_INTRA + 1,
    CTX_RQT_ROOT_CBF = CTX_CHROMA_MODE + 1,
    CTX_MERGE_FLAG = CTX_RQT_ROOT_CBF + 1,
    CTX_MERGE_IDX = CTX_MERGE_FLAG + 1,
    CTX_INTER_PRED = CTX_MERGE_IDX + 1,
    CTX_REF_IDX = CTX_INTER_PRED + 5,
    CTX_MVP_FLAG = CTX_REF_IDX + 2,
    CTX_MVD = CTX_MVP_FLAG + 1,  // abs_mvd_greater0_flag, abs_mvd_greater1_flag
    CTX_SPLIT_TF = CTX_MVD + 2,
    CTX_CBF_LUMA = CTX_SPLIT_TF + 3,
    CTX_CBF_CHROMA = CTX_CBF_LUMA + 2,
    CTX_QP_DELTA = CTX_CBF_CHROMA + 5,
    CTX_TS = CTX_QP_DELTA + 2,
    CTX_LAST_X = CTX_TS + 2,
    CTX_LAST_Y = CTX_LAST_X + 18,
    CTX_CSBF = CTX_LAST_Y + 18,
    CTX_SIG = CTX_CSBF + 4,
    CTX_G1 = CTX_SIG + 44,
    CTX_G2 = CTX_G1 + 24,
    CTX_RDPCM_FLAG = CTX_G2 + 6,
    CTX_RDPCM_DIR = CTX_RDPCM_FLAG + 2,
    CTX_RES_SCALE = CTX_RDPCM_DIR + 2,
    CTX_RES_SIGN = CTX_RES_SCALE + 8,
    CTX_ACT = CTX_RES_SIGN + 2,
    CTX_COUNT = CTX_ACT + 1
};

// One syntax element's contexts: name, offset, count (for the table pins).
struct CtxElement {
    const char* name;
    int off, count;
};
extern const CtxElement kCtxElements[];
extern const int kNumCtxElements;

// initValue of every context for initType 0 (I), 1 and 2 (9.3.2.2)
const uint8_t (&ctx_init_table())[3][CTX_COUNT];
// rangeTabLps row of pStateIdx s (Table 9-52), transIdxLps / transIdxMps
const uint8_t* cabac_range_lps_row(int s);
uint8_t cabac_trans_lps(int s);
uint8_t cabac_trans_mps(int s);

// H.265 Table 9-52 (rangeTabLps) and Table 9-53 (transIdxLps)
extern const uint8_t kCabacRangeLps[64][4];
extern const uint8_t kCabacTransLps[64];
// one lookup per decision, built from those two tables: for a packed
// context (valMps << 6 | pStateIdx) and range class, rangeLps | the packed
// context after an MPS << 8 | after an LPS << 16
extern uint32_t kCabacDecision[128][4];

// packed context: bits 0-5 pStateIdx, bit 6 valMps
struct Contexts {
    uint8_t s[CTX_COUNT];
    uint8_t stat_coeff[4];  // persistent Rice statistics (9.3.2.2, RExt)
    void init(int init_type, int slice_qp);
};

// ---------------------------------------------------------------------------
// arithmetic decoding engine (9.3.4.3)
// ---------------------------------------------------------------------------
struct CabacDec {
    // The 9-bit ivlOffset register and the stream bits after it share one
    // 64-bit window: val = ivlOffset << k | the next k unread bits.  Reading
    // n more bits into ivlOffset (renormalisation, bypass) only moves the
    // boundary (k -= n), and ivlOffset >= x is val >= x << k, so a bin needs
    // no bit reads; the window takes 32 more bits whenever k drops below 24.
    // Past the end the stream reads as one bits (as the reference reads it).
    const uint8_t* data = nullptr;
    size_t len = 0;         // bytes of the slice data
    size_t loaded = 0;      // stream bits taken into the window
    uint64_t val = 0;
    int k = 0;              // unread bits held below ivlOffset
    uint32_t range = 510;
    bool overrun = false;

    // bits consumed: ivlOffset's bits included, the window's unread ones not
    size_t pos() const { return loaded - (size_t)k; }
    uint32_t offset() const { return (uint32_t)(val >> k); }

    // k <= 23 on entry: val < 2^(9 + k) has room for 32 more bits
    __attribute__((always_inline)) void refill() {
        const size_t byte = loaded >> 3;  // loaded is a multiple of 32
        uint32_t w;
        if (byte + 4 <= len) {
            w = (uint32_t)data[byte] << 24 | (uint32_t)data[byte + 1] << 16 | (uint32_t)data[byte + 2] << 8 |
                (uint32_t)data[byte + 3];
        } else {
            w = 0;
            for (int i = 0; i < 4; i++) w = w << 8 | (byte + i < len ? data[byte + i] : 0xffu);
        }
        val = val << 32 | w;
        k += 32;
        loaded += 32;
    }
    __attribute__((always_inline)) void keep_filled() {
        if (k < 24) refill();
    }
    // n in 0..24 raw bits after ivlOffset (PCM samples, alignment); they do
    // not enter ivlOffset
    uint32_t read_bits(int n) {
        if (n == 0) return 0;
        if (k < n) refill();
        const uint64_t low = val & ((1ull << (k - n)) - 1ull);
        const uint32_t v = (uint32_t)((val >> (k - n)) & ((1ull << n) - 1ull));
        val = (val >> k << (k - n)) | low;
        k -= n;
        keep_filled();
        return v;
    }
    // read past the end of the data, or flagged malformed by a caller
    bool failed() const { return overrun || pos() > len * 8; }
    void start(const uint8_t* d, size_t n) {
        data = d;
        len = n;
        loaded = 0;
        val = 0;
        k = 0;
        overrun = false;
        refill();
        init_engine();
    }
    void init_engine() {  // 9.3.2.5: ivlOffset = the next 9 bits
        range = 510;
        if (k < 9) refill();
        val &= (1ull << k) - 1ull;
        k -= 9;
        keep_filled();
    }
    // DecodeDecision (9.3.4.3.2) with the state transition (9.3.4.3.2.2)
    // and renormalisation (9.3.4.3.3), without data-dependent branches
    __attribute__((always_inline)) int decode(uint8_t& ctx) {
        const uint32_t st = ctx;
        const uint32_t e = kCabacDecision[st & 127][(range >> 6) & 3];
        const uint32_t lps = e & 0xffu;
        range -= lps;
        const uint64_t scaled = (uint64_t)range << k;
        const uint32_t m = 0u - (uint32_t)(val >= scaled);  // all ones on an LPS
        val -= scaled & (uint64_t)(int64_t)(int32_t)m;
        range ^= (range ^ lps) & m;
        ctx = (uint8_t)(((e >> 8) ^ (((e >> 8) ^ (e >> 16)) & m)) & 0xffu);
        const int n = __builtin_clz(range) - 23;  // 0 when range >= 256
        range <<= n;
        k -= n;
        keep_filled();
        return (int)((st >> 6) ^ (m & 1u));
    }
    // DecodeBypass (9.3.4.3.4)
    __attribute__((always_inline)) int bypass() {
        k -= 1;
        const uint64_t scaled = (uint64_t)range << k;
        const uint32_t m = 0u - (uint32_t)(val >= scaled);
        val -= scaled & (uint64_t)(int64_t)(int32_t)m;
        keep_filled();
        return (int)(m & 1u);
    }
    // n bypass bins, first bin most significant.  With ivlCurrRange fixed,
    // n bypass steps are the long division of ivlOffset << n | the next n
    // bits by ivlCurrRange: the quotient is the bins, the remainder the new
    // ivlOffset (ivlOffset < ivlCurrRange keeps the quotient below 2^n).
    __attribute__((always_inline)) uint32_t bypass_bits(int n) {
        uint32_t v = 0;
        if (n < 8) {  // a divide costs more than a few single bins
            for (int i = 0; i < n; i++) v = v << 1 | (uint32_t)bypass();
            return v;
        }
        while (n > 0) {
            const int c = n < 16 ? n : 16;
            const int sh = k - c;
            const uint32_t top = (uint32_t)(val >> sh);  // < 2^(9 + c)
            const uint32_t q = top / range, r = top - q * range;
            val = ((uint64_t)r << sh) | (val & ((1ull << sh) - 1ull));
            k = sh;
            keep_filled();
            v = v << c | q;
            n -= c;
        }
        return v;
    }
    int terminate() {
        range -= 2;
        if (val >= ((uint64_t)range << k)) return 1;
        renorm();
        return 0;
    }
    void renorm() {
        if (range < 256) {
            const int n = __builtin_clz(range) - 23;  // shifts that bring range to [256, 510]
            range <<= n;
            k -= n;
            keep_filled();
        }
    }
    void align_bypass() { range = 256; }
    // to the next byte boundary after a terminating bin (pcm_flag, end of
    // a sub-stream): the bits the engine holds are given back
    void byte_align() {
        const int drop = (int)((8 - (pos() & 7)) & 7);
        if (drop) read_bits(drop);
    }
};

// ---------------------------------------------------------------------------
// arithmetic encoding engine (9.3.5, bit-serial with outstanding bits)
// ---------------------------------------------------------------------------
struct CabacEnc {
    BitWriter* w = nullptr;
    uint32_t low = 0, range = 510;
    uint32_t outstanding = 0;
    bool first = true;

    void start(BitWriter* out) {
        w = out;
        low = 0;
        range = 510;
        outstanding = 0;
        first = true;
    }
    void put(uint32_t b) {
        if (first) first = false;
        else w->bit(b);
        while (outstanding) {
            w->bit(1 - b);
            outstanding--;
        }
    }
    void renorm() {
        while (range < 256) {
            if (low < 256) put(0);
            else if (low >= 512) {
                low -= 512;
                put(1);
            } else {
                low -= 256;
                outstanding++;
            }
            range <<= 1;
            low <<= 1;
        }
    }
    void encode(uint8_t& ctx, int bin);
    void bypass(int bin) {
        low <<= 1;
        if (bin) low += range;
        if (low >= 1024) {
            put(1);
            low -= 1024;
        } else if (low < 512) put(0);
        else {
            low -= 512;
            outstanding++;
        }
    }
    void bypass_bits(uint32_t v, int n) {
        for (int i = n - 1; i >= 0; i--) bypass((v >> i) & 1);
    }
    // terminating bin; a one flushes the engine, whose last written bit is
    // the rbsp_stop_one_bit / alignment bit that follows
    void terminate(int bin) {
        range -= 2;
        if (bin) {
            low += range;
            range = 2;
            renorm();
            put((low >> 9) & 1);
            w->bits(((low >> 7) & 3) | 1, 2);
        } else {
            renorm();
        }
    }
    void align_bypass() { range = 256; }
};

// ---------------------------------------------------------------------------
// residual_coding
// ---------------------------------------------------------------------------
constexpr int SCAN_DIAG = 0, SCAN_HOR = 1, SCAN_VER = 2;

struct ResidualParams {
    int log2n = 2;
    bool luma = true;
    int scan = SCAN_DIAG;
    bool ts_allowed = false;        // transform_skip_enabled && !tqb && log2n <= Log2MaxTransformSkipSize
    bool tqb = false;
    bool explicit_rdpcm = false;    // explicit_rdpcm_enabled_flag && inter CU
    bool implicit_rdpcm = false;    // implicit_rdpcm_enabled_flag && intra && mode 10 / 26
    bool sign_hiding = false;
    bool persistent_rice = false;
    bool ts_context = false;
    bool bypass_alignment = false;
    bool extended_precision = false;
    int bit_depth = 8;
};

// decoded / encoded residual of one TB: levels n*n row-major
struct ResidualOut {
    bool ts = false;
    int rdpcm = -1;  // -1 none, 0 horizontal, 1 vertical (explicit)
    // nonzero positions (raster) of the decoded levels, when the caller gives
    // the buffer (n*n entries); the decoder then writes only those levels
    // and expects the rest of the array to be zero already
    uint16_t* nz = nullptr;
    int nnz = -1;
};

void residual_decode(CabacDec& d, Contexts& c, const ResidualParams& p, int32_t* levels, ResidualOut& r);
void residual_encode(CabacEnc& e, Contexts& c, const ResidualParams& p, const int32_t* levels, const ResidualOut& r);
// whether the sign of sub-block coefficients at the lowest scan position is
// hidden; the producer makes each such sign follow the parity of the sum of
// the sub-block's magnitudes (7.3.8.11 signHidden)
void residual_hide_signs(const ResidualParams& p, const ResidualOut& r, int32_t* levels);
// scan order positions (x + y * w) of a w x w grid, w in {1, 2, 4, 8}
const uint8_t* scan_order(int log2w, int scan);
// mode-dependent scan (7.4.9.11; src/decode.rs:4481 luma_scan, :4495 chroma_scan)
int scan_for(int mode, int log2n, bool luma, bool chroma444);

}  // namespace hpvcd
