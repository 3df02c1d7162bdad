// kernels.hip — gfx950 (CDNA4, wave64) reconstruction kernels.
//
//   k_residual<R, LOG2N>  dequantisation + inverse DCT/DST or transform skip /
//                         bypass, rotation, RDPCM, cross-component prediction;
//                         batched over every TB of one size class
//                         (reference: src/transform.rs:489-910,
//                         src/decode.rs:4724-4870, 5280-5504, 8418)
//   k_intra<T, R>         intra prediction + residual add as a persistent
//                         CTU wavefront; Y, Cb and Cr planes are independent
//                         units (reference: src/intra.rs:84-374,
//                         src/decode.rs:4641-4990, 5170-5280)
//   k_deblock<T>          vertical then horizontal edge passes, each over
//                         luma and chroma together (reference:
//                         src/decode.rs:1774-2143, src/deblock.rs:216-330)
//   k_sao<T>              band / edge offset, out of place from the
//                         deblocked planes (reference: src/decode.rs:2258,
//                         src/sao.rs:179-470)
//   k_rgb<T, ORIENTED>    YUV -> RGB8/RGB16 with crop, grid placement and orientation
//                         (reference: src/yuv.rs:187-396, src/lib.rs:1610)
//
// None of this is a dense contraction; all arithmetic is integer and stays
// bit-exact with the CPU reference.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.h"

namespace hpvcd {

// ---------------------------------------------------------------------------
// shared helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int64_t clampl(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int32_t wrap16(int32_t v) { return (int32_t)(int16_t)(uint16_t)(uint32_t)v; }
__device__ __forceinline__ int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

// Orders this wave's LDS traffic: DS operations of one wave execute in
// order, so a compiler barrier is enough between a write and another lane's
// read of the same location.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// HEVC transform matrix entries as compile-time functions: after unrolling,
// every coefficient of the inverse transforms below is an immediate operand.
// cos64(j) is the normative integer approximation of 64*sqrt(2)*cos(j*pi/64)
// (H.265 8.6.4.2); the n-point matrix row k, column m is the 32-point matrix
// row k*32/n: c(k*32/n*(2m+1) mod 128), folded into 0..32 with its sign.
__host__ __device__ constexpr int cos64(int j) {
    constexpr int8_t t[33] = {64, 90, 90, 90, 89, 88, 87, 85, 83, 82, 80, 78, 75, 73, 70, 67, 64,
                              61, 57, 54, 50, 46, 43, 38, 36, 31, 25, 22, 18, 13, 9,  4,  0};
    return t[j];
}
__host__ __device__ constexpr int dct_coef(int n, int k, int m) {
    const int a = (k * (32 / n) * (2 * m + 1)) & 127;
    return a < 32 ? cos64(a) : a < 64 ? -cos64(64 - a) : a < 96 ? -cos64(a - 64) : cos64(128 - a);
}
// 4-point DST-VII of 4x4 intra luma, row k column m (H.265 8.6.4.2 eq. 8-315)
__host__ __device__ constexpr int dst4_coef(int k, int m) {
    constexpr int8_t t[16] = {29, 55, 74, 84, 74, 74, 0, -74, 84, -29, -74, 55, 55, -84, 74, -29};
    return t[k * 4 + m];
}
static_assert(dct_coef(32, 1, 0) == 90 && dct_coef(32, 31, 31) == -4 && dct_coef(32, 31, 15) == -90 && dct_coef(4, 1, 1) == 36 &&
                  dct_coef(4, 3, 0) == 36 && dct_coef(4, 2, 1) == -64 && dct_coef(8, 1, 3) == 18,
              "HEVC DCT matrix entries");
__constant__ int16_t kLevelScale[6] = {40, 45, 51, 57, 64, 72};
__constant__ uint8_t kBeta[52] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  6,  7,
                                  8,  9,  10, 11, 12, 13, 14, 15, 16, 17, 18, 20, 22, 24, 26, 28, 30, 32,
                                  34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64};
__constant__ uint8_t kTc[54] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  0,  1,  1,  1,  1,  1,  1,  1,  1,  1,
                                2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 5, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 22, 24};
__constant__ int8_t kIntraAngle[33] = {32,  26,  21,  17,  13,  9,  5,  2,  0,  -2, -5, -9, -13, -17, -21, -26, -32,
                                       -26, -21, -17, -13, -9, -5, -2, 0,  2,  5,  9,  13,  17,  21,  26,  32};
__constant__ int16_t kInvAngle[15] = {-4096, -1638, -910, -630, -482, -390, -315, -256,
                                      -315,  -390,  -482, -630, -910, -1638, -4096};

__device__ __forceinline__ int max_range_of(int bd, bool ep) { return ep ? (bd + 6 > 15 ? bd + 6 : 15) : 15; }

__device__ __forceinline__ int qpc_of(int qpi, int chroma_idc, int bd) {
    qpi = clampi(qpi, -6 * (bd - 8), 57);
    if (chroma_idc != 1) return qpi < 51 ? qpi : 51;
    if (qpi < 30) return qpi;
    if (qpi > 43) return qpi - 6;
    const int t[14] = {29, 30, 31, 32, 33, 33, 34, 34, 35, 35, 36, 36, 37, 37};
    return t[qpi - 30];
}

// ---------------------------------------------------------------------------
// K1: residual (dequant + inverse transform / TS / bypass + RExt ops + CCP)
// ---------------------------------------------------------------------------
// wrapping 32-bit sum and 24-bit multiply-add (inputs of the inverse
// transforms are clipped to at most 23 bits; the accumulation wraps as the
// reference's i32 arithmetic does)
__device__ __forceinline__ int32_t mad24w(int32_t acc, int c, int32_t x) { return wadd(acc, __mul24(c, x)); }

// n-point inverse DCT of one column or row (src/transform.rs:489): the even
// outputs come from the n/2-point inverse of the even inputs, the odd half
// from the odd basis rows, mirrored (the matrix rows are symmetric /
// antisymmetric), which is the direct product's value exactly.
template <int N> struct InvDct {
    __device__ __forceinline__ static void run(const int32_t (&in)[N], int32_t (&out)[N]) {
        constexpr int H = N / 2;
        int32_t ev[H], e[H];
#pragma unroll
        for (int j = 0; j < H; j++) ev[j] = in[2 * j];
        InvDct<H>::run(ev, e);
#pragma unroll
        for (int k = 0; k < H; k++) {
            int32_t o = 0;
#pragma unroll
            for (int j = 1; j < N; j += 2) o = mad24w(o, dct_coef(N, j, k), in[j]);
            out[k] = wadd(e[k], o);
            out[N - 1 - k] = (int32_t)((uint32_t)e[k] - (uint32_t)o);
        }
    }
};
template <> struct InvDct<2> {
    __device__ __forceinline__ static void run(const int32_t (&in)[2], int32_t (&out)[2]) {
        out[0] = (int32_t)(64u * (uint32_t)wadd(in[0], in[1]));
        out[1] = (int32_t)(64u * ((uint32_t)in[0] - (uint32_t)in[1]));
    }
};

__device__ __forceinline__ void inv_dst4(const int32_t (&in)[4], int32_t (&out)[4]) {
#pragma unroll
    for (int m = 0; m < 4; m++) {
        int32_t acc = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) acc = mad24w(acc, dst4_coef(k, m), in[k]);
        out[m] = acc;
    }
}

// One row of n residual samples to the pool: 16- or 8-byte stores when the
// row is whole and aligned, else one sample at a time.
template <int N, typename R>
__device__ __forceinline__ void store_res_row(R* p, const int32_t (&v)[N], int vw) {
    constexpr int BYTES = N * (int)sizeof(R);
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    if (vw == N && BYTES >= 16 && (a & 15) == 0) {
        constexpr int PER = 16 / (int)sizeof(R);
#pragma unroll
        for (int q = 0; q < N / PER; q++) {
            uint32_t w[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                if (sizeof(R) == 2)
                    w[t] = (uint32_t)(uint16_t)v[q * PER + 2 * t] | ((uint32_t)(uint16_t)v[q * PER + 2 * t + 1] << 16);
                else
                    w[t] = (uint32_t)v[q * PER + t];
            }
            reinterpret_cast<uint4*>(p)[q] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    } else if (vw == N && BYTES == 8 && (a & 7) == 0) {
        reinterpret_cast<uint2*>(p)[0] =
            make_uint2((uint32_t)(uint16_t)v[0] | ((uint32_t)(uint16_t)v[1] << 16),
                       (uint32_t)(uint16_t)v[2] | ((uint32_t)(uint16_t)v[3] << 16));
    } else {
#pragma unroll
        for (int m = 0; m < N; m++)
            if (m < vw) p[m] = (R)v[m];
    }
}

// N lanes of one wave per TB: a TB's samples sit in LDS rows of N + 1 words
// (conflict-free column and row access); lane j owns column j for the
// vertical pass and row j for the horizontal pass, so the TB needs no
// barrier beyond its wave.  BLOCK threads per workgroup (64, 128 or 256).
template <typename R, int LOG2N, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_residual(const DevResJob* __restrict__ jobs, uint32_t count,
                                                    const uint32_t* __restrict__ coeffs,
                                                    const uint8_t* __restrict__ scaling, R* __restrict__ res_pool) {
    constexpr int N = 1 << LOG2N;
    constexpr int NN = N * N;
    constexpr int P = N + 1;        // LDS row pitch in words
    constexpr int TBS = BLOCK / N;  // TBs per workgroup
    __shared__ int32_t blk[TBS * N * P];

    const int tid = threadIdx.x;
    const int slot = tid >> LOG2N, j = tid & (N - 1);
    const uint32_t i = blockIdx.x * TBS + slot;
    const bool valid = i < count;
    int32_t* B = blk + slot * N * P;

    // the job is the only load the coefficient addresses wait for
    DevResJob J = jobs[valid ? i : 0];
    if (!valid) J.flags = 0;
#pragma unroll
    for (int m = 0; m < N; m++) B[j * P + m] = 0;

    const uint32_t jf = J.flags;
    const int bd_l = 8 + (J.bds & 15), bd_c = 8 + (J.bds >> 4);
    const int bd = (jf & RJ_CHROMA) ? bd_c : bd_l;
    const bool ep = (jf & RJ_EP) != 0;
    const bool ts = (jf & RJ_TS) != 0;
    const bool tqb = (jf & RJ_TQB) != 0;
    const bool cbf = (jf & RJ_CBF) != 0;
    const bool pcm = (jf & RJ_PCM) != 0;
    const bool i16 = (jf & RJ_I16) != 0;
    const int max_range = max_range_of(bd, ep);
    const bool transform = (jf & RJ_TRANSFORM) != 0;
    const bool use_dst = (jf & RJ_DST) != 0;
    const bool rotate = (jf & RJ_ROTATE) != 0;
    const int rdpcm = (jf & RJ_RDPCM) ? ((jf & RJ_RDPCM_V) ? 1 : 0) : -1;
    wave_sync();

    // scatter + dequantise the nonzero levels (src/transform.rs:764-910); a
    // rotated block (src/decode.rs:4356) takes each level at its mirrored
    // position
    if (cbf) {
        const int64_t bd_shift = bd + LOG2N + 10 - max_range;
        const int qp = J.qp;
        const int64_t factor = (int64_t)kLevelScale[qp % 6] * ((int64_t)1 << (qp / 6)) * 16;
        const int64_t add = (int64_t)1 << (bd_shift - 1);
        const int64_t cmin = -((int64_t)1 << max_range), cmax = ((int64_t)1 << max_range) - 1;
        int tr_shift = max_range - bd - LOG2N;
        if (ep && tr_shift < 0) tr_shift = 0;
        const int32_t tr_add = tr_shift > 0 ? 1 << (tr_shift - 1) : 0;
        const uint8_t* sc = J.scaling >= 0 ? scaling + J.scaling : nullptr;
        const uint32_t* co = coeffs + J.coeff;
        for (uint32_t k = j; k < J.coeff_cnt; k += N) {
            uint32_t word = co[k];
            int pos = HPVCD_COEFF_POS(word);
            int32_t level = HPVCD_COEFF_LEVEL(word);
            int32_t v;
            if (tqb) {
                v = i16 ? clampi(level, -32768, 32767) : level;
            } else if (pcm) {
                v = level << qp;  // pcm_sample << (BitDepth - PcmBitDepth), src/decode.rs:2788
            } else {
                int64_t f = factor;
                if (sc) {
                    int64_t m;
                    if (LOG2N == 2) m = sc[pos];
                    else if (pos == 0 && LOG2N >= 4) m = sc[64];
                    else {
                        int yy = pos >> LOG2N, xx = pos & (N - 1);
                        int down = LOG2N > 3 ? LOG2N - 3 : 0;
                        m = sc[(yy >> down) * 8 + (xx >> down)];
                    }
                    f = (factor / 16) * m;
                }
                int64_t l = clampl(level, cmin, cmax);
                v = (int32_t)clampl((l * f + add) >> bd_shift, cmin, cmax);
                if (ts) {
                    v = tr_shift >= 0 ? (wadd(v, tr_add) >> tr_shift) : (int32_t)((uint32_t)v << (-tr_shift));
                    if (i16) v = wrap16(v);
                }
            }
            if (rotate) pos = NN - 1 - pos;
            B[(pos >> LOG2N) * P + (pos & (N - 1))] = v;
        }
    }
    wave_sync();

    // vertical pass, lane j = column j: first inverse transform stage (only
    // the nx columns that hold levels), or vertical RDPCM accumulation
    // (src/decode.rs:8418)
    const int32_t cl_min = -(1 << max_range), cl_max = (1 << max_range) - 1;
    if (transform) {
        int32_t c[N], o[N];
#pragma unroll
        for (int k = 0; k < N; k++) c[k] = B[k * P + j];
        if constexpr (N == 4) {
            if (use_dst) inv_dst4(c, o);
            else InvDct<4>::run(c, o);
        } else {
            InvDct<N>::run(c, o);
        }
        const bool live = j < (int)J.nx;
#pragma unroll
        for (int m = 0; m < N; m++) B[m * P + j] = live ? clampi(wadd(o[m], 64) >> 7, cl_min, cl_max) : 0;
    } else if (rdpcm == 1) {
        int32_t acc = B[j];
#pragma unroll
        for (int y = 1; y < N; y++) {
            acc = wadd(B[y * P + j], acc);
            if (i16) acc = wrap16(acc);
            B[y * P + j] = acc;
        }
    }
    wave_sync();
    if (!valid) return;

    // horizontal pass, lane j = row j: second inverse transform stage, or
    // horizontal RDPCM accumulation; then cross-component prediction
    const int shift2 = max_range + 5 - bd;
    const int32_t add2 = 1 << (shift2 - 1);
    int32_t row[N], v[N];
#pragma unroll
    for (int m = 0; m < N; m++) row[m] = B[j * P + m];
    if (transform) {
        if constexpr (N == 4) {
            if (use_dst) inv_dst4(row, v);
            else InvDct<4>::run(row, v);
        } else {
            InvDct<N>::run(row, v);
        }
#pragma unroll
        for (int m = 0; m < N; m++) {
            v[m] = wadd(v[m], add2) >> shift2;
            if (i16) v[m] = wrap16(v[m]);
        }
    } else {
        v[0] = row[0];
#pragma unroll
        for (int m = 1; m < N; m++) {
            int32_t a = row[m];
            if (rdpcm == 0) {
                a = wadd(a, v[m - 1]);
                if (i16) a = wrap16(a);
            }
            v[m] = a;
        }
    }
    if (j >= (int)J.vh) return;
    R* out = res_pool + J.out + (size_t)j * J.stride;
    if (jf & RJ_CCP) {
        // src/decode.rs:5488 apply_cross_comp_residual
        const R* lres = out - J.luma_delta;
#pragma unroll
        for (int m = 0; m < N; m++) {
            if (m >= (int)J.vw) break;
            int64_t ly = (int64_t)lres[m];
            int64_t sy = (ly << bd_c) >> bd_l;
            int64_t delta = ((int64_t)J.res_scale * sy) >> 3;
            v[m] = (int32_t)clampl((int64_t)v[m] + delta, INT32_MIN, INT32_MAX);
        }
    }
    store_res_row<N, R>(out, v, J.vw);
}

// ---------------------------------------------------------------------------
// K1b: inverse adaptive colour transform (src/act.rs:43, YCgCo-R lifting) of
// the co-located 4:4:4 residual triples of ACT TUs, in place in the residual
// pool after the cross-component pass.  One wave per TU; list holds the luma
// TB of each triple.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) k_act(const DevPic* __restrict__ pics, const hpvcd_tb* __restrict__ tbs,
                                             const uint32_t* __restrict__ tb_pic, const uint32_t* __restrict__ list,
                                             uint32_t count, R* __restrict__ res_pool) {
    const int lane = threadIdx.x & 63;
    const uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= count) return;
    const uint32_t tbi = list[i];
    const hpvcd_tb tb = tbs[tbi];
    const DevPic& P = pics[tb_pic[tbi]];
    const int log2n = tb.log2n, n = 1 << log2n;
    R* r0 = res_pool + P.res_off[0];
    R* r1 = res_pool + P.res_off[1];
    R* r2 = res_pool + P.res_off[2];
    for (int e = lane; e < n * n; e += 64) {
        const int px = tb.x + (e & (n - 1)), py = tb.y + (e >> log2n);
        if (px >= P.w || py >= P.h) continue;
        const size_t o = (size_t)py * P.w + px;  // 4:4:4: the three planes share one geometry
        const int32_t y = r0[o], cg = r1[o], co = r2[o];
        const int32_t t = (int32_t)((uint32_t)y - (uint32_t)(cg >> 1));
        const int32_t b = (int32_t)((uint32_t)t - (uint32_t)(co >> 1));
        r0[o] = (R)wadd(cg, t);
        r1[o] = (R)b;
        r2[o] = (R)wadd(b, co);
    }
}

// ---------------------------------------------------------------------------
// K1c: motion compensation (src/decode.rs:7000 motion_compensate, src/mc.rs)
//
// One wave per block of at most 8 x 8 luma samples of a PU (four blocks per
// workgroup); for each plane and list the clamped reference window ((bh +
// taps - 1) x (bw + taps - 1), src/mc.rs:310) is staged in the wave's LDS,
// filtered horizontally into an LDS
// row buffer (>> BitDepth - 8) and vertically (>> 6), and narrowed to i16 as
// the reference's 14-bit intermediates are.  The separable form equals the
// reference's full-sample / one-direction special cases exactly (the
// identity tap 64 is a shift by 6 on either side).  The lists then combine
// (uni / bi / explicit weighted, src/mc.rs:538-733) and the TB's residual
// is added (src/decode.rs:4724: Clip(pred + res); the residual pool is zero
// where a TB has none), so an inter CU leaves this kernel reconstructed and
// the wavefront only publishes its border samples.  Reference samples are
// the final planes of earlier pictures of the job, read in place.
// ---------------------------------------------------------------------------
__constant__ int8_t kLumaFilter[4][8] = {{0, 0, 0, 64, 0, 0, 0, 0},
                                         {-1, 4, -10, 58, 17, -5, 1, 0},
                                         {-1, 4, -11, 40, 40, -11, 4, -1},
                                         {0, 1, -5, 17, 58, -10, 4, -1}};
__constant__ int8_t kChromaFilter[8][4] = {{0, 64, 0, 0},    {-2, 58, 10, -2}, {-4, 54, 16, -2}, {-6, 46, 28, -4},
                                           {-4, 36, 36, -4}, {-4, 28, 46, -6}, {-2, 16, 54, -4}, {-2, 10, 58, -2}};

// Workgroups are dispatched to the 8 XCDs in turn (workgroup i to XCD i % 8),
// each XCD with its own L2.  xcd_block maps workgroup i to a block index such
// that each XCD takes one contiguous range of blocks: neighbouring blocks of
// a picture, whose reference windows overlap, then meet in one L2 (10 % off
// the 1080p batch's motion compensation).  A bijection on [0, n), so
// correctness never depends on the dispatch order.  Kernels whose grids put
// the plane or picture in a slow dimension keep their order: contiguous
// ranges there would give some XCDs only luma work.
__device__ __forceinline__ uint32_t xcd_block(uint32_t i, uint32_t n) {
    constexpr uint32_t kXcds = 8;
    const uint32_t q = n / kXcds, r = n % kXcds, x = i % kXcds, k = i / kXcds;
    return x * q + min(x, r) + k;
}

// The waves of a workgroup share nothing, so a wave needs no barrier beyond
// its own.
constexpr int kMcBlock = 8;
constexpr int kMcWaves = 4;
constexpr int kMcWinStride = 16;  // window rows: up to 8 + 7 samples
constexpr int kMcRowStride = 8;   // horizontally filtered rows: up to 8 samples

template <typename T, typename R>
__global__ void __launch_bounds__(64 * kMcWaves) k_mc(const DevPic* __restrict__ pics, const DevPu* __restrict__ pus,
                                                     const DevMcTask* __restrict__ tasks, uint32_t ntasks,
                                                     T* __restrict__ cur, const T* __restrict__ refs,
                                                     const R* __restrict__ res) {
    __shared__ int32_t win_all[kMcWaves][kMcWinStride * kMcWinStride];
    __shared__ int32_t rows_all[kMcWaves][kMcWinStride * kMcRowStride];
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const uint32_t ti = xcd_block(blockIdx.x, gridDim.x) * kMcWaves + (uint32_t)wave;
    if (ti >= ntasks) return;
    int32_t* win = win_all[wave];
    int32_t* rows = rows_all[wave];
    const DevMcTask tk = tasks[ti];
    const DevPu& U = pus[tk.pu];
    const DevPic& P = pics[tk.pic];
    const int flags = (int)U.flags;
    const int bw_l = min(kMcBlock, U.w - (int)tk.ox), bh_l = min(kMcBlock, U.h - (int)tk.oy);
    const int nplanes = P.chroma_idc ? 3 : 1;
    // the lane's sample of the block
    const int sx = lane & 7, sy = lane >> 3;
    for (int pl = 0; pl < nplanes; pl++) {
        const int sw = pl ? P.sub_w_log2 : 0, sh = pl ? P.sub_h_log2 : 0;
        const int bw = bw_l >> sw, bh = bh_l >> sh;
        const int bx = (U.x >> sw) + ((int)tk.ox >> sw), by = (U.y >> sh) + ((int)tk.oy >> sh);
        const int pw = pl ? P.cw : P.w, ph = pl ? P.ch : P.h;
        const int bd = pl ? P.bd_c : P.bd, s1 = bd - 8;
        const int taps = pl ? 4 : 8, half = taps / 2 - 1;
        const int ww = bw + taps - 1, wh = bh + taps - 1;
        const bool mine = sx < bw && sy < bh;
        int32_t pred[2] = {0, 0};
        for (int l = 0; l < 2; l++) {
            if (!((flags >> l) & 1)) continue;
            const T* rp = refs + pics[U.ref[l]].plane_off[pl];
            int mx = U.mv[l][0], my = U.mv[l][1], fx, fy, ix, iy;
            if (pl == 0) {
                ix = bx + (mx >> 2); fx = mx & 3;
                iy = by + (my >> 2); fy = my & 3;
            } else {
                mx = mx * 2 / (1 << sw);   // chroma vector, src/decode.rs:7293
                my = my * 2 / (1 << sh);
                ix = bx + (mx >> 3); fx = mx & 7;
                iy = by + (my >> 3); fy = my & 7;
            }
            // clamped reference window (RefPlane::sample, src/mc.rs:310)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int e = lane + 64 * k, r = e >> 4, c = e & 15;
                if (r < wh && c < ww) {
                    const int x = min(max(ix - half + c, 0), pw - 1), y = min(max(iy - half + r, 0), ph - 1);
                    win[e] = (int32_t)rp[(size_t)y * pw + x];
                }
            }
            wave_sync();
            const int8_t* hf = pl ? kChromaFilter[fx] : kLumaFilter[fx];
            const int8_t* vf = pl ? kChromaFilter[fy] : kLumaFilter[fy];
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const int e = lane + 64 * k, r = e >> 3, c = e & 7;
                if (r < wh && c < bw) {
                    int32_t acc = 0;
                    for (int t = 0; t < taps; t++) acc += (int32_t)hf[t] * win[r * kMcWinStride + c + t];
                    rows[r * kMcRowStride + c] = acc >> s1;
                }
            }
            wave_sync();
            if (mine) {
                int32_t acc = 0;
                for (int t = 0; t < taps; t++) acc += (int32_t)vf[t] * rows[(sy + t) * kMcRowStride + sx];
                pred[l] = wrap16(acc >> 6);
            }
            wave_sync();
        }
        const int px = bx + sx, py = by + sy;
        if (!mine || px >= pw || py >= ph) continue;
        const int32_t maxv = (1 << bd) - 1;
        const int comp = pl, dn = pl ? 1 : 0;
        int32_t v;
        if ((flags & 3) == 3) {
            if (U.weighted) {
                const int log2wd = U.log2_denom[dn] + 14 - bd;
                const int64_t rnd = ((int64_t)U.offset[0][comp] + U.offset[1][comp] + 1) << log2wd;
                v = (int32_t)(((int64_t)pred[0] * U.weight[0][comp] + (int64_t)pred[1] * U.weight[1][comp] + rnd) >>
                              (log2wd + 1));
            } else {
                const int shift = 15 - bd;
                v = (pred[0] + pred[1] + (1 << (shift - 1))) >> shift;
            }
        } else if (flags & 3) {
            const int l = (flags & 1) ? 0 : 1;
            if (U.weighted) {
                const int log2wd = U.log2_denom[dn] + 14 - bd;
                v = log2wd >= 1 ? ((pred[l] * U.weight[l][comp] + (1 << (log2wd - 1))) >> log2wd) + U.offset[l][comp]
                                : pred[l] * U.weight[l][comp] + U.offset[l][comp];
            } else {
                const int shift = 14 - bd;
                v = (pred[l] + (shift > 0 ? 1 << (shift - 1) : 0)) >> shift;
            }
        } else {
            v = 1 << (bd - 1);  // no usable reference: mid-grey (src/decode.rs:7206)
        }
        const int32_t pv = clampi(v, 0, maxv);
        const int32_t rv = (int32_t)res[P.res_off[pl] + (size_t)py * pw + px];
        cur[P.plane_off[pl] + (size_t)py * pw + px] = (T)clampi(wadd(pv, rv), 0, maxv);
    }
}

// ---------------------------------------------------------------------------
// K2: intra prediction + reconstruction, persistent CTU wavefront
//
// Work unit = one (picture, plane, CTB).  Units are handed out by an atomic
// ticket in an order where every unit a ticket reads from has a smaller
// ticket, so every wait ends (the producer already runs).  A unit needs only
// the bottom row of the three CTBs above it and the right column of the CTB
// to its left; each unit publishes exactly those samples as 32-bit words
// (run tag << 16 | sample) with relaxed agent-scope atomics, and consumers
// poll the words they need until the tag matches.  Per-word coherence is all
// that is required, so no fences (and no L2 write-back / invalidate) are
// needed.  Inside a unit one wave predicts the TBs one after another in
// dependency order while a second wave polls the neighbour words the next TBs
// read into the LDS tile; the static part of every TB (availability,
// substituted reference addresses, per-sample reference words) comes from
// k_prep, which runs before the wavefront on the records alone.
// ---------------------------------------------------------------------------
#include "intra.hip"

// Starts a run: zeroes the error word and the ticket counters and advances
// the border tag (never 0; border words start zeroed).
__global__ void k_begin_run(uint32_t* sync, uint32_t ncounters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        uint32_t g = (sync[1] + 1) & 0xffffu;
        sync[1] = g ? g : 1;
        sync[0] = 0;
    }
    if (i < ncounters) sync[2 + i] = 0;
}

// ---------------------------------------------------------------------------
// K3: deblocking (src/decode.rs:1774 serial order: luma V, luma H, chroma V,
// chroma H; exempt TQB/PCM cells keep their input values)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t cell_of(const DevPic& P, const uint32_t* info, int x, int y) {
    return info[P.cell_begin + (size_t)(y >> 2) * P.gw + (x >> 2)];
}

__device__ __forceinline__ bool suppressed(const DevPic& P, uint32_t inf) {
    return (inf & HPVCD_CELL_TQB) || ((P.flags & HPVCD_PF_PCM_LF_DISABLED) && (inf & HPVCD_CELL_PCM));
}

__device__ __forceinline__ hpvcd_slice_params slice_params(const DevPic& P, uint32_t inf) {
    int s = HPVCD_CELL_SLICE(inf);
    if (s < P.num_slices) return P.slices[s];
    hpvcd_slice_params d;
    d.disabled = (P.flags & HPVCD_PF_DEBLOCK_DISABLED) ? 1 : 0;
    d.beta_offset_div2 = 0;
    d.tc_offset_div2 = 0;
    d.lf_across = 1;
    return d;
}

// bS of the edge between P sample (xp, yp) and Q sample (xq, yq), with the
// slice / tile gate (src/decode.rs:1673, :1692, :1712)
__device__ __forceinline__ int edge_bs(const DevPic& P, const uint32_t* info, bool vertical, int xq, int yq) {
    if (xq >= P.w || yq >= P.h) return 0;
    if (vertical ? xq == 0 : yq == 0) return 0;
    uint32_t iq = cell_of(P, info, xq, yq);
    int bs = vertical ? HPVCD_CELL_BSV(iq) : HPVCD_CELL_BSH(iq);
    if (!bs) return 0;
    int xp = vertical ? xq - 1 : xq, yp = vertical ? yq : yq - 1;
    uint32_t ip = cell_of(P, info, xp, yp);
    int sq = HPVCD_CELL_SLICE(iq);
    bool q_across = sq < P.num_slices ? P.slices[sq].lf_across != 0 : (P.flags & HPVCD_PF_LF_ACROSS_SLICES) != 0;
    if (!q_across && HPVCD_CELL_SLICE(ip) != (uint32_t)sq) return 0;
    if ((P.flags & HPVCD_PF_TILES) && !(P.flags & HPVCD_PF_LF_ACROSS_TILES) && HPVCD_CELL_TILE(ip) != HPVCD_CELL_TILE(iq))
        return 0;
    return bs;
}

__device__ __forceinline__ int qp_at(const DevPic& P, const uint32_t* info, int px, int py) {
    int gx = px >> 2, gy = py >> 2;
    if (gx > P.gw - 1) gx = P.gw - 1;
    if (gy > P.gh - 1) gy = P.gh - 1;
    return HPVCD_CELL_QP(info[P.cell_begin + (size_t)gy * P.gw + gx]);
}

// Vector of 4 samples.
template <typename T> struct Vec4;
template <> struct Vec4<uint8_t> { using type = uchar4; };
template <> struct Vec4<uint16_t> { using type = ushort4; };

// One 4-line luma edge segment.  base points at q0 of the first line; a
// vertical edge's lines are rows (taps along the row), a horizontal edge's
// lines are columns (taps down the column).  Either way the 8 x 4 samples
// p3..q3 come in as eight aligned 4-sample vectors (edges sit on multiples of
// 8, lines on multiples of 4, rows start at multiples of 4 samples) and go
// back the same way; a thread owns p3..q3 of its segment, so rewriting the
// samples the filter leaves is safe.
template <typename T>
__device__ __forceinline__ void luma_segment_load(const T* base, bool vertical, int stride, int32_t (&s)[4][8]) {
    using V = typename Vec4<T>::type;
    if (vertical) {
#pragma unroll
        for (int l = 0; l < 4; l++) {
            const V a = *reinterpret_cast<const V*>(base + l * stride - 4), b = *reinterpret_cast<const V*>(base + l * stride);
            s[l][0] = a.x; s[l][1] = a.y; s[l][2] = a.z; s[l][3] = a.w;
            s[l][4] = b.x; s[l][5] = b.y; s[l][6] = b.z; s[l][7] = b.w;
        }
    } else {
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const V a = *reinterpret_cast<const V*>(base + (t - 4) * stride);
            s[0][t] = a.x; s[1][t] = a.y; s[2][t] = a.z; s[3][t] = a.w;
        }
    }
}

template <typename T>
__device__ void luma_segment(T* base, bool vertical, int stride, const int32_t (&s)[4][8], int beta, int tc,
                             int maxv, bool wp, bool wq) {
    // src/deblock.rs:216 luma_decision + :241 deblock_luma_segment
    using V = typename Vec4<T>::type;
    auto dpf = [&](int l) { int v = s[l][1] - 2 * s[l][2] + s[l][3]; return v < 0 ? -v : v; };
    auto dqf = [&](int l) { int v = s[l][6] - 2 * s[l][5] + s[l][4]; return v < 0 ? -v : v; };
    int dp0 = dpf(0), dp3 = dpf(3), dq0 = dqf(0), dq3 = dqf(3);
    if (dp0 + dp3 + dq0 + dq3 >= beta) return;
    auto dsam = [&](int l, int dpq) {
        int a = s[l][0] - s[l][3], b = s[l][4] - s[l][7], c = s[l][3] - s[l][4];
        a = a < 0 ? -a : a; b = b < 0 ? -b : b; c = c < 0 ? -c : c;
        return 2 * dpq < (beta >> 2) && a + b < (beta >> 3) && c < ((5 * tc + 1) >> 1);
    };
    const bool strong = dsam(0, dp0 + dq0) && dsam(3, dp3 + dq3);
    const int side_thr = (beta + (beta >> 1)) >> 3;
    const bool do_p1 = dp0 + dp3 < side_thr, do_q1 = dq0 + dq3 < side_thr;
    const int tc2 = tc >> 1;
    int32_t o[4][8];
#pragma unroll
    for (int l = 0; l < 4; l++) {
        const int32_t p3 = s[l][0], p2 = s[l][1], p1 = s[l][2], p0 = s[l][3];
        const int32_t q0 = s[l][4], q1 = s[l][5], q2 = s[l][6], q3 = s[l][7];
#pragma unroll
        for (int t = 0; t < 8; t++) o[l][t] = s[l][t];
        if (strong) {
            auto sc = [&](int32_t orig, int32_t v) { return clampi(clampi(v, orig - 2 * tc, orig + 2 * tc), 0, maxv); };
            if (wp) {
                o[l][3] = sc(p0, (p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3);
                o[l][2] = sc(p1, (p2 + p1 + p0 + q0 + 2) >> 2);
                o[l][1] = sc(p2, (2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3);
            }
            if (wq) {
                o[l][4] = sc(q0, (p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3);
                o[l][5] = sc(q1, (p0 + q0 + q1 + q2 + 2) >> 2);
                o[l][6] = sc(q2, (p0 + q0 + q1 + 3 * q2 + 2 * q3 + 4) >> 3);
            }
        } else {
            const int32_t delta0 = (9 * (q0 - p0) - 3 * (q1 - p1) + 8) >> 4;
            if ((delta0 < 0 ? -delta0 : delta0) < tc * 10) {
                const int32_t delta = clampi(delta0, -tc, tc);
                if (wp) o[l][3] = clampi(p0 + delta, 0, maxv);
                if (wq) o[l][4] = clampi(q0 - delta, 0, maxv);
                if (do_p1 && wp) o[l][2] = clampi(p1 + clampi((((p2 + p0 + 1) >> 1) - p1 + delta) >> 1, -tc2, tc2), 0, maxv);
                if (do_q1 && wq) o[l][5] = clampi(q1 + clampi((((q2 + q0 + 1) >> 1) - q1 - delta) >> 1, -tc2, tc2), 0, maxv);
            }
        }
    }
    if (vertical) {
#pragma unroll
        for (int l = 0; l < 4; l++) {
            V a, b;
            a.x = (T)o[l][0]; a.y = (T)o[l][1]; a.z = (T)o[l][2]; a.w = (T)o[l][3];
            b.x = (T)o[l][4]; b.y = (T)o[l][5]; b.z = (T)o[l][6]; b.w = (T)o[l][7];
            *reinterpret_cast<V*>(base + l * stride - 4) = a;
            *reinterpret_cast<V*>(base + l * stride) = b;
        }
    } else {
#pragma unroll
        for (int t = 1; t < 7; t++) {
            V a;
            a.x = (T)o[0][t]; a.y = (T)o[1][t]; a.z = (T)o[2][t]; a.w = (T)o[3][t];
            *reinterpret_cast<V*>(base + (t - 4) * stride) = a;
        }
    }
}

template <typename T>
__device__ __forceinline__ void deblock_luma(const DevPic* __restrict__ pics, const uint32_t* __restrict__ info,
                                             T* __restrict__ samples, int vertical, uint32_t pic_base, uint32_t bx) {
    const DevPic& P = pics[pic_base + blockIdx.y];
    if (!P.deblock_on) return;
    const int edge_max = vertical ? P.w : P.h, scan_max = vertical ? P.h : P.w;
    const int n_e = edge_max >= 12 ? (edge_max - 4) / 8 : 0;
    const int n_s = scan_max / 4;
    const uint32_t idx = bx * blockDim.x + threadIdx.x;
    if (idx >= (uint32_t)(n_e * n_s)) return;
    int e, s;
    if (vertical) { s = idx / n_e; e = idx - s * n_e; } else { e = idx / n_s; s = idx - e * n_s; }
    const int edge = 8 * (e + 1), scan = 4 * s, mid = scan + 1;
    const int xq = vertical ? edge : mid, yq = vertical ? mid : edge;
    // the segment's samples load while the edge's strength and QP are decided
    T* const seg = samples + P.plane_off[0] + (vertical ? (size_t)scan * P.w + edge : (size_t)edge * P.w + scan);
    int32_t px[4][8];
    luma_segment_load<T>(seg, vertical != 0, P.w, px);
    const int bs = edge_bs(P, info, vertical, xq, yq);
    if (!bs) return;
    const uint32_t iq = cell_of(P, info, xq, yq);
    const hpvcd_slice_params sd = slice_params(P, iq);
    if (sd.disabled) return;
    const int qp_p = vertical ? qp_at(P, info, edge - 1, mid) : qp_at(P, info, mid, edge - 1);
    const int qp_q = HPVCD_CELL_QP(iq);
    const int avg = (qp_p + qp_q + 1) >> 1;
    const int bi = clampi(avg + sd.beta_offset_div2 * 2, 0, 51);
    const int ti = clampi(avg + 2 * (bs - 1) + sd.tc_offset_div2 * 2, 0, 53);
    const int beta = kBeta[bi] << (P.bd - 8), tc = kTc[ti] << (P.bd - 8);
    if (!tc) return;
    bool wp = true, wq = true;
    if (P.suppress_any) {
        const int xp = vertical ? edge - 1 : scan, yp = vertical ? scan : edge - 1;
        wp = !suppressed(P, cell_of(P, info, xp, yp));
        wq = !suppressed(P, cell_of(P, info, vertical ? edge : scan, vertical ? scan : edge));
    }
    luma_segment<T>(seg, vertical != 0, P.w, px, beta, tc, (1 << P.bd) - 1, wp, wq);
}

template <typename T>
__device__ __forceinline__ void deblock_chroma(const DevPic* __restrict__ pics, const uint32_t* __restrict__ info,
                                               T* __restrict__ samples, int vertical, uint32_t pic_base, uint32_t bx) {
    const DevPic& P = pics[pic_base + blockIdx.y];
    if (!P.deblock_on || P.chroma_idc == 0) return;
    const int cw = P.cw, ch = P.ch;
    const int edge_max = vertical ? cw : ch, scan_max = vertical ? ch : cw;
    const int n_e = edge_max >= 10 ? (edge_max - 2) / 8 : 0;
    const int n_s = scan_max / 4;
    const uint32_t idx = bx * blockDim.x + threadIdx.x;
    if (idx >= (uint32_t)(n_e * n_s)) return;
    int e, s;
    if (vertical) { s = idx / n_e; e = idx - s * n_e; } else { e = idx / n_s; s = idx - e * n_s; }
    const int edge = 8 * (e + 1), scan = 4 * s, mid = scan + 1;
    const int sw = P.sub_w_log2, sh = P.sub_h_log2;
    const int lex = vertical ? edge << sw : mid << sw, ley = vertical ? mid << sh : edge << sh;
    const int bs = edge_bs(P, info, vertical, lex, ley);
    if (bs < 2) return;
    const int lqx = lex < P.w - 1 ? lex : P.w - 1, lqy = ley < P.h - 1 ? ley : P.h - 1;
    const hpvcd_slice_params sd = slice_params(P, cell_of(P, info, lqx, lqy));
    if (sd.disabled) return;
    int ppx, ppy;
    if (vertical) { ppx = lex > 0 ? lex - 1 : 0; if (ppx > P.w - 1) ppx = P.w - 1; ppy = lqy; }
    else { ppx = lqx; ppy = ley > 0 ? ley - 1 : 0; if (ppy > P.h - 1) ppy = P.h - 1; }
    const int avg = (qp_at(P, info, ppx, ppy) + qp_at(P, info, lqx, lqy) + 1) >> 1;
    const int maxv = (1 << P.bd_c) - 1;
    for (int pln = 0; pln < 2; pln++) {
        const int off = pln == 0 ? P.pps_cb_qp_offset : P.pps_cr_qp_offset;
        const int qpc = qpc_of(avg + off, P.chroma_idc, P.bd_c);
        const int tci = clampi(qpc + 2 + sd.tc_offset_div2 * 2, 0, 53);
        const int tcc = kTc[tci] << (P.bd_c - 8);
        if (!tcc) continue;
        T* pl = samples + P.plane_off[1 + pln];
        // p1 p0 q0 q1 of the 4 lines from aligned 4-sample vectors (edges
        // on multiples of 8, lines on multiples of 4, rows start at
        // multiples of 4 samples); a thread owns the 8 samples around its
        // edge on each line, so whole vectors go back
        using V = typename Vec4<T>::type;
        int32_t p1[4], p0[4], q0[4], q1[4];
        V pv[4], qv[4];
        T* base = vertical ? pl + (size_t)scan * cw + edge : pl + (size_t)edge * cw + scan;
        if (vertical) {
#pragma unroll
            for (int l = 0; l < 4; l++) {
                pv[l] = *reinterpret_cast<const V*>(base + l * cw - 4);
                qv[l] = *reinterpret_cast<const V*>(base + l * cw);
                p1[l] = pv[l].z; p0[l] = pv[l].w; q0[l] = qv[l].x; q1[l] = qv[l].y;
            }
        } else {
            const V a = *reinterpret_cast<const V*>(base - 2 * cw), b = *reinterpret_cast<const V*>(base - cw);
            const V c = *reinterpret_cast<const V*>(base), d = *reinterpret_cast<const V*>(base + cw);
            p1[0] = a.x; p1[1] = a.y; p1[2] = a.z; p1[3] = a.w;
            p0[0] = b.x; p0[1] = b.y; p0[2] = b.z; p0[3] = b.w;
            q0[0] = c.x; q0[1] = c.y; q0[2] = c.z; q0[3] = c.w;
            q1[0] = d.x; q1[1] = d.y; q1[2] = d.z; q1[3] = d.w;
        }
        int32_t np[4], nq[4];
        bool any = false;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int l = scan + k;
            np[k] = p0[k];
            nq[k] = q0[k];
            const int32_t delta = clampi(((q0[k] - p0[k]) * 4 + p1[k] - q1[k] + 4) >> 3, -tcc, tcc);
            if (!delta) continue;
            bool wp = true, wq = true;
            if (P.suppress_any) {
                int cpx = vertical ? edge - 1 : l, cpy = vertical ? l : edge - 1;
                int cqx = vertical ? edge : l, cqy = vertical ? l : edge;
                wp = !suppressed(P, cell_of(P, info, cpx << sw, cpy << sh));
                wq = !suppressed(P, cell_of(P, info, cqx << sw, cqy << sh));
            }
            if (wp) np[k] = clampi(p0[k] + delta, 0, maxv);
            if (wq) nq[k] = clampi(q0[k] - delta, 0, maxv);
            any = true;
        }
        if (!any) continue;
        if (vertical) {
#pragma unroll
            for (int l = 0; l < 4; l++) {
                V a = pv[l], b = qv[l];
                a.w = (T)np[l];
                b.x = (T)nq[l];
                *reinterpret_cast<V*>(base + l * cw - 4) = a;
                *reinterpret_cast<V*>(base + l * cw) = b;
            }
        } else {
            V a, b;
            a.x = (T)np[0]; a.y = (T)np[1]; a.z = (T)np[2]; a.w = (T)np[3];
            b.x = (T)nq[0]; b.y = (T)nq[1]; b.z = (T)nq[2]; b.w = (T)nq[3];
            *reinterpret_cast<V*>(base - cw) = a;
            *reinterpret_cast<V*>(base) = b;
        }
    }
}

// One pass (vertical or horizontal edges) over luma and chroma together: the
// first luma_blocks blocks take luma segments, the rest chroma (the planes'
// edges are independent; horizontal edges run after vertical ones).
template <typename T>
__global__ void __launch_bounds__(256) k_deblock(const DevPic* __restrict__ pics, const uint32_t* __restrict__ info,
                                                 T* __restrict__ samples, int vertical, uint32_t pic_base,
                                                 uint32_t luma_blocks) {
    if (blockIdx.x < luma_blocks) deblock_luma<T>(pics, info, samples, vertical, pic_base, blockIdx.x);
    else deblock_chroma<T>(pics, info, samples, vertical, pic_base, blockIdx.x - luma_blocks);
}

// ---------------------------------------------------------------------------
// K4: SAO, out of place (src/decode.rs:2258; gated EO src/sao.rs:433)
//
// One workgroup per part of a luma CTB's rows (the CTB whole, or 2 or 4
// bands of rows when the launch holds few CTBs) and per chroma plane of a CTB, so the CTB's
// SAO parameters are uniform; each thread takes quads of 4 horizontally adjacent samples
// (plane widths are multiples of 4 and planes start 16-sample aligned, so a
// quad is one aligned vector load/store).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_sao(const DevPic* __restrict__ pics, const uint32_t* __restrict__ info,
                                             const hpvcd_sao* __restrict__ sao, const T* __restrict__ src_pool,
                                             T* __restrict__ dst_pool, uint32_t pic_base, int luma_parts) {
    const DevPic& P = pics[pic_base + blockIdx.y];
    // z [0, luma_parts): parts of the luma CTB's rows; then Cb, Cr
    const int z = (int)blockIdx.z;
    const int c = z < luma_parts ? 0 : z - luma_parts + 1, part = z < luma_parts ? z : 0;
    const int nparts = c ? 1 : luma_parts;
    if (c > 0 && P.chroma_idc == 0) return;
    const int ctb_i = blockIdx.x;
    if (ctb_i >= P.ctb_cols * P.ctb_rows) return;
    const int ry = ctb_i / P.ctb_cols, rx = ctb_i - ry * P.ctb_cols;
    const int pw = c ? P.cw : P.w, ph = c ? P.ch : P.h;
    const int sw = c ? P.sub_w_log2 : 0, sh = c ? P.sub_h_log2 : 0;
    const int ctb = 1 << P.ctb_log2;
    const int xs = (rx * ctb) >> sw, ys0 = (ry * ctb) >> sh;
    const int xe = min(((rx + 1) * ctb) >> sw, pw), ye0 = min(((ry + 1) * ctb) >> sh, ph);
    const int prow = (ye0 - ys0 + nparts - 1) / nparts;
    const int ys = ys0 + part * prow, ye = min(ye0, ys + prow);
    const int qw = (xe - xs) >> 2;
    if (qw <= 0 || ye <= ys) return;
    const T* src = src_pool + P.plane_off[c];
    T* dst = dst_pool + P.plane_off[c];
    using V = typename Vec4<T>::type;
    // quad (qx, qy) of the thread, advanced by the block size each pass
    // without a division per pass
    const int rows = ye - ys, sy = (int)blockDim.x / qw, sx = (int)blockDim.x - sy * qw;
    const int qy0 = (int)threadIdx.x / qw, qx0 = (int)threadIdx.x - qy0 * qw;
    const bool on = c == 0 ? (P.flags & HPVCD_PF_SAO_LUMA) != 0 : ((P.flags & HPVCD_PF_SAO_CHROMA) != 0 && P.cw && P.ch);
    const hpvcd_sao& S = sao[P.ctb_begin + ctb_i];
    const int type = on ? S.type_idx[c] : 0;
    if (type == 0) {
        for (int qy = qy0, qx = qx0; qy < rows;) {
            const int y = ys + qy, x = xs + qx * 4;
            *reinterpret_cast<V*>(dst + (size_t)y * pw + x) = *reinterpret_cast<const V*>(src + (size_t)y * pw + x);
            qx += sx;
            qy += sy;
            if (qx >= qw) { qx -= qw; qy++; }
        }
        return;
    }
    const int bd = c ? P.bd_c : P.bd;
    const int32_t maxv = (1 << bd) - 1;
    const int32_t o0 = S.offsets[c][0], o1 = S.offsets[c][1], o2 = S.offsets[c][2], o3 = S.offsets[c][3];
    const int bp = S.band_pos[c];
    const int ec = S.eo_class[c];
    const int dx = ec == 1 ? 0 : 1, dy = ec == 0 ? 0 : (ec == 3 ? -1 : 1);
    const bool across_slices = P.sao_across_slices != 0;
    const bool across_tiles = !(P.flags & HPVCD_PF_TILES) || (P.flags & HPVCD_PF_LF_ACROSS_TILES);
    const bool gated = P.suppress_any || !across_slices || !across_tiles;
    for (int qy = qy0, qx = qx0; qy < rows;) {
        const int y = ys + qy, x = xs + qx * 4;
        qx += sx;
        qy += sy;
        if (qx >= qw) { qx -= qw; qy++; }
        const V sv = *reinterpret_cast<const V*>(src + (size_t)y * pw + x);
        const int32_t s[4] = {sv.x, sv.y, sv.z, sv.w};
        int32_t out[4] = {s[0], s[1], s[2], s[3]};
        if (type == 1) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint8_t rel = (uint8_t)((uint8_t)(s[k] >> (bd - 5)) - bp);
                const int32_t o = rel == 0 ? o0 : rel == 1 ? o1 : rel == 2 ? o2 : o3;
                if (rel < 4) out[k] = clampi(s[k] + o, 0, maxv);
            }
        } else {
            const int y1 = y + dy, y2 = y - dy;
            const bool rows_ok = y1 >= 0 && y1 < ph && y2 >= 0 && y2 < ph;
            const T* r1 = src + (size_t)(rows_ok ? y1 : y) * pw;
            const T* r2 = src + (size_t)(rows_ok ? y2 : y) * pw;
            if (!gated) {
                // no gates: branch-free, neighbour columns clamped into the
                // row (a sample whose neighbour lies outside keeps its value)
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int xk = x + k, x1 = xk + dx, x2 = xk - dx;
                    const bool ok = rows_ok && x2 >= 0 && x1 < pw;
                    const int32_t n1 = r1[min(x1, pw - 1)], n2 = r2[max(x2, 0)];
                    const int ei = ((s[k] > n1) - (s[k] < n1)) + ((s[k] > n2) - (s[k] < n2)) + 2;
                    const int32_t o = ei == 0 ? o0 : ei == 1 ? o1 : ei == 3 ? o2 : ei == 4 ? o3 : 0;
                    out[k] = ok ? clampi(s[k] + o, 0, maxv) : s[k];
                }
            } else
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int xk = x + k, x1 = xk + dx, x2 = xk - dx;
                bool ok = rows_ok && x1 >= 0 && x1 < pw && x2 >= 0 && x2 < pw;
                if (ok && gated) {
                    // src/sao.rs:378 luma_neighbor_ok (center exemption, slice and tile gates)
                    const uint32_t ic = cell_of(P, info, xk << sw, y << sh);
                    if (suppressed(P, ic)) ok = false;
                    if (ok && (!across_slices || !across_tiles)) {
                        const uint32_t i1 = cell_of(P, info, x1 << sw, y1 << sh), i2 = cell_of(P, info, x2 << sw, y2 << sh);
                        if (!across_slices && (HPVCD_CELL_SLICE(i1) != HPVCD_CELL_SLICE(ic) || HPVCD_CELL_SLICE(i2) != HPVCD_CELL_SLICE(ic))) ok = false;
                        if (!across_tiles && (HPVCD_CELL_TILE(i1) != HPVCD_CELL_TILE(ic) || HPVCD_CELL_TILE(i2) != HPVCD_CELL_TILE(ic))) ok = false;
                    }
                }
                if (ok) {
                    const int32_t n1 = r1[x1], n2 = r2[x2];
                    const int ei = ((s[k] > n1) - (s[k] < n1)) + ((s[k] > n2) - (s[k] < n2)) + 2;
                    const int32_t o = ei == 0 ? o0 : ei == 1 ? o1 : ei == 3 ? o2 : ei == 4 ? o3 : 0;
                    if (o) out[k] = clampi(s[k] + o, 0, maxv);
                }
            }
        }
        if (P.suppress_any) {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (suppressed(P, cell_of(P, info, (x + k) << sw, y << sh))) out[k] = s[k];
        }
        V ov;
        ov.x = (T)out[0];
        ov.y = (T)out[1];
        ov.z = (T)out[2];
        ov.w = (T)out[3];
        *reinterpret_cast<V*>(dst + (size_t)y * pw + x) = ov;
    }
}

// ---------------------------------------------------------------------------
// K5: YUV -> RGB with crop window and grid placement (src/yuv.rs:187-396)
//
// One thread per 4 horizontally adjacent output pixels; a workgroup covers
// 64 such quads of blockDim / 64 output rows (grid: quads of a row / 64,
// rows / (blockDim / 64), pictures).  The Q13 products fit 32 bits for bit depths up to 12
// (|coefficient| < 2^15, sample < 2^12).
// ---------------------------------------------------------------------------
// Whether a picture's output orientation moves its pixels (the oriented
// store of k_rgb); the other orientations store the canvas as it is.
__host__ __device__ __forceinline__ bool orient_moves(int o) {
    return o >= HPVCD_ORIENT_FLIP_H && o != HPVCD_ORIENT_TRANSPOSE && o != HPVCD_ORIENT_TRANSVERSE;
}

// ORIENTED: the launch holds a picture whose orientation moves pixels (the
// per-pixel oriented store is compiled in only then).
template <typename T, bool ORIENTED>
__global__ void __launch_bounds__(256) k_rgb(const DevPic* __restrict__ pics, const T* __restrict__ pool,
                                             uint8_t* __restrict__ out_pool, uint32_t pic_base) {
    const DevPic& P = pics[pic_base + blockIdx.z];
    if (HPVCD_OUT_IS_YUV(P.out_kind)) return;  // raw planes: k_yuv
    const int qx = (P.copy_w + 3) >> 2;
    const int xq = (int)(blockIdx.x * 64 + (threadIdx.x & 63));
    const int yy = (int)(blockIdx.y * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (xq >= qx || yy >= P.copy_h) return;
    const int xx0 = xq * 4;
    const int npx = min(4, P.copy_w - xx0);
    const int w = P.w, h = P.h;
    const int x0 = P.crop_left < w - 1 ? P.crop_left : w - 1, y0 = P.crop_top < h - 1 ? P.crop_top : h - 1;
    const int src_y = y0 + yy;
    const int y_row = src_y < h - 1 ? src_y : h - 1;
    const int bd = P.bd;
    const int32_t scale = 1 << (bd - 8);
    const bool full = (P.flags & HPVCD_PF_FULL_RANGE) != 0;
    const int32_t ky = full ? 8192 : 9539, kc = full ? 8192 : 9326;
    const int32_t y_black = full ? 0 : 16 * scale, neutral = 128 * scale;
    const bool mono = P.chroma_idc == 0;
    // luma outputs hold one channel; RGB outputs three (a 4:0:0 picture as gray,
    // the decode_rgb8 expansion of src/lib.rs:1978)
    const int chn = (P.out_kind == HPVCD_OUT_LUMA8 || P.out_kind == HPVCD_OUT_LUMA16) ? 1 : 3;
    const int32_t cmax = bd == 8 ? 255 : (1 << bd) - 1;
    const T* Yp = pool + P.plane_off[0] + (size_t)y_row * w;
    const int cw = P.cw, ch = P.ch;
    const int crow = (src_y >> P.sub_h_log2) < ch - 1 ? (src_y >> P.sub_h_log2) : ch - 1;
    const T* Up = pool + P.plane_off[1] + (size_t)crow * cw;
    const T* Vp = pool + P.plane_off[2] + (size_t)crow * cw;
    int32_t crr = 0, cbg = 0, crg = 0, cbb = 0;
    if (!mono && P.matrix != 0 && P.matrix != 8) {
        int32_t m0, m1, m2, m3;
        if (P.matrix == 5 || P.matrix == 6) { m0 = 11485; m1 = -2819; m2 = -5851; m3 = 14516; }
        else if (P.matrix == 9) { m0 = 12080; m1 = -1348; m2 = -4681; m3 = 17546; }
        else { m0 = 12901; m1 = -1534; m2 = -3835; m3 = 15201; }
        crr = (m0 * kc + 4096) >> 13; cbg = (m1 * kc + 4096) >> 13;
        crg = (m2 * kc + 4096) >> 13; cbb = (m3 * kc + 4096) >> 13;
    }
    // samples of the quad: aligned vector loads inside the picture (rows
    // start at multiples of 4 samples), one sample at a time at its edges
    using V4 = typename Vec4<T>::type;
    const int sx0 = x0 + xx0, csw = P.sub_w_log2;
    const bool vec = npx == 4 && (sx0 & 3) == 0 && sx0 + 3 <= w - 1;
    int32_t Ls[4], Us[4] = {0, 0, 0, 0}, Vs[4] = {0, 0, 0, 0};
    if (vec) {
        const V4 q = *reinterpret_cast<const V4*>(Yp + sx0);
        Ls[0] = q.x; Ls[1] = q.y; Ls[2] = q.z; Ls[3] = q.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int src_x = sx0 + min(k, npx - 1);
            Ls[k] = Yp[src_x < w - 1 ? src_x : w - 1];
        }
    }
    if (!mono) {
        const int c0 = sx0 >> csw;
        if (vec && csw == 1 && c0 + 1 <= cw - 1) {
            Us[0] = Us[1] = Up[c0];
            Us[2] = Us[3] = Up[c0 + 1];
            Vs[0] = Vs[1] = Vp[c0];
            Vs[2] = Vs[3] = Vp[c0 + 1];
        } else if (vec && csw == 0 && c0 + 3 <= cw - 1) {
            const V4 u = *reinterpret_cast<const V4*>(Up + c0), vv = *reinterpret_cast<const V4*>(Vp + c0);
            Us[0] = u.x; Us[1] = u.y; Us[2] = u.z; Us[3] = u.w;
            Vs[0] = vv.x; Vs[1] = vv.y; Vs[2] = vv.z; Vs[3] = vv.w;
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int src_x = sx0 + min(k, npx - 1);
                const int cc = (src_x >> csw) < cw - 1 ? (src_x >> csw) : cw - 1;
                Us[k] = Up[cc];
                Vs[k] = Vp[cc];
            }
        }
    }
    int32_t v[4][3];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int32_t L = Ls[k];
        if (mono) {
            v[k][0] = v[k][1] = v[k][2] = (ky * (L - y_black) + 4096) >> 13;
            continue;
        }
        const int32_t U = Us[k], Vv = Vs[k];
        if (P.matrix == 0) {
            if (full) { v[k][0] = Vv; v[k][1] = L; v[k][2] = U; }
            else {
                v[k][0] = (kc * (Vv - y_black) + 4096) >> 13;
                v[k][1] = (ky * (L - y_black) + 4096) >> 13;
                v[k][2] = (kc * (U - y_black) + 4096) >> 13;
            }
        } else if (P.matrix == 8) {
            const int32_t cg = U - neutral, co = Vv - neutral, t = L - cg;
            v[k][0] = t + co; v[k][1] = L + cg; v[k][2] = t - co;
        } else {
            const int32_t yt = ky * (L - y_black), cbc = U - neutral, crc = Vv - neutral;
            v[k][0] = (yt + crr * crc + 4096) >> 13;
            v[k][1] = (yt + cbg * cbc + crg * crc + 4096) >> 13;
            v[k][2] = (yt + cbb * cbc + 4096) >> 13;
        }
    }
    const bool wide = P.out_kind == HPVCD_OUT_RGB16 || P.out_kind == HPVCD_OUT_LUMA16;
    const int orient = P.orient;
    if (ORIENTED && orient_moves(orient)) {
        // oriented image (src/lib.rs:2065 rotate_luma, :2116 rotate_buf): each
        // pixel of the canvas goes to its place in the oriented buffer
        const int W = P.out_width, H = P.out_height, Y = P.dst_y + yy;
        const int sh = (wide || bd == 8) ? 0 : bd - 8;
        for (int k = 0; k < npx; k++) {
            const int X = P.dst_x + xx0 + k;
            size_t idx;
            if (orient == HPVCD_ORIENT_FLIP_H) idx = (size_t)Y * W + (W - 1 - X);
            else if (orient == HPVCD_ORIENT_ROTATE_180) idx = (size_t)(H - 1 - Y) * W + (W - 1 - X);
            else if (orient == HPVCD_ORIENT_FLIP_V) idx = (size_t)(H - 1 - Y) * W + X;
            else if (orient == HPVCD_ORIENT_ROTATE_90) idx = (size_t)X * H + (H - 1 - Y);
            else idx = (size_t)(W - 1 - X) * H + Y;
            for (int j = 0; j < chn; j++) {
                const uint16_t sv = (uint16_t)clampi(v[k][j], 0, cmax);
                if (wide) reinterpret_cast<uint16_t*>(out_pool + P.out_off)[idx * chn + j] = sv;
                else out_pool[P.out_off + idx * chn + j] = (uint8_t)(sv >> sh);
            }
        }
        return;
    }
    const size_t o = ((size_t)(P.dst_y + yy) * P.out_width + (P.dst_x + xx0)) * chn;
    if (wide) {
        uint16_t* d = reinterpret_cast<uint16_t*>(out_pool + P.out_off) + o;
        uint16_t b[12];
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int j = 0; j < 3; j++) b[k * 3 + j] = (uint16_t)clampi(v[k][j], 0, cmax);
        if (chn == 3 && npx == 4 && (((uintptr_t)d) & 7) == 0) {
            uint2* d64 = reinterpret_cast<uint2*>(d);
#pragma unroll
            for (int m = 0; m < 3; m++)
                d64[m] = make_uint2((uint32_t)b[4 * m] | ((uint32_t)b[4 * m + 1] << 16),
                                    (uint32_t)b[4 * m + 2] | ((uint32_t)b[4 * m + 3] << 16));
        } else {
            for (int k = 0; k < npx; k++)
                for (int j = 0; j < chn; j++) d[k * chn + j] = b[k * 3 + j];
        }
    } else {
        uint8_t* d = out_pool + P.out_off + o;
        const int sh = bd == 8 ? 0 : bd - 8;
        uint8_t b[12];
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int j = 0; j < 3; j++) b[k * 3 + j] = (uint8_t)((uint16_t)clampi(v[k][j], 0, cmax) >> sh);
        if (chn == 3 && npx == 4 && (((uintptr_t)d) & 3) == 0) {
            uint32_t* d32 = reinterpret_cast<uint32_t*>(d);
#pragma unroll
            for (int m = 0; m < 3; m++)
                d32[m] = (uint32_t)b[4 * m] | ((uint32_t)b[4 * m + 1] << 8) | ((uint32_t)b[4 * m + 2] << 16) |
                         ((uint32_t)b[4 * m + 3] << 24);
        } else {
            for (int k = 0; k < npx; k++)
                for (int j = 0; j < chn; j++) d[k * chn + j] = b[k * 3 + j];
        }
    }
}

// Rows of a picture's raw-plane copy (src/lib.rs:143 plane_window, :882
// stitch_native_yuv_band): its luma rows, then its Cb and Cr rows.  The
// chroma copy of a grid tile lands at (col * tcw, row * tch), tcw =
// ceil(max(win_w, copy_w) / sub_w), clipped to the output's chroma planes.
struct YuvCopy {
    int cdx, cdy, ccw, cch, cwo, cho;
};
__host__ __device__ __forceinline__ YuvCopy yuv_copy(const DevPic& P) {
    YuvCopy c{0, 0, 0, 0, 0, 0};
    if (P.chroma_idc == 0 || P.win_w <= 0 || P.win_h <= 0) return c;
    const int sw = 1 << P.sub_w_log2, sh = 1 << P.sub_h_log2;
    c.cwo = (P.out_width + sw - 1) / sw;
    c.cho = (P.out_height + sh - 1) / sh;
    const int tcw = ((P.win_w > P.copy_w ? P.win_w : P.copy_w) + sw - 1) / sw;
    const int tch = ((P.win_h > P.copy_h ? P.win_h : P.copy_h) + sh - 1) / sh;
    c.cdx = (P.dst_x / P.win_w) * tcw;
    c.cdy = (P.dst_y / P.win_h) * tch;
    c.ccw = c.cwo - c.cdx < tcw ? c.cwo - c.cdx : tcw;
    c.cch = c.cho - c.cdy < tch ? c.cho - c.cdy : tch;
    if (c.ccw < 0 || P.copy_w <= 0) c.ccw = 0;
    if (c.cch < 0 || P.copy_h <= 0) c.cch = 0;
    return c;
}

// K5 for raw planes (Decoder::decode_yuv, src/lib.rs:762): coded sample
// values copied to planar Y, Cb, Cr outputs, source coordinates clamped to
// the coded plane; 4 samples of a row per thread (grid: quads of a row / 64,
// rows / (blockDim / 64), pictures).
template <typename T>
__global__ void __launch_bounds__(256) k_yuv(const DevPic* __restrict__ pics, const T* __restrict__ pool,
                                             uint8_t* __restrict__ out_pool, uint32_t pic_base) {
    const DevPic& P = pics[pic_base + blockIdx.z];
    if (!HPVCD_OUT_IS_YUV(P.out_kind)) return;
    const int xq = (int)(blockIdx.x * 64 + (threadIdx.x & 63));
    const int r = (int)(blockIdx.y * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const YuvCopy c = yuv_copy(P);
    int plane, y, n, dx, dy, sw_, sh_, x0, y0, ow;
    size_t obase;
    const size_t luma_px = (size_t)P.out_width * P.out_height;
    if (r < P.copy_h) {
        plane = 0; y = r; n = P.copy_w; dx = P.dst_x; dy = P.dst_y; sw_ = P.w; sh_ = P.h; ow = P.out_width;
        x0 = P.crop_left < P.w - 1 ? P.crop_left : P.w - 1;
        y0 = P.crop_top < P.h - 1 ? P.crop_top : P.h - 1;
        obase = 0;
    } else if (r < P.copy_h + 2 * c.cch) {
        const int k = (r - P.copy_h) / c.cch;
        plane = 1 + k; y = r - P.copy_h - k * c.cch; n = c.ccw; dx = c.cdx; dy = c.cdy; sw_ = P.cw; sh_ = P.ch;
        ow = c.cwo;
        const int cl = P.crop_left >> P.sub_w_log2, ct = P.crop_top >> P.sub_h_log2;
        x0 = cl < P.cw - 1 ? cl : P.cw - 1;
        y0 = ct < P.ch - 1 ? ct : P.ch - 1;
        obase = luma_px + (size_t)k * c.cwo * c.cho;
    } else {
        return;
    }
    const int x = xq * 4;
    if (x >= n) return;
    const int sy = y0 + y < sh_ - 1 ? y0 + y : sh_ - 1;
    const T* src = pool + P.plane_off[plane] + (size_t)sy * sw_;
    T* dst = reinterpret_cast<T*>(out_pool + P.out_off) + obase + (size_t)(dy + y) * ow + dx + x;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (x + k < n) {
            const int sx = x0 + x + k < sw_ - 1 ? x0 + x + k : sw_ - 1;
            dst[k] = src[sx];
        }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// explicit instantiations (sample type x residual type)
// ---------------------------------------------------------------------------
#define HPVCD_INST_RES1(R, L, B) \
    template __global__ void k_residual<R, L, B>(const DevResJob*, uint32_t, const uint32_t*, const uint8_t*, R*);
#define HPVCD_INST_RES(R)                                                                            \
    HPVCD_INST_RES1(R, 2, 64) HPVCD_INST_RES1(R, 2, 128) HPVCD_INST_RES1(R, 2, 256)                 \
    HPVCD_INST_RES1(R, 3, 64) HPVCD_INST_RES1(R, 3, 128) HPVCD_INST_RES1(R, 3, 256)                 \
    HPVCD_INST_RES1(R, 4, 64) HPVCD_INST_RES1(R, 4, 128) HPVCD_INST_RES1(R, 4, 256)                 \
    HPVCD_INST_RES1(R, 5, 64) HPVCD_INST_RES1(R, 5, 128) HPVCD_INST_RES1(R, 5, 256)
HPVCD_INST_RES(int16_t)
HPVCD_INST_RES(int32_t)

#define HPVCD_INST_INTRA(T, R)                                                                                   \
    template __global__ void k_intra<T, R>(const DevPic*, const DevUnit*, const uint32_t*, uint32_t, const uint4*,  \
                                           const uint16_t*, const uint32_t*, T*, const R*, uint32_t*, uint32_t*,   \
                                           const uint32_t*, uint32_t*);
HPVCD_INST_INTRA(uint8_t, int16_t)
HPVCD_INST_INTRA(uint8_t, int32_t)
HPVCD_INST_INTRA(uint16_t, int16_t)
HPVCD_INST_INTRA(uint16_t, int32_t)

#define HPVCD_INST_T(T)                                                                                          \
    template __global__ void k_deblock<T>(const DevPic*, const uint32_t*, T*, int, uint32_t, uint32_t);          \
    template __global__ void k_sao<T>(const DevPic*, const uint32_t*, const hpvcd_sao*, const T*, T*, uint32_t, int); \
    template __global__ void k_rgb<T, false>(const DevPic*, const T*, uint8_t*, uint32_t);                       \
    template __global__ void k_rgb<T, true>(const DevPic*, const T*, uint8_t*, uint32_t);                        \
    template __global__ void k_yuv<T>(const DevPic*, const T*, uint8_t*, uint32_t);
HPVCD_INST_T(uint8_t)
HPVCD_INST_T(uint16_t)

}  // namespace hpvcd
