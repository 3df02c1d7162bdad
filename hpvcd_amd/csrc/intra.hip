// intra.hip — K2: availability masks and the intra prediction wavefront
// (included by kernels.hip inside namespace hpvcd).

constexpr int kMaxCtb = 64;
// Tile layout: the row above the unit (x0-1 .. x0+2*ctb-1, padded), then 64
// body rows, each the left neighbour followed by the unit's samples.  The
// padding puts every body row's first sample on an 8-byte boundary, and the
// row stride of 72 samples (36 banks) moves each row 4 banks on, so the
// 64 samples of an 8 x 8 TB, or of four 4 x 4 TBs of one 8 x 8 block, fall
// in distinct banks.
constexpr int kTopW = 2 * kMaxCtb + 3;            // row above the unit: x0-1 .. x0+2*ctb-1 (+2 unused)
constexpr int kRowW = kMaxCtb + 8;                // body row stride (left neighbour at the sample before)
static_assert((kTopW + 1) % 4 == 0 && kRowW % 4 == 0, "body rows start on 8-byte boundaries");
constexpr int kNeutral = kTopW + kMaxCtb * kRowW; // tile slot holding 1 << (bitDepth - 1)
constexpr int kZero = kNeutral + 1;               // tile slot holding 0: the I_PCM predictor
constexpr int kSink = kNeutral + 2;               // write-only slot of the lanes of a step that hold no sample
constexpr int kTileWords = kNeutral + 3;
constexpr int kFStride = 132;                     // filtered references of one TB (4 * 32 + 1)
constexpr int kFBase = kTileWords;                // the filtered reference row follows the tile
constexpr int kMaxUnitTbs = 256;
constexpr int kMaxUnitRefs = 256 * 17;            // reference entries of one unit (all 4x4 TBs is the maximum)
constexpr int kBorderWords = 2 * kMaxCtb;         // bottom row + right column per unit
constexpr int kIntraThreads = 128;               // wave 0 predicts, wave 1 fetches neighbour border samples
constexpr uint32_t kSpinLimit = 1u << 19;
constexpr int kCopyBatch = 16;                   // global loads in flight per thread when staging a unit
constexpr int kFetchAhead = 8;                   // TBs ahead of the prediction wave whose neighbour words are polled

constexpr int kLdsSamples = kFBase + kFStride;

// prediction classes of the TB descriptor
enum { kClsAngular = 0, kClsHv = 1, kClsPlanar = 2, kClsDc = 3 };

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
}

__device__ __forceinline__ uint32_t border_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void border_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits for a tagged border word and returns its sample.
__device__ __forceinline__ uint16_t border_wait(const uint32_t* p, uint32_t tag, uint32_t* err) {
    uint32_t spins = 0;
    uint32_t v = border_load(p);
    while ((v >> 16) != tag) {
        ++spins;
        if (spins > kSpinLimit || ((spins & 1023u) == 0 && border_load(err) != 0)) {
            atomicOr(err, 1u);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        v = border_load(p);
    }
    return (uint16_t)(v & 0xffffu);
}

// Sum over the 64 lanes of a wave with DPP row shifts and row broadcasts
// (no LDS traffic); the result is returned to every lane.
__device__ __forceinline__ int32_t wave_sum(int32_t v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ int32_t mul24(int32_t a, int32_t b) { return __mul24(a, b); }




// Inclusive prefix maximum over the 64 lanes of a wave (DPP row shifts and
// row broadcasts).
__device__ __forceinline__ uint32_t wave_prefix_max(uint32_t v) {
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true));   // row_shr:1
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true));   // row_shr:2
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true));   // row_shr:4
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true));   // row_shr:8
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return v;
}

// Static shape of one TB's prediction: everything that depends on the
// records only (mode, size, picture flags), src/intra.rs:151 for the filter
// decision.
struct TbShape {
    int n, log2n, mode, cls, angle, inv;
    bool filt, edge_on, vertical;
};

__device__ __forceinline__ TbShape tb_shape(int log2n, int mode, int plane, uint32_t pflags, int chroma_idc,
                                            const int16_t* angles, const int16_t* inv_angles) {
    TbShape t;
    t.log2n = log2n;
    t.n = 1 << log2n;
    t.mode = mode;
    t.cls = mode == 0 ? kClsPlanar : mode == 1 ? kClsDc : (mode == 10 || mode == 26) ? kClsHv : kClsAngular;
    t.angle = mode >= 2 ? angles[mode - 2] : 0;
    t.inv = (mode >= 11 && mode <= 25) ? inv_angles[mode - 11] : 0;
    t.vertical = mode >= 18;
    const bool smoothing = !(pflags & HPVCD_PF_INTRA_SMOOTHING_DISABLED);
    const bool filt_plane = plane == 0 || chroma_idc == 3;
    t.filt = false;
    if (smoothing && filt_plane && t.n != 4 && mode != 1) {
        const int d26 = mode > 26 ? mode - 26 : 26 - mode, d10 = mode > 10 ? mode - 10 : 10 - mode;
        const int min_dist = d26 < d10 ? d26 : d10;
        const int thresh = t.n == 8 ? 7 : (t.n == 16 ? 1 : (t.n == 32 ? 0 : 8));
        t.filt = mode == 0 || min_dist > thresh;
    }
    t.edge_on = plane == 0 && !(pflags & HPVCD_PF_INTRA_BOUNDARY_FILTER_OFF) && t.n < 32;
    return t;
}

// Tile address of the substituted reference at s-index k (src/intra.rs:84):
// s 0..2n-1 runs up the left column, 2n is the corner, 2n+1..4n runs along the
// row above; an unavailable sample takes the nearest available one below it
// in s order, or the first available one, or the neutral value.
__device__ __forceinline__ int ref_tile_addr(int k, uint64_t amask, int n, int x0, int y0, int lgv, int lgh) {
    const int n2 = 2 * n, ngl = n2 >> lgv;
    const int b = k < n2 ? k >> lgv : (k == n2 ? ngl : ngl + 1 + ((k - n2 - 1) >> lgh));
    int src = k;
    if (!((amask >> b) & 1ull)) {
        const uint64_t below = amask & ((1ull << b) - 1ull);
        if (below) {
            const int bb = 63 - __builtin_clzll(below);
            src = bb < ngl ? ((bb + 1) << lgv) - 1 : (bb == ngl ? n2 : n2 + ((bb - ngl) << lgh));
        } else if (amask) {
            const int fb = __builtin_ctzll(amask);
            src = fb < ngl ? fb << lgv : (fb == ngl ? n2 : n2 + 1 + ((fb - ngl - 1) << lgh));
        } else {
            return kNeutral;
        }
    }
    if (src < n2) return kTopW + (y0 + n2 - 1 - src) * kRowW + x0;
    return y0 == 0 ? x0 + src - n2 : kTopW + (y0 - 1) * kRowW + x0 + src - n2;
}

// s-index of the extended main reference ref[k], k in [-n, 2n], of an
// angular mode (negative k projects the side reference with invAngle).
__device__ __forceinline__ int ext_sidx(int k, int n, int inv, bool vertical) {
    int off = k;
    if (k < 0) {
        const int idx = (mul24(k, inv) + 128) >> 8;
        off = -(idx > 2 * n ? 2 * n : (idx < 0 ? 0 : idx));
    }
    return vertical ? 2 * n + off : 2 * n - off;
}

// Per-sample reference word: a0 [0,13), a1 [13,26), [26,31) the angular
// fraction or the DC edge kind, bit 31 the H/V edge flag.  a0/a1 are tile
// addresses of substituted references (M, by s-index), or s-indices into the
// TB's filtered reference row when the references are filtered (M null).
__device__ __forceinline__ uint32_t sample_word(const TbShape& t, int x, int y, const uint16_t* M) {
    const int n = t.n, n2 = 2 * n;
    auto A = [&](int s) -> uint32_t { return M ? (uint32_t)M[s] : (uint32_t)s; };
    if (t.cls == kClsAngular) {
        const int a = t.vertical ? y : x, c = t.vertical ? x : y;
        const int pos = mul24(a + 1, t.angle), ii = pos >> 5, frac = pos & 31;
        const int k0 = c + ii + 1;
        const uint32_t a0 = A(ext_sidx(k0, n, t.inv, t.vertical));
        const uint32_t a1 = frac ? A(ext_sidx(k0 + 1, n, t.inv, t.vertical)) : a0;
        return a0 | (a1 << 13) | ((uint32_t)frac << 26);
    }
    if (t.cls == kClsHv) {
        const bool ver = t.mode == 26;
        const int s0 = ver ? n2 + 1 + x : n2 - 1 - y, s1 = ver ? n2 - 1 - y : n2 + 1 + x;
        const bool e = t.edge_on && (ver ? x == 0 : y == 0);
        return A(s0) | (A(s1) << 13) | (e ? 1u << 31 : 0u);
    }
    if (t.cls == kClsPlanar) return A(n2 - 1 - y) | (A(n2 + 1 + x) << 13);
    const uint32_t kind = !t.edge_on ? 0u : (x == 0 && y == 0 ? 2u : ((x == 0 || y == 0) ? 1u : 0u));
    const int s0 = kind == 2 ? n2 - 1 : (y == 0 ? n2 + 1 + x : n2 - 1 - y);
    return A(s0) | (A(n2 + 1) << 13) | (kind << 26);
}

// Static preparation of every TB for the wavefront: one block per unit.
// Everything here depends on the records only, so it runs fully parallel
// before the wavefront:
//  * reference availability (src/decode.rs:4651 luma_avail, :4669
//    chroma_avail) as a mask over 4-sample groups - bits [0, ngl) the left
//    groups from the bottom up, bit ngl the corner, then the above groups;
//    availability is constant within a group because a group never spans
//    two 4x4 luma cells;
//  * the substituted reference map (tile address per s-index, src/intra.rs:84);
//  * the per-sample reference words;
//  * the TB descriptor, with the neighbour samples the TB reads.
// The unit's window of 4x4 luma cells (one cell left/above, two CTB widths
// right/below) is staged in LDS first, so availability reads no global memory.
constexpr int kWin = 2 * kMaxCtb / 4 + 1;

struct PrepLds {
    uint32_t ord[kWin * kWin];
    uint32_t inf[kWin * kWin];
    PrepTb tb[kMaxUnitTbs];
    uint32_t ro[kMaxUnitTbs];
    uint16_t rmap[kMaxUnitRefs];
    uint16_t job[kMaxUnitTbs];       // TBs by size class (4x4, 8x8, larger), each by prediction class
    uint32_t cnt[12];                // TBs per (size class, prediction class)
};

constexpr int kPrepThreads = 256;

__global__ void __launch_bounds__(kPrepThreads) k_prep(const DevPic* __restrict__ pics, const DevUnit* __restrict__ units,
                                                       const uint32_t* __restrict__ prep_list,
                                                       const PrepTb* __restrict__ unit_recs,
                                                       const uint32_t* __restrict__ unit_refoff,
                                                       const uint32_t* __restrict__ cell_order,
                                                       const uint32_t* __restrict__ cell_info, uint4* __restrict__ desc_g,
                                                       uint16_t* __restrict__ rmap_g, uint32_t* __restrict__ smap_g) {
    __shared__ PrepLds L;
    __shared__ int16_t angles[33], inv_angles[15];
    const int tid = threadIdx.x, wave = uni(tid >> 6), lane = tid & 63;
    constexpr int kWaves = kPrepThreads / 64;
    const DevUnit U = units[prep_list[blockIdx.x]];
    const DevPic* Pp = pics + uni(U.pic);
    // picture fields the loops use, held in registers
    const int plane = uni((int)U.plane);
    const int ctb_log2 = Pp->ctb_log2, chroma_idc = Pp->chroma_idc, num_slices = Pp->num_slices;
    const int sw = plane ? Pp->sub_w_log2 : 0, sh = plane ? Pp->sub_h_log2 : 0;
    const int lw = Pp->w, lh = Pp->h, gw = Pp->gw, gh_cells = Pp->gh;
    const int pw = plane ? Pp->cw : lw, ph = plane ? Pp->ch : lh;
    const uint32_t pflags = Pp->flags, cell_begin = Pp->cell_begin;
    const int ctb = 1 << ctb_log2, ctbw = ctb >> sw, ctbh = ctb >> sh;
    const int ux0 = U.cx * ctbw, uy0 = U.cy * ctbh;
    const int gx0 = ((int)U.cx << (ctb_log2 - 2)) - 1, gy0 = ((int)U.cy << (ctb_log2 - 2)) - 1;
    const int ntb = U.ntb;
    const uint32_t rbase = U.rbase, sbase = U.sbase, tb_list = U.tb_list;
    const bool need_info = (pflags & (HPVCD_PF_TILES | HPVCD_PF_CONSTRAINED_INTRA)) || num_slices > 1;
    const uint32_t* ord_g = cell_order + cell_begin;
    const uint32_t* inf_g = cell_info + cell_begin;
    if (tid < 33) angles[tid] = kIntraAngle[tid];
    if (tid < 15) inv_angles[tid] = kInvAngle[tid];
    // The window's cells that can precede the unit's TBs in decode order:
    // the row above (the CTBs above and above-right), the column to the left
    // (the left CTB and the corner) and the CTB itself.  Every other cell
    // belongs to a CTB decoded later (to the right, or a CTB row below) and is
    // stored as never available without a load.
    const int cc = ctb >> 2;  // the CTB's width in 4x4 luma cells
    for (int c = tid; c < kWin * kWin; c += kPrepThreads) L.ord[c] = 0xffffffffu;
    __syncthreads();
    {
        const int nneed = kWin + cc + cc * cc;
        constexpr int kNeedPer = (kWin + 16 + 16 * 16 + kPrepThreads - 1) / kPrepThreads;
        const int lcc = ctb_log2 - 2;
        uint32_t vo[kNeedPer], vi[kNeedPer];
        int wc[kNeedPer];
        bool win[kNeedPer];
#pragma unroll
        for (int k = 0; k < kNeedPer; k++) {
            const int t = tid + k * kPrepThreads;
            int wx, wy;
            if (t < kWin) { wx = t; wy = 0; }
            else if (t < kWin + cc) { wx = 0; wy = t - kWin + 1; }
            else { const int i = t - kWin - cc; wx = 1 + (i & (cc - 1)); wy = 1 + (i >> lcc); }
            const int gx = gx0 + wx, gy = gy0 + wy;
            win[k] = t < nneed && gx >= 0 && gy >= 0 && gx < gw && gy < gh_cells;
            wc[k] = wy * kWin + wx;
            const int g = win[k] ? gy * gw + gx : 0;
            vo[k] = ord_g[g];
            vi[k] = need_info ? inf_g[g] : 0u;
        }
#pragma unroll
        for (int k = 0; k < kNeedPer; k++) {
            if (tid + k * kPrepThreads < nneed) {
                L.ord[wc[k]] = win[k] ? vo[k] : 0xffffffffu;
                L.inf[wc[k]] = win[k] ? vi[k] : 0u;
            }
        }
    }
    if (tid < 12) L.cnt[tid] = 0;
    __syncthreads();
    // job order: TBs by size class, and within one by prediction class, so
    // that the TBs sharing a wave mostly take one branch of the word loop
    static_assert(kMaxUnitTbs <= kPrepThreads, "one TB per thread");
    int key = 0, rank = 0;
    if (tid < ntb) {
        const PrepTb t = unit_recs[tb_list + tid];
        L.tb[tid] = t;
        L.ro[tid] = unit_refoff[tb_list + tid];
        const int m = (t.flags & HPVCD_TB_PCM) ? 2 : t.mode;
        const int pc = m == 0 ? kClsPlanar : m == 1 ? kClsDc : (m == 10 || m == 26) ? kClsHv : kClsAngular;
        key = (t.log2n == 2 ? 0 : t.log2n == 3 ? 4 : 8) + pc;
        rank = (int)atomicAdd(&L.cnt[key], 1u);
    }
    __syncthreads();
    if (tid < ntb) {
        int base = 0;
        for (int k = 0; k < key; k++) base += (int)L.cnt[k];
        L.job[base + rank] = (uint16_t)tid;
    }
    const int n4 = (int)(L.cnt[0] + L.cnt[1] + L.cnt[2] + L.cnt[3]);
    const int n8 = (int)(L.cnt[4] + L.cnt[5] + L.cnt[6] + L.cnt[7]);
    __syncthreads();
    const int lgv = 2 - sh, lgh = 2 - sw;
    const int gv = 4 >> sh, gh = 4 >> sw;
    const bool strong_on = plane == 0 && (pflags & HPVCD_PF_STRONG_INTRA_SMOOTHING);
    // lanes per TB: enough for the TB's reference groups (2n / 4-sample
    // group height on the left, the corner, 2n / group width above), so 4x4
    // TBs go eight to a wave where the groups fit 8 lanes (else four), 8x8
    // TBs four where they fit 16 (else two), the others one
    const int lg4 = (8 >> lgv) + 1 + (8 >> lgh) <= 8 ? 3 : 4;
    const int lg8 = (16 >> lgv) + 1 + (16 >> lgh) <= 16 ? 4 : 5;
    const int t4 = 64 >> lg4, t8 = 64 >> lg8;
    const int j4 = (n4 + t4 - 1) / t4, j8 = (n8 + t8 - 1) / t8, njobs = j4 + j8 + (ntb - n4 - n8);
    // per job: availability mask (one lane per 4-sample group), substituted
    // reference map, descriptor, per-sample reference words of each TB
    for (int j = wave; j < njobs; j += kWaves) {
        const int lg = j < j4 ? lg4 : (j < j4 + j8 ? lg8 : 6);  // log2 of the lanes per TB
        const int G = 1 << lg, gl = lane & (G - 1), gbase = lane & ~(G - 1);
        int first, end;  // the job's first TB slot (always present) and the end of its size class
        if (j < j4) { first = t4 * j; end = n4; }
        else if (j < j4 + j8) { first = n4 + t8 * (j - j4); end = n4 + n8; }
        else { first = n4 + n8 + (j - j4 - j8); end = ntb; }
        const int slot = first + (lane >> lg);
        const bool has = slot < end;
        const int e = L.job[has ? slot : first];
        const bool small = lg < 6;
        const PrepTb tb = L.tb[e];
        const int log2n = tb.log2n, mode = tb.mode, n = 1 << log2n, n2 = 2 * n;
        const int x = tb.x, y = tb.y, x0 = x - ux0, y0 = y - uy0;
        const uint32_t order = tb.order, ro = L.ro[e];
        const int roff = ro & 0x1fff, soff = ((ro >> 13) & 0xff) << 4;
        const int ngl = n2 / gv, nga = n2 / gh;
        auto cell = [&](int px, int py) -> int {
            return (((py << sh) >> 2) - gy0) * kWin + (((px << sw) >> 2) - gx0);
        };
        const uint32_t cur = L.inf[cell(x, y)];
        int px = x - 1, py = y - 1;
        if (gl < ngl) py = y + n2 - (gl + 1) * gv;
        else if (gl > ngl) { px = x + (gl - ngl - 1) * gh; }
        bool av = has && gl < ngl + 1 + nga && px >= 0 && py >= 0 && px < pw && py < ph && (px << sw) < lw &&
                  (py << sh) < lh;
        if (av) {
            const int c = cell(px, py);
            av = L.ord[c] < order;
            if (av && need_info) {
                const uint32_t i = L.inf[c];
                if ((pflags & HPVCD_PF_TILES) && HPVCD_CELL_TILE(i) != HPVCD_CELL_TILE(cur)) av = false;
                if (num_slices > 1 && HPVCD_CELL_SLICE(i) != HPVCD_CELL_SLICE(cur)) av = false;
                if ((pflags & HPVCD_PF_CONSTRAINED_INTRA) && !(i & HPVCD_CELL_INTRA)) av = false;
            }
        }
        const uint64_t mall = __ballot(av);
        const uint64_t m = small ? (mall >> gbase) & ((1ull << G) - 1ull) : mall;
        // substituted reference map
        uint16_t* M = L.rmap + roff;
        if (has)
            for (int k = gl; k <= 2 * n2; k += G) {
                const uint16_t addr = (uint16_t)ref_tile_addr(k, m, n, x0, y0, lgv, lgh);
                M[k] = addr;
                rmap_g[rbase + roff + k] = addr;
            }
        wave_sync();
        if (!has) continue;
        // border samples read (substitution only ever reads available
        // samples): the row above for a TB on the unit's top edge, the left
        // column for one on its left edge
        uint32_t top_need = 0, left_need = 0;
        const uint64_t left_bits = m & ((1ull << ngl) - 1ull);
        const uint64_t above_bits = (m >> (ngl + 1)) & ((1ull << nga) - 1ull);
        const bool corner = (m >> ngl) & 1ull;
        if (y0 == 0) {
            if (above_bits) top_need = (uint32_t)(x0 + ((64 - __builtin_clzll(above_bits)) << lgh) + 1);
            else if (corner) top_need = (uint32_t)(x0 + 1);
        }
        if (x0 == 0) {
            if (left_bits) left_need = (uint32_t)(y0 + n2 - ((int)__builtin_ctzll(left_bits) << lgv));
            if (corner && y0 > 0) left_need = max(left_need, (uint32_t)y0);
        }
        // I_PCM blocks read no neighbours: their words name, through the
        // angular form with fraction 0, the zero slot (the samples arrive as
        // the residual)
        const bool pcm = (tb.flags & HPVCD_TB_PCM) != 0;
        if (pcm) top_need = left_need = 0;
        TbShape s = tb_shape(log2n, mode, plane, pflags, chroma_idc, angles, inv_angles);
        if (pcm) {
            s.cls = kClsAngular;
            s.filt = false;
        }
        const uint16_t* Ms = s.filt ? nullptr : M;
        auto A = [&](int s_) -> uint32_t { return Ms ? (uint32_t)Ms[s_] : (uint32_t)s_; };
        uint32_t u0 = 0, u1 = 0;
        if (s.cls == kClsHv) {
            u0 = A(mode == 26 ? n2 + 1 : n2 - 1);
            u1 = A(n2);
        } else if (s.cls == kClsPlanar) {
            u0 = A(n2 + 1 + n);
            u1 = A(n2 - 1 - n);
        }
        if (gl == 0) {
            // a step's first TB also carries the size of the step after it
            // and whether any TB of its step writes the unit's bottom row or
            // right column
            // (step bits marked by the host on the step's first TB)
            const int gsz = max(1, (int)((tb.flags >> TBX_GROUP_SHIFT) & 7u));
            const uint32_t gnext = ((tb.flags >> TBX_NEXT_SHIFT) & 3u) + 1;
            const bool pub = (tb.flags & TBX_STEP_PUB) != 0, plain = (tb.flags & TBX_STEP_PLAIN) != 0;
            const bool fast = n <= 8 && !s.filt && s.cls != kClsDc;
            uint4 d;
            d.x = (uint32_t)x0 | ((uint32_t)y0 << 6) | ((uint32_t)(log2n - 2) << 12) | ((uint32_t)s.cls << 14) |
                  (s.filt ? 1u << 17 : 0u) |
                  ((s.filt && strong_on && n == 32) ? 1u << 18 : 0u) | (y0 + n == ctbh ? 1u << 19 : 0u) |
                  (x0 + n == ctbw ? 1u << 20 : 0u) | (fast ? 1u << 21 : 0u) |
                  ((uint32_t)((tb.flags >> TBX_GROUP_SHIFT) & 7u) << 22) | ((gnext > 1 ? 3u : 0u) << 25) |
                  (pub ? 1u << 28 : 0u) | (plain ? 1u << 29 : 0u) |
                  ((uint32_t)((gsz > 1 ? gsz : (n == 8 ? 4 : 1)) - 1) << 30);
            d.y = (ro & 0x1fffu) | ((uint32_t)soff << 13);
            d.z = u0 | (u1 << 13);
            d.w = top_need | (left_need << 8) | ((uint32_t)(e + gsz) << 16) | (fast ? 0u : 1u << 25);
            desc_g[tb_list + e] = d;
        }
        uint32_t* dst = smap_g + sbase + soff;
        for (int p = gl; p < n * n; p += G)
            dst[p] = pcm ? (uint32_t)kZero | ((uint32_t)kZero << 13) : sample_word(s, p & (n - 1), p >> log2n, Ms);
    }
    __syncthreads();
}

// Per-TB descriptor, decoded once per TB with scalar bit-field extracts.
//   x: [0,6) x0  [6,12) y0 (unit relative)  [12,14) log2n-2  [14,16) class
//      17 filtered references  18 strong-filter candidate
//      19 writes the unit's bottom row  20 writes its right column
//      21 fast path: at most 8 x 8, references unfiltered, not DC
//      [22,25) size of the group of independent 4 x 4 TBs this TB leads
//      (step leaders only:) [25,27) 3 if the next step is a group of
//      several TBs, else 0  28 some TB of
//      the step writes the unit's bottom row or right column  29 some TB of
//      the step is planar or horizontal / vertical  [30,32) sixteens of
//      samples in the step minus one (TBs of at most 8 x 8)
//   y: [0,13) reference map offset  [13,25) sample word offset
//   z: uniform references of the class: [0,13) u0  [13,26) u1
//   w: [0,8) row-above words read  [8,16) left-column rows read
//      (step leaders:) [16,25) index of the TB after the step  25 the
//      step is not fast (bit 21 clear)
struct IntraLds {
    uint16_t tile[kLdsSamples];
    alignas(16) uint32_t smap[kMaxCtb * kMaxCtb];
    uint4 desc[kMaxUnitTbs + 4];      // a not-fast descriptor follows the last TB
    alignas(16) uint16_t refmap[kMaxUnitRefs];
    uint32_t pneed[kMaxUnitTbs];      // prefix maxima over the TB order of the neighbour samples read:
                                      // row-above words | left-column rows << 16
    uint32_t imask[8];                // intra 4 x 4 blocks of an inter unit: row r in bits 16 (r & 1) of word r / 2
    uint32_t bready;                  // TBs whose neighbour samples are all in the tile
    uint32_t unit;
};

// Prediction of one n x n TB (n = 1 << LOG2N) from its per-sample reference
// words (src/intra.rs:257), residual add and write into the unit tile.  Lane l
// predicts samples l, l + 64, ... in raster order.  S indexes the tile and the
// filtered reference rows (fbase selects the row of this wave); u0/u1 are the
// class's uniform references.  For n <= 8 the word and residual of the lane's
// sample may come prefetched (pf).
template <int LOG2N, int CLS, typename R>
__device__ __forceinline__ void intra_predict(const uint16_t* __restrict__ S, const uint32_t* __restrict__ W,
                                              uint16_t* __restrict__ dst, const R* __restrict__ rs,
                                              bool pf, uint32_t w_pf, int32_t r_pf, int lane, int fbase, int u0,
                                              int u1, int32_t dc, int32_t maxv, uint32_t* __restrict__ bw,
                                              int pub_b, int pub_r, uint32_t tagw) {
    constexpr int n = 1 << LOG2N, nn = n * n;
    constexpr int NQ = nn < 64 ? 1 : nn / 64;
    constexpr int CH = NQ < 16 ? NQ : 16;  // samples per lane whose LDS reads overlap
    const int32_t ru0 = (CLS == kClsHv || CLS == kClsPlanar) ? (int32_t)S[u0] : 0;
    const int32_t ru1 = (CLS == kClsHv || CLS == kClsPlanar) ? (int32_t)S[u1] : 0;
#pragma unroll 1
    for (int q0 = 0; q0 < NQ; q0 += CH) {
        uint32_t w[CH];
        int32_t res[CH], v[CH];
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const int p = ((q0 + c) * 64 + lane) & (nn - 1);
            if (NQ == 1 && pf) {
                w[c] = w_pf;
                res[c] = r_pf;
            } else {
                w[c] = W[p];
                // zero where no residual; 16-bit residuals wait in the tile at the sample itself
                res[c] = sizeof(R) == 2 ? (int32_t)(int16_t)dst[(p >> LOG2N) * kRowW + (p & (n - 1))]
                                        : (int32_t)rs[(p >> LOG2N) * kMaxCtb + (p & (n - 1))];
            }
        }
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const int p = ((q0 + c) * 64 + lane) & (nn - 1);
            const int x = p & (n - 1), y = p >> LOG2N;
            const int32_t r0 = S[(w[c] & 0x1fffu) + fbase], r1 = S[((w[c] >> 13) & 0x1fffu) + fbase];
            const int k = (w[c] >> 26) & 31;
            int32_t r;
            if (CLS == kClsAngular) {
                r = (mul24(32 - k, r0) + mul24(k, r1) + 16) >> 5;
            } else if (CLS == kClsHv) {
                r = (int32_t)(w[c] >> 31) ? clampi(ru0 + ((r1 - ru1) >> 1), 0, maxv) : r0;
            } else if (CLS == kClsPlanar) {
                r = (mul24(n - 1 - x, r0) + mul24(x + 1, ru0) + mul24(n - 1 - y, r1) + mul24(y + 1, ru1) + n) >>
                    (LOG2N + 1);
            } else {
                // kind 0: dc; 1: (r0 + 3 dc + 2) >> 2; 2: (r0 + 2 dc + r1 + 2) >> 2
                r = ((k ? r0 : 0) + (4 - k) * dc + ((k >> 1) ? r1 : 0) + 2) >> 2;
            }
            v[c] = clampi(wadd(r, res[c]), 0, maxv);
        }
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const int p = ((q0 + c) * 64 + lane) & (nn - 1);
            if (nn >= 64 || lane < nn) {
                const int x = p & (n - 1), y = p >> LOG2N;
                dst[y * kRowW + x] = (uint16_t)v[c];
                // the unit's bottom row and right column go out at once as
                // tagged words for the units below and to the right
                if (pub_b >= 0 && y == n - 1) border_store(bw + pub_b + x, tagw | (uint32_t)v[c]);
                if (pub_r >= 0 && x == n - 1) border_store(bw + pub_r + y, tagw | (uint32_t)v[c]);
            }
        }
    }
}

// One TB step of the wavefront for TBs of size n = 1 << LOG2N: filtered
// references ([1 2 1] or strong, src/intra.rs:151) into this wave's row when
// the TB filters, the DC value, then prediction.
template <int LOG2N, typename R>
__device__ __forceinline__ void intra_tb(IntraLds& L, const R* __restrict__ resb, uint32_t d0, uint32_t d1, uint32_t d2,
                                         bool pf, uint32_t w_pf, int32_t r_pf, int lane, int bd, int32_t maxv,
                                         uint32_t* unit_border, uint32_t tagw) {
    constexpr int n = 1 << LOG2N, n2 = 2 * n, n4 = 4 * n;
    constexpr int NIT = (n4 + 64) / 64;  // passes over the 4n + 1 reference slots
    uint16_t* S = L.tile;
    const int x0 = d0 & 63, y0 = (d0 >> 6) & 63, cls = (d0 >> 14) & 3;
    const bool filt = (d0 >> 17) & 1, scand = (d0 >> 18) & 1;
    const uint16_t* M = L.refmap + (d1 & 0x1fff);
    const uint32_t* W = L.smap + ((d1 >> 13) & 0xfff);
    int fbase = 0;
    if (LOG2N > 2 && filt) {
        uint16_t* F = S + kFBase;
        bool strong = false;
        int32_t bl = 0, tl = 0, tr = 0;
        if (LOG2N == 5 && scand) {
            bl = S[M[0]];
            tl = S[M[n2]];
            tr = S[M[n4]];
            const int thr = 1 << (bd - 5);
            const int a = bl + tl - 2 * (int32_t)S[M[n]], b = tr + tl - 2 * (int32_t)S[M[3 * n]];
            strong = uni((int)((a < 0 ? -a : a) < thr && (b < 0 ? -b : b) < thr)) != 0;
        }
        if (LOG2N == 5 && strong) {
#pragma unroll
            for (int it = 0; it < NIT; it++) {
                const int i = it * 64 + lane;
                if (i <= n4)
                    F[i] = (uint16_t)(i < n2 ? (i * tl + (n2 - i) * bl + n) >> 6 : ((n4 - i) * tl + (i - n2) * tr + n) >> 6);
            }
        } else {
#pragma unroll
            for (int it = 0; it < NIT; it++) {
                const int i = it * 64 + lane;
                if (i <= n4) {
                    const bool endp = i == 0 || i == n4;  // end samples stay unfiltered
                    const int i0 = endp ? i : i - 1, i2 = endp ? i : i + 1;
                    F[i] = (uint16_t)(((int32_t)S[M[i0]] + 2 * (int32_t)S[M[i]] + (int32_t)S[M[i2]] + 2) >> 2);
                }
            }
        }
        wave_sync();
        fbase = kFBase;
    }
    int32_t dc = 0;
    if (cls == kClsDc) {  // DC never filters: sum of the n left and n above references
        int32_t v = 0;
        if (lane < n2) v = S[M[lane < n ? n2 - n + lane : n2 + 1 + lane - n]];
        dc = (wave_sum(v) + n) >> (LOG2N + 1);
    }
    const int u0 = (int)(d2 & 0x1fff) + fbase, u1 = (int)((d2 >> 13) & 0x1fff) + fbase;
    uint16_t* dst = S + kTopW + y0 * kRowW + x0 + 1;
    const R* rs = sizeof(R) == 2 ? resb : resb + y0 * kMaxCtb + x0;
    const int pb = ((d0 >> 19) & 1) ? x0 : -1;            // bottom-row word of column x0
    const int pr = ((d0 >> 20) & 1) ? kMaxCtb + y0 : -1;  // right-column word of row y0
    switch (cls) {
        case kClsAngular:
            intra_predict<LOG2N, kClsAngular, R>(S, W, dst, rs, pf, w_pf, r_pf, lane, fbase, u0, u1, dc, maxv, unit_border, pb, pr, tagw);
            break;
        case kClsHv:
            intra_predict<LOG2N, kClsHv, R>(S, W, dst, rs, pf, w_pf, r_pf, lane, fbase, u0, u1, dc, maxv, unit_border, pb, pr, tagw);
            break;
        case kClsPlanar:
            intra_predict<LOG2N, kClsPlanar, R>(S, W, dst, rs, pf, w_pf, r_pf, lane, fbase, u0, u1, dc, maxv, unit_border, pb, pr, tagw);
            break;
        default:
            intra_predict<LOG2N, kClsDc, R>(S, W, dst, rs, pf, w_pf, r_pf, lane, fbase, u0, u1, dc, maxv, unit_border, pb, pr, tagw);
    }
    wave_sync();
}

// Staging of a unit: the 16-byte vectors of its descriptors, reference map
// and reference words (at most 256 + 4352 / 8 + 4096 / 4), and its residual
// block in 4-sample vectors, all loads of a thread in flight at once.
constexpr int kStageVec = (kMaxUnitTbs + kMaxUnitRefs / 8 + kMaxCtb * kMaxCtb / 4 + kIntraThreads - 1) / kIntraThreads;
constexpr int kStageRes = kMaxCtb * kMaxCtb / 4 / kIntraThreads;

// Four consecutive plane samples as one load.
__device__ __forceinline__ void load_quad(const uint8_t* p, uint32_t (&v)[4]) {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(p);
    v[0] = w & 0xff; v[1] = (w >> 8) & 0xff; v[2] = (w >> 16) & 0xff; v[3] = w >> 24;
}
__device__ __forceinline__ void load_quad(const uint16_t* p, uint32_t (&v)[4]) {
    const uint2 w = *reinterpret_cast<const uint2*>(p);
    v[0] = w.x & 0xffff; v[1] = w.x >> 16; v[2] = w.y & 0xffff; v[3] = w.y >> 16;
}

// Four consecutive plane samples as one store.
__device__ __forceinline__ void store_quad(uint8_t* p, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    *reinterpret_cast<uint32_t*>(p) = a | (b << 8) | (c << 16) | (d << 24);
}
__device__ __forceinline__ void store_quad(uint16_t* p, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    *reinterpret_cast<uint2*>(p) = make_uint2(a | (b << 16), c | (d << 16));
}

// Four residual samples as one load: 8 bytes of int16, 16 of int32.
template <typename R> struct Quad;
template <> struct Quad<int16_t> { using V = uint2; };
template <> struct Quad<int32_t> { using V = uint4; };
template <typename Q> __device__ __forceinline__ Q quad_zero();
template <> __device__ __forceinline__ uint2 quad_zero<uint2>() { return make_uint2(0, 0); }
template <> __device__ __forceinline__ uint4 quad_zero<uint4>() { return make_uint4(0, 0, 0, 0); }

// Four 16-bit residual samples into the tile, one 8-byte store (t is a
// body sample whose x is a multiple of 4).
__device__ __forceinline__ void store_res_tile(uint16_t* t, uint2 q) { *reinterpret_cast<uint2*>(t) = q; }

// The reference samples of a fast step's lane: the angular pair (r0, r1) and,
// for planar and H/V, the class's uniform pair (ru0, ru1).
struct StepRefs {
    int32_t r0, r1, ru0, ru1;
};

// The prediction wave's registers between fast steps: the step's first TB
// index and descriptor (wave-uniform), and per lane its word, residual, tile
// index and TB descriptor words (w, r, ti, l0, l2) and the next step's
// descriptor and lane descriptor words as loaded (n*, nl*).  The two kinds
// are kept apart so that the uniform ones stay in scalar registers.
struct FastUni {
    int i;
    uint32_t d0, d1, d2, d3;
};
struct FastLane {
    uint32_t w;
    int32_t r;
    int ti;
    uint32_t l0, l2, n0, n1, n2, n3, nl0, nl2;
};

// The common step: one TB of at most 8 x 8 or a group of up to four
// independent 4 x 4 TBs of one dependency level, references unfiltered, not
// DC, one sample per lane.  Lane l takes sample l of the step (a group's
// samples run TB after TB, 16 each); l0/l2 are the descriptor words of its
// TB, w/rv its prefetched reference word and residual, (x, y) its sample in
// the TB and ti the sample's tile index.  Lanes past the step's samples
// write to the sink slot.  Planar and H/V need two more reference slots
// (descriptor word z), read only when `mixed` says some lane needs them.
__device__ __forceinline__ void intra_step_fast(uint16_t* __restrict__ S, uint32_t l0, const StepRefs& q, uint32_t w,
                                                int32_t rv, bool b8, bool mixed, int x, int y, int ti, bool active,
                                                int32_t maxv) {
    const int32_t r0 = q.r0, r1 = q.r1;
    const int k = (w >> 26) & 31;
    int32_t r = (mul24(32 - k, r0) + mul24(k, r1) + 16) >> 5;
    if (mixed) {
        const int n = b8 ? 8 : 4, sh = b8 ? 4 : 3;
        const int cls = (l0 >> 14) & 3;
        const int32_t ru0 = q.ru0, ru1 = q.ru1;
        const int32_t hv = (int32_t)(w >> 31) ? max(0, min(ru0 + ((r1 - ru1) >> 1), maxv)) : r0;
        const int32_t pl = (mul24(n - 1 - x, r0) + mul24(x + 1, ru0) + mul24(n - 1 - y, r1) + mul24(y + 1, ru1) + n) >> sh;
        r = cls == kClsAngular ? r : (cls == kClsHv ? hv : pl);
    }
    const int32_t v = max(0, min((int32_t)wadd(r, rv), maxv));
    S[active ? ti : kSink] = (uint16_t)v;
}

// The border samples of a fast step: the unit's bottom row and right column
// go out at once as tagged words for the units below and to the right.
__device__ __forceinline__ void intra_step_publish(const uint16_t* __restrict__ S, uint32_t l0, bool b8, int x, int y,
                                                   int ti, bool active, uint32_t* __restrict__ bw, uint32_t tagw) {
    const int n = b8 ? 8 : 4, x0 = l0 & 63, y0 = (l0 >> 6) & 63;
    if (!active) return;
    const uint32_t v = S[ti];
    if (((l0 >> 19) & 1) && y == n - 1) border_store(bw + x0 + x, tagw | v);
    if (((l0 >> 20) & 1) && x == n - 1) border_store(bw + kMaxCtb + y0 + y, tagw | v);
}

// TBs of one step of the prediction wave: a group of up to four independent
// 4 x 4 TBs (descriptor bits 22-24 of its first TB), else one TB.
__device__ __forceinline__ int step_size(uint32_t d0) { return max(1, (int)((d0 >> 22) & 7)); }
// 3 if the step after the one d0 leads is a group of several TBs, else 0
// (descriptor bits 25-26): the mask of the lane's TB in it.
__device__ __forceinline__ int next_step_spread(uint32_t d0) { return (int)((d0 >> 25) & 3u); }
// Index of the TB after the step whose leader's descriptor word w is d3.
__device__ __forceinline__ int step_end(uint32_t d3) { return (int)((d3 >> 16) & 0x1ffu); }
// step_end, plus 512 when the step is not fast: a fast run goes on while the
// next step's key is at most the count of ready TBs.
__device__ __forceinline__ int step_key(uint32_t d3) { return (int)((d3 >> 16) & 0x3ffu); }
// Whether lane l (lane16 = l / 16) holds a sample of a step of TBs of at
// most 8 x 8.
__device__ __forceinline__ bool step_active(uint32_t d0, int lane16) { return lane16 <= (int)(d0 >> 30); }

// step_prefetch without the size test, for a step known to be fast (its
// loads then have a fixed count, which lets the reference reads issued before
// them be waited for alone).
template <typename R>
__device__ __forceinline__ void step_prefetch_fast(const IntraLds& L, const R* __restrict__ resb, uint32_t d0,
                                                   uint32_t d1, uint32_t l0, int lane, int lt4, int lt8,
                                                   int lr4, int lr8, uint32_t& w, int32_t& r, int& ti) {
    const bool b8 = (d0 >> 12) & 1;
    w = L.smap[min((int)((d1 >> 13) & 0xfff) + lane, kMaxCtb * kMaxCtb - 1)];
    ti = kTopW + 1 + mul24((l0 >> 6) & 63, kRowW) + (int)(l0 & 63) + (b8 ? lt8 : lt4);
    if constexpr (sizeof(R) == 2) r = reinterpret_cast<const int16_t*>(L.tile)[ti];
    else r = (int32_t)resb[(l0 & 0xfffu) + (b8 ? lr8 : lr4)];
}

// The lane's reference word, residual and tile index of a step of TBs of at
// most 8 x 8: lane l takes word l of the step (a group's words are
// contiguous, 16 per TB).  d0/d1 are the first TB's descriptor words, l0 the
// lane's own TB's; lt4/lt8 (lr4/lr8) are the lane's tile (residual block)
// offsets in a 4 x 4 / 8 x 8 TB.  16-bit residuals wait in the tile at their
// samples, 32-bit ones in the residual block; both hold zeros where a TB has
// no residual.
template <typename R>
__device__ __forceinline__ void step_prefetch(const IntraLds& L, const R* __restrict__ resb, uint32_t d0,
                                              uint32_t d1, uint32_t l0, int lane, int lt4, int lt8,
                                              int lr4, int lr8, uint32_t& w, int32_t& r, int& ti) {
    // (a group's TBs are 4 x 4)
    if (!((d0 >> 13) & 1)) {
        const bool b8 = (d0 >> 12) & 1;
        w = L.smap[min((int)((d1 >> 13) & 0xfff) + lane, kMaxCtb * kMaxCtb - 1)];
        ti = kTopW + 1 + mul24((l0 >> 6) & 63, kRowW) + (int)(l0 & 63) + (b8 ? lt8 : lt4);
        if constexpr (sizeof(R) == 2) r = reinterpret_cast<const int16_t*>(L.tile)[ti];
        else r = (int32_t)resb[(l0 & 0xfffu) + (b8 ? lr8 : lr4)];
    }
}

// Index of the descriptor of the lane's TB in a step that starts at TB i,
// spread the step's next_step_spread mask (lanes past a group read a later
// TB, or the descriptors after the last; their results go to the sink
// slot).
__device__ __forceinline__ int lane_desc(int i, int spread, int lane16) { return i + (lane16 & spread); }

__device__ __forceinline__ void lds_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local"); }
__device__ __forceinline__ void lds_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local"); }

// The intra wavefront.  A persistent grid takes (picture, plane, CTB) units
// in ticket order (a topological order of the CTB dependencies).  Per unit,
// wave 0 reconstructs the TBs one after the other in an order where every TB
// follows the TBs it references, and publishes the unit's bottom row and
// right column as tagged border words the moment they are written; wave 1
// polls the neighbours' border words the unit reads and moves them into the
// unit tile, counting in L.bready how many TBs can start.  The two waves meet
// only at the start and end of a unit.
template <typename T, typename R>
__global__ void __launch_bounds__(kIntraThreads)
    k_intra(const DevPic* __restrict__ pics, const DevUnit* __restrict__ units,
            const uint32_t* __restrict__ ticket_units, uint32_t n_units, const uint4* __restrict__ desc_g,
            const uint16_t* __restrict__ rmap_g, const uint32_t* __restrict__ smap_g, T* __restrict__ samples,
            const R* __restrict__ res_pool, uint32_t* __restrict__ border, uint32_t* counter, const uint32_t* gen,
            uint32_t* err) {
    __shared__ IntraLds L;
    // 32-bit residuals have a block of their own; 16-bit ones are staged
    // into the tile at their samples and replaced there by the reconstruction
    constexpr bool kResInTile = sizeof(R) == 2;
    __shared__ __align__(16) R resb[kResInTile ? 8 : kMaxCtb * kMaxCtb];
    const int tid = threadIdx.x, wave = uni(tid >> 6), lane = tid & 63;
    const uint32_t tag = *gen & 0xffffu;

    for (;;) {
        if (tid == 0) L.unit = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t ticket = uni(L.unit);
        __syncthreads();
        if (ticket >= n_units) return;
        // unit and picture descriptors are wave-uniform: scalar loads and
        // scalar control flow from here on
        const uint32_t ui = uni(ticket_units[ticket]);
        const DevUnit U = units[ui];
        const DevPic& P = pics[uni(U.pic)];
        const int plane = uni((int)U.plane);
        const int sw = plane ? P.sub_w_log2 : 0, sh = plane ? P.sub_h_log2 : 0;
        const int ctb = 1 << P.ctb_log2;
        const int ctbw = ctb >> sw, ctbh = ctb >> sh;       // unit size in plane samples
        const int ctbw_log2 = P.ctb_log2 - sw;
        const int pw = plane ? P.cw : P.w, ph = plane ? P.ch : P.h;
        const int ux0 = U.cx * ctbw, uy0 = U.cy * ctbh;     // unit origin (plane samples)
        const int bd = plane ? P.bd_c : P.bd;
        const int32_t maxv = (1 << bd) - 1;
        T* pl = samples + P.plane_off[plane];
        const R* rp = res_pool + P.res_off[plane];
        const int ntb = U.ntb;

        // stage the unit: TB descriptors, reference maps and words (all
        // prepared before the wavefront) and residuals, in 16-byte (words,
        // maps) and 4-sample (residual) vectors.  Unit map offsets are
        // multiples of 8 entries, sample word offsets of 16; residual rows of
        // a plane start at multiples of 4 samples and its width is a
        // multiple of 4.
        {
            // one batch: every thread issues all its loads, then stores
            const int v1 = ntb, v2 = v1 + ((int)(U.nref + 7) >> 3), v3 = v2 + ((int)U.nsamp >> 2);
            const uint4* rm16 = reinterpret_cast<const uint4*>(rmap_g + U.rbase);
            const uint4* sm16 = reinterpret_cast<const uint4*>(smap_g + U.sbase);
            using Q = typename Quad<R>::V;
            const int nres = ctbw * ctbh;
            uint4 va[kStageVec];
            Q vr[kStageRes];
#pragma unroll
            for (int b = 0; b < kStageVec; b++) {
                const int e = min(tid + b * kIntraThreads, v3 - 1);
                va[b] = e < v1 ? desc_g[U.desc_base + e] : (e < v2 ? rm16[e - v1] : sm16[e - v2]);
            }
#pragma unroll
            for (int b = 0; b < kStageRes; b++) {
                const int e = 4 * (tid + b * kIntraThreads);
                const int ry = e >> ctbw_log2, rx = e - (ry << ctbw_log2);
                const int px = ux0 + rx, py = uy0 + ry;
                const bool in = e < nres && px < pw && py < ph;
                const Q q = *reinterpret_cast<const Q*>(rp + (in ? (size_t)py * pw + px : 0));
                vr[b] = in ? q : quad_zero<Q>();
            }
#pragma unroll
            for (int b = 0; b < kStageVec; b++) {
                const int e = tid + b * kIntraThreads;
                if (e < v1) L.desc[e] = va[b];
                else if (e < v2) reinterpret_cast<uint4*>(L.refmap)[e - v1] = va[b];
                else if (e < v3) reinterpret_cast<uint4*>(L.smap)[e - v2] = va[b];
            }
#pragma unroll
            for (int b = 0; b < kStageRes; b++) {
                const int e = 4 * (tid + b * kIntraThreads);
                if (e < nres) {
                    if constexpr (kResInTile) {
                        if (!P.inter) store_res_tile(L.tile + kTopW + 1 + (e >> ctbw_log2) * kRowW + (e & (ctbw - 1)), vr[b]);
                    } else {
                        *reinterpret_cast<Q*>(resb + (e >> ctbw_log2) * kMaxCtb + (e & (ctbw - 1))) = vr[b];
                    }
                }
            }
            if (P.inter) {
                // motion-compensated prediction of the unit, the predictor of its inter TBs
                for (int e0 = tid; e0 < nres; e0 += kCopyBatch * kIntraThreads) {
                    T v[kCopyBatch];
#pragma unroll
                    for (int b = 0; b < kCopyBatch; b++) {
                        const int e = e0 + b * kIntraThreads;
                        const int ry = e >> ctbw_log2, rx = e - (ry << ctbw_log2);
                        const int px = ux0 + rx, py = uy0 + ry;
                        v[b] = (e < nres && px < pw && py < ph) ? pl[(size_t)py * pw + px] : (T)0;
                    }
#pragma unroll
                    for (int b = 0; b < kCopyBatch; b++) {
                        const int e = e0 + b * kIntraThreads;
                        if (e < nres) L.tile[kTopW + (e >> ctbw_log2) * kRowW + (e & (ctbw - 1)) + 1] = (uint16_t)v[b];
                    }
                }
                if constexpr (kResInTile) {
                    // residuals into the tile over the predictor, at the intra TBs only
                    if (tid < 8) L.imask[tid] = 0;
                    __syncthreads();
                    for (int i = tid; i < ntb; i += kIntraThreads) {
                        const uint32_t d0 = L.desc[i].x;
                        const int bx = (d0 & 63) >> 2, by = ((d0 >> 6) & 63) >> 2, nb = 1 << ((d0 >> 12) & 3);
                        const uint32_t run = ((1u << nb) - 1u) << bx;
                        for (int r = by; r < by + nb; r++) atomicOr(&L.imask[r >> 1], run << ((r & 1) * 16));
                    }
                    __syncthreads();
#pragma unroll
                    for (int b = 0; b < kStageRes; b++) {
                        const int e = 4 * (tid + b * kIntraThreads);
                        const int ry = e >> ctbw_log2, rx = e & (ctbw - 1);
                        if (e < nres && ((L.imask[(ry >> 3) & 7] >> (((ry >> 2) & 1) * 16 + (rx >> 2))) & 1u))
                            store_res_tile(L.tile + kTopW + 1 + ry * kRowW + rx, vr[b]);
                    }
                }
            }
        }
        if (tid == 0) {
            // after the last TB: a descriptor that is not fast (key 512, above
            // any count of ready TBs), which ends a fast run
            L.desc[ntb] = make_uint4(0u, 0u, 0u, 1u << 25);
            L.bready = 0;
            L.tile[kNeutral] = (uint16_t)(1 << (bd - 1));
            L.tile[kZero] = 0;
        }
        __syncthreads();

        uint32_t* const bw = border + (size_t)ui * kBorderWords;
        const uint32_t tagw = tag << 16;
        if (wave == 1) {
            // ---- neighbour samples ----
            // prefix maxima of the per-TB needs over the TB order
            uint32_t carry_t = 0, carry_l = 0;
            for (int c = 0; c < ntb; c += 64) {
                const int i = c + lane;
                const uint32_t w = i < ntb ? L.desc[i].w : 0u;
                const uint32_t vt = max(wave_prefix_max(w & 0xffu), carry_t);
                const uint32_t vl = max(wave_prefix_max((w >> 8) & 0xffu), carry_l);
                if (i < ntb) L.pneed[i] = vt | (vl << 16);
                carry_t = (uint32_t)__builtin_amdgcn_readlane((int)vt, 63);
                carry_l = (uint32_t)__builtin_amdgcn_readlane((int)vl, 63);
            }
            // lane l owns row-above words l, l + 64, l + 128 and left row l;
            // tile[e] of the row above is sample (ux0 - 1 + e, uy0 - 1), the
            // bottom row of unit (ncx, cy - 1); left rows come from the right
            // column of unit (cx - 1, cy)
            const int all_top = (int)carry_t, all_left = (int)carry_l;
            uint32_t valid = 0, done = 0;
#pragma unroll
            for (int w = 0; w < 3; w++)
                if (lane + 64 * w < all_top && ux0 - 1 + lane + 64 * w >= 0) valid |= 1u << w;
            if (lane < all_left) valid |= 8u;
            auto wptr = [&](int w) -> const uint32_t* {
                if (w == 3) return border + (size_t)(ui - 1) * kBorderWords + kMaxCtb + lane;
                const int px = ux0 - 1 + lane + 64 * w, ncx = px >> ctbw_log2;
                const uint32_t nu = ui - (uint32_t)U.cx + (uint32_t)ncx - (uint32_t)P.ctb_cols;
                return border + (size_t)nu * kBorderWords + (px - (ncx << ctbw_log2));
            };
            int ready = 0;
            uint32_t spins = 0;
            for (;;) {
                // poll only the missing words that the next few TBs read
                const int horizon = min(ntb, ready + kFetchAhead) - 1;
                const uint32_t hn = L.pneed[horizon];
                const int want_top = (int)(hn & 0xffffu), want_left = (int)(hn >> 16);
                uint32_t want = 0;
#pragma unroll
                for (int w = 0; w < 3; w++)
                    if (lane + 64 * w < want_top) want |= 1u << w;
                if (lane < want_left) want |= 8u;
                const uint32_t poll = valid & ~done & want;
                uint32_t hv[4];
#pragma unroll
                for (int w = 0; w < 4; w++) hv[w] = ((poll >> w) & 1u) ? border_load(wptr(w)) : 0u;
#pragma unroll
                for (int w = 0; w < 4; w++)
                    if (((poll >> w) & 1u) && (hv[w] >> 16) == tag) {
                        L.tile[w == 3 ? kTopW + lane * kRowW : lane + 64 * w] = (uint16_t)(hv[w] & 0xffffu);
                        done |= 1u << w;
                    }
                // contiguous prefixes of the row above / left column now present
                const uint32_t miss = valid & ~done;
                int top_have = all_top, left_have = all_left;
#pragma unroll
                for (int w = 0; w < 3; w++) {
                    const uint64_t m = __ballot((miss >> w) & 1u);
                    if (m) top_have = min(top_have, 64 * w + (int)__builtin_ctzll(m));
                }
                {
                    const uint64_t m = __ballot((miss >> 3) & 1u);
                    if (m) left_have = min(left_have, (int)__builtin_ctzll(m));
                }
                // TBs whose reads are all covered
                int nr = ready;
                for (;;) {
                    const int i = nr + lane;
                    const uint32_t pn = i < ntb ? L.pneed[i] : 0u;
                    const bool ok = i < ntb && (int)(pn & 0xffffu) <= top_have && (int)(pn >> 16) <= left_have;
                    const uint64_t okm = __ballot(ok);
                    const int cnt = okm == ~0ull ? 64 : (int)__builtin_ctzll(~okm);
                    nr += cnt;
                    if (cnt < 64 || nr >= ntb) break;
                }
                const bool progress = nr > ready;
                if (progress) {
                    ready = nr;
                    lds_release();
                    if (lane == 0) __hip_atomic_store(&L.bready, (uint32_t)ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                if (ready >= ntb) break;
                if (progress) continue;
                if (++spins > kSpinLimit || ((spins & 1023u) == 0 && border_load(err) != 0)) {
                    // give up: release the prediction wave (the error flag fails the run)
                    atomicOr(err, 1u);
                    lds_release();
                    if (lane == 0) __hip_atomic_store(&L.bready, (uint32_t)ntb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    break;
                }
                __builtin_amdgcn_s_sleep(8);  // back off: the words come from units still running
            }
        } else {
            // ---- prediction ----
            // inter samples on the unit's bottom row and right column are
            // final already: publish them before any intra TB
            if (U.inter_bottom | U.inter_right) {
                if ((U.inter_bottom >> lane) & 1ull)
                    border_store(bw + lane, tagw | L.tile[kTopW + (ctbh - 1) * kRowW + lane + 1]);
                if ((U.inter_right >> lane) & 1ull)
                    border_store(bw + kMaxCtb + lane, tagw | L.tile[kTopW + lane * kRowW + ctbw]);
            }
            // the descriptor of the next TB and, for TBs of up to 8 x 8, its
            // reference words and residuals are loaded while the current TB runs
            uint32_t cd0 = 0, cd1 = 0, cd2 = 0, cw = 0, cl0 = 0, cl2 = 0;
            int32_t cr = 0;
            int cti = kSink;
            // the lane's sample in a 4 x 4 TB (lane % 16) and an 8 x 8 TB,
            // its tile and residual-block offsets from the TB origin
            const int lane16 = lane >> 4;
            const int x4 = lane & 3, y4 = (lane >> 2) & 3, x8 = lane & 7, y8 = lane >> 3;
            const int lt4 = y4 * kRowW + x4, lt8 = y8 * kRowW + x8;
            const int lr4 = y4 * kMaxCtb + x4, lr8 = y8 * kMaxCtb + x8;
            // TBs run in steps: a group of up to four independent 4 x 4 TBs
            // (size in the leader's descriptor bits 22-24), or one TB.  The
            // first TB's descriptor of the next step (nr*) and the descriptor
            // of the lane's TB in it (nl*) are loaded one step ahead of their
            // use, its reference words and residuals while the current step
            // runs.  Only the descriptor words used are loaded: the register
            // of an unused word would hold up the next load into it.
            uint32_t cd3 = 0, nr0 = 0, nr1 = 0, nr2 = 0, nr3 = 0, nl0 = 0, nl2 = 0;
            const uint32_t* dw = reinterpret_cast<const uint32_t*>(L.desc);
            if (ntb > 0) {
                cd0 = uni(dw[0]);
                cd1 = uni(dw[1]);
                cd2 = uni(dw[2]);
                cd3 = uni(dw[3]);
                const int li = lane_desc(0, step_size(cd0) > 1 ? 3 : 0, lane16);
                cl0 = dw[4 * li];
                cl2 = dw[4 * li + 2];
                step_prefetch(L, resb, cd0, cd1, cl0, lane, lt4, lt8, lr4, lr8, cw, cr, cti);
                const int ni = step_end(cd3), nli = lane_desc(ni, next_step_spread(cd0), lane16);
                nr0 = dw[4 * ni];
                nr1 = dw[4 * ni + 1];
                nr2 = dw[4 * ni + 2];
                nr3 = dw[4 * ni + 3];
                nl0 = dw[4 * nli];
                nl2 = dw[4 * nli + 2];
            }
            int ready = 0;
            for (int i = 0; i < ntb;) {
                const int g = step_size(cd0);
                if (i + g > ready) {
                    uint32_t spins = 0;
                    for (;;) {
                        ready = (int)uni(__hip_atomic_load(&L.bready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
                        if (ready >= i + g) break;
                        if (++spins > (kSpinLimit << 4)) {
                            atomicOr(err, 1u);
                            ready = ntb;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    lds_acquire();
                }
                if ((cd0 >> 21) & 1) {
                    // a run of fast steps, each ready and followed by a fast
                    // one, unrolled by two over two register sets
                    FastUni Au{i, cd0, cd1, cd2, cd3}, Bu;
                    FastLane Av{cw, cr, cti, cl0, cl2, nr0, nr1, nr2, nr3, nl0, nl2}, Bv;
                    auto step = [&](const FastUni& cu, const FastLane& c, FastUni& nu, FastLane& n) -> bool {
                        const bool mixed = (cu.d0 >> 29) & 1;
                        // this step's angular reference pair is read first
                        // (the reads follow the previous step's write); the
                        // next step's loads, the planar / H-V pair and the
                        // loop control are issued while it is in flight.
                        // Past the last TB the loads read the descriptors
                        // after it and the slots they name, and their
                        // values go unused.
                        StepRefs q;
                        q.r0 = L.tile[c.w & 0x1fffu];
                        q.r1 = L.tile[(c.w >> 13) & 0x1fffu];
                        __builtin_amdgcn_sched_barrier(0);
                        nu.i = step_end(cu.d3);
                        nu.d0 = uni(c.n0);
                        nu.d1 = uni(c.n1);
                        nu.d2 = uni(c.n2);
                        nu.d3 = uni(c.n3);
                        n.l0 = c.nl0;
                        n.l2 = c.nl2;
                        step_prefetch_fast(L, resb, nu.d0, nu.d1, n.l0, lane, lt4, lt8, lr4, lr8, n.w, n.r, n.ti);
                        const int key = step_key(nu.d3), nn = key & 0x1ff;
                        const uint4 m = L.desc[nn];
                        n.n0 = m.x;
                        n.n1 = m.y;
                        n.n2 = m.z;
                        n.n3 = m.w;
                        const int nli = lane_desc(nn, next_step_spread(nu.d0), lane16);
                        n.nl0 = dw[4 * nli];
                        n.nl2 = dw[4 * nli + 2];
                        if (mixed) {
                            q.ru0 = L.tile[c.l2 & 0x1fffu];
                            q.ru1 = L.tile[(c.l2 >> 13) & 0x1fffu];
                        }
                        __builtin_amdgcn_sched_barrier(0);
                        const bool b8 = (cu.d0 >> 12) & 1;
                        const int x = b8 ? x8 : x4, y = b8 ? y8 : y4;
                        const bool active = step_active(cu.d0, lane16);
                        intra_step_fast(L.tile, c.l0, q, c.w, c.r, b8, mixed, x, y, c.ti, active, maxv);
                        if ((cu.d0 >> 28) & 1) intra_step_publish(L.tile, c.l0, b8, x, y, c.ti, active, bw, tagw);
                        // the next step runs here if it is fast (the
                        // descriptor after the last TB is not) and ready
                        return key <= ready;
                    };
                    auto leave = [&](const FastUni& u, const FastLane& v) {
                        i = u.i;
                        cd0 = u.d0;
                        cd1 = u.d1;
                        cd2 = u.d2;
                        cd3 = u.d3;
                        cw = v.w;
                        cr = v.r;
                        cti = v.ti;
                        cl0 = v.l0;
                        cl2 = v.l2;
                        nr0 = v.n0;
                        nr1 = v.n1;
                        nr2 = v.n2;
                        nr3 = v.n3;
                        nl0 = v.nl0;
                        nl2 = v.nl2;
                    };
                    for (;;) {
                        if (!step(Au, Av, Bu, Bv)) {
                            leave(Bu, Bv);
                            break;
                        }
                        if (!step(Bu, Bv, Au, Av)) {
                            leave(Au, Av);
                            break;
                        }
                    }
                } else {
                    const int next = step_end(cd3);
                    const uint32_t nd0 = uni(nr0), nd1 = uni(nr1), nd2 = uni(nr2), nd3 = uni(nr3);
                    uint32_t nw = 0, mr0 = nr0, mr1 = nr1, mr2 = nr2, mr3 = nr3, ml0 = nl0, ml2 = nl2;
                    int32_t nrr = 0;
                    int nti = kSink;
                    if (next < ntb) {
                        step_prefetch(L, resb, nd0, nd1, nl0, lane, lt4, lt8, lr4, lr8, nw, nrr, nti);
                        const int nn = step_end(nd3), nli = lane_desc(nn, next_step_spread(nd0), lane16);
                        mr0 = dw[4 * nn];
                        mr1 = dw[4 * nn + 1];
                        mr2 = dw[4 * nn + 2];
                        mr3 = dw[4 * nn + 3];
                        ml0 = dw[4 * nli];
                        ml2 = dw[4 * nli + 2];
                    }
                    switch ((cd0 >> 12) & 3) {
                        case 0: intra_tb<2, R>(L, resb, cd0, cd1, cd2, true, cw, cr, lane, bd, maxv, bw, tagw); break;
                        case 1: intra_tb<3, R>(L, resb, cd0, cd1, cd2, true, cw, cr, lane, bd, maxv, bw, tagw); break;
                        case 2: intra_tb<4, R>(L, resb, cd0, cd1, cd2, false, 0, 0, lane, bd, maxv, bw, tagw); break;
                        default: intra_tb<5, R>(L, resb, cd0, cd1, cd2, false, 0, 0, lane, bd, maxv, bw, tagw); break;
                    }
                    i = next;
                    cd0 = nd0;
                    cd1 = nd1;
                    cd2 = nd2;
                    cd3 = nd3;
                    cw = nw;
                    cr = nrr;
                    cti = nti;
                    cl0 = nl0;
                    cl2 = nl2;
                    nr0 = mr0;
                    nr1 = mr1;
                    nr2 = mr2;
                    nr3 = mr3;
                    nl0 = ml0;
                    nl2 = ml2;
                }
            }
        }
        __syncthreads();
        // the whole CTB to the plane, four samples per store (plane rows
        // start at multiples of 4 samples)
        for (int e = 4 * tid; e < ctbw * ctbh; e += 4 * kIntraThreads) {
            const int ry = e >> ctbw_log2, rx = e - (ry << ctbw_log2);
            const int px = ux0 + rx, py = uy0 + ry;
            if (px < pw && py < ph) {
                const uint2 q = *reinterpret_cast<const uint2*>(L.tile + kTopW + 1 + ry * kRowW + rx);
                store_quad(pl + (size_t)py * pw + px, q.x & 0xffffu, q.x >> 16, q.y & 0xffffu, q.y >> 16);
            }
        }
        __syncthreads();
    }
}
