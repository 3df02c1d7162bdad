// intra.hip — K2: availability masks and the intra prediction wavefront
// (included by kernels.hip inside namespace hpvcd).

// Optional cycle profiling of the wavefront (build with -DHPVCD_INTRA_PROFILE):
// per unit and wave, shader-clock cycles spent in each phase.  Only records
// timestamps; the computation is unchanged.
#ifdef HPVCD_INTRA_PROFILE
__device__ uint64_t* g_intra_prof = nullptr;
#define HPVCD_PROF_DECL uint64_t prof_[8] = {0, 0, 0, 0, 0, 0, 0, 0}; uint64_t prof_t_ = clock64();
#define HPVCD_PROF(k) { const uint64_t c_ = clock64(); prof_[k] += c_ - prof_t_; prof_t_ = c_; }
#define HPVCD_PROF_FLUSH(u)                                                                                 \
    if (lane == 0 && g_intra_prof)                                                                          \
        for (int k_ = 0; k_ < 8; k_++) g_intra_prof[((size_t)(u) * 4 + wave) * 8 + k_] = prof_[k_];
#else
#define HPVCD_PROF_DECL
#define HPVCD_PROF(k)
#define HPVCD_PROF_FLUSH(u)
#endif

constexpr int kMaxCtb = 64;
constexpr int kTopW = 2 * kMaxCtb + 1;     // row above the unit: x0-1 .. x0+2*ctb-1
constexpr int kRowW = kMaxCtb + 1;         // body rows: left neighbour + the unit's samples
constexpr int kMaxUnitTbs = 256;
constexpr int kMaxUnitRefs = 256 * 17;    // reference entries of one unit (all 4x4 TBs is the maximum)
constexpr int kBorderWords = 2 * kMaxCtb;  // bottom row + right column per unit
constexpr uint32_t kSpinLimit = 1u << 19;

// Reference availability of one TB as a bit mask over 4-sample groups
// (src/decode.rs:4651 luma_avail, :4669 chroma_avail): bits [0, ngl) are the
// left groups from the bottom up, bit ngl the corner, then the above groups
// left to right.  Availability is constant within a group because a group
// never spans two 4x4 luma cells.  Computed for every TB in parallel before
// the wavefront so that the wavefront does no availability work.
__global__ void __launch_bounds__(256) k_avail(const DevPic* __restrict__ pics, const hpvcd_tb* __restrict__ tbs,
                                               const uint32_t* __restrict__ tb_pic, uint32_t ntbs,
                                               const uint32_t* __restrict__ cell_order,
                                               const uint32_t* __restrict__ cell_info, uint64_t* __restrict__ masks) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntbs) return;
    const hpvcd_tb tb = tbs[t];
    const DevPic& P = pics[tb_pic[t]];
    const int comp = tb.comp;
    const int sw = comp ? P.sub_w_log2 : 0, sh = comp ? P.sub_h_log2 : 0;
    const int pw = comp ? P.cw : P.w, ph = comp ? P.ch : P.h;
    const int n = 1 << tb.log2n, n2 = 2 * n;
    const int gv = 4 >> sh, gh = 4 >> sw, ngl = n2 / gv, nga = n2 / gh;
    const int x0 = tb.x, y0 = tb.y;
    const bool need_info = (P.flags & (HPVCD_PF_TILES | HPVCD_PF_CONSTRAINED_INTRA)) || P.num_slices > 1;
    const uint32_t* ord = cell_order + P.cell_begin;
    const uint32_t* inf = cell_info + P.cell_begin;
    const uint32_t cur = need_info ? inf[((y0 << sh) >> 2) * P.gw + ((x0 << sw) >> 2)] : 0u;
    auto avail = [&](int px, int py) -> bool {
        if (px < 0 || py < 0 || px >= pw || py >= ph) return false;
        const int lx = px << sw, ly = py << sh;
        if (lx >= P.w || ly >= P.h) return false;
        const int c = (ly >> 2) * P.gw + (lx >> 2);
        if (ord[c] >= tb.order) return false;
        if (need_info) {
            const uint32_t i = inf[c];
            if ((P.flags & HPVCD_PF_TILES) && HPVCD_CELL_TILE(i) != HPVCD_CELL_TILE(cur)) return false;
            if (P.num_slices > 1 && HPVCD_CELL_SLICE(i) != HPVCD_CELL_SLICE(cur)) return false;
            if ((P.flags & HPVCD_PF_CONSTRAINED_INTRA) && !(i & HPVCD_CELL_INTRA)) return false;
        }
        return true;
    };
    uint64_t m = 0;
    for (int g = 0; g < ngl; g++)
        if (avail(x0 - 1, y0 + n2 - (g + 1) * gv)) m |= 1ull << g;
    if (avail(x0 - 1, y0 - 1)) m |= 1ull << ngl;
    for (int g = 0; g < nga; g++)
        if (avail(x0 + g * gh, y0 - 1)) m |= 1ull << (ngl + 1 + g);
    masks[t] = m;
}

struct IntraLds {
    uint16_t tile[kTopW + kMaxCtb * kRowW];
    int32_t raw[4][132];
    int32_t ref[4][132];
    hpvcd_tb tb[kMaxUnitTbs];
    uint64_t mask[kMaxUnitTbs];
    uint16_t lev[kMaxUnitTbs + 1];
    uint16_t refoff[kMaxUnitTbs];
    uint16_t refmap[kMaxUnitRefs];
    int16_t angle[33];
    int16_t inv_angle[15];
    uint32_t unit;
};

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

template <typename T, typename R>
__global__ void __launch_bounds__(256) k_intra(const DevPic* __restrict__ pics, const hpvcd_tb* __restrict__ tbs,
                                               const uint64_t* __restrict__ masks, const DevUnit* __restrict__ units,
                                               const uint32_t* __restrict__ ticket_units, uint32_t n_units,
                                               const uint32_t* __restrict__ unit_tbs,
                                               const uint16_t* __restrict__ level_off,
                                               const uint16_t* __restrict__ unit_refoff, T* __restrict__ samples,
                                               const R* __restrict__ res_pool, uint32_t* __restrict__ border,
                                               uint32_t* counter, const uint32_t* gen, uint32_t* err) {
    __shared__ IntraLds L;
    __shared__ R resb[kMaxCtb * kMaxCtb];
    const int tid = threadIdx.x, wave = uni(tid >> 6), lane = tid & 63;
    const uint32_t tag = *gen & 0xffffu;
    if (tid < 33) L.angle[tid] = kIntraAngle[tid];
    if (tid < 15) L.inv_angle[tid] = kInvAngle[tid];

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
        const int pw = plane ? P.cw : P.w, ph = plane ? P.ch : P.h;
        const int ux0 = U.cx * ctbw, uy0 = U.cy * ctbh;     // unit origin (plane samples)
        const int bd = plane ? P.bd_c : P.bd;
        const int32_t maxv = (1 << bd) - 1;
        T* pl = samples + P.plane_off[plane];
        const R* rp = res_pool + P.res_off[plane];
        const int ntb = U.ntb, nlev = U.nlevels;
        HPVCD_PROF_DECL

        // stage what does not depend on neighbours: TB records, masks,
        // levels and residuals
        for (int e = tid; e < ntb; e += 256) {
            const uint32_t t = unit_tbs[U.tb_list + e];
            L.tb[e] = tbs[t];
            L.mask[e] = masks[t];
            L.refoff[e] = unit_refoff[U.tb_list + e];
        }
        for (int e = tid; e <= nlev; e += 256) L.lev[e] = level_off[U.level_list + e];
        for (int e = tid; e < ctbw * ctbh; e += 256) {
            int ry = e / ctbw, rx = e - ry * ctbw;
            int px = ux0 + rx, py = uy0 + ry;
            resb[ry * kMaxCtb + rx] = (px < pw && py < ph) ? rp[(size_t)py * pw + px] : (R)0;
        }
        __syncthreads();
        // Reference address maps of every TB of the unit: tile address of the
        // substituted reference for each s-index (src/intra.rs:84), 0xffff for
        // the neutral value.  Availability does not depend on sample values,
        // so this runs before the neighbours are ready and is off the
        // dependency chain.  s-index 0..2n-1 runs up the left column, 2n is
        // the corner, 2n+1..4n runs along the row above.
        {
            const int lgv = 2 - sh, lgh = 2 - sw;
            for (int t = wave; t < ntb; t += 4) {
                const hpvcd_tb tb = L.tb[t];
                const uint64_t amask = uni64(L.mask[t]);
                const int n = 1 << uni((int)tb.log2n), n2 = 2 * n, nref = 4 * n + 1;
                const int x0 = uni((int)tb.x) - ux0, y0 = uni((int)tb.y) - uy0;
                const int ngl = n2 >> lgv;
                const int roff = uni((int)L.refoff[t]);
                const int first_b = amask ? __builtin_ctzll(amask) : -1;
                const int first = first_b < 0 ? -1
                                  : (first_b < ngl ? first_b << lgv
                                                   : (first_b == ngl ? n2 : n2 + 1 + ((first_b - ngl - 1) << lgh)));
                for (int k = lane; k < nref; k += 64) {
                    const int b = k < n2 ? k >> lgv : (k == n2 ? ngl : ngl + 1 + ((k - n2 - 1) >> lgh));
                    int src = k;
                    if (!((amask >> b) & 1ull)) {
                        const uint64_t below = amask & ((1ull << b) - 1ull);
                        if (below) {
                            const int bb = 63 - __builtin_clzll(below);
                            src = bb < ngl ? ((bb + 1) << lgv) - 1 : (bb == ngl ? n2 : n2 + ((bb - ngl) << lgh));
                        } else {
                            src = first;
                        }
                    }
                    int addr = 0xffff;
                    if (src >= 0)
                        addr = src < n2 ? kTopW + (y0 + n2 - 1 - src) * kRowW + x0
                                        : (src == n2 ? (y0 == 0 ? x0 : kTopW + (y0 - 1) * kRowW + x0)
                                                     : (y0 == 0 ? x0 + src - n2 : kTopW + (y0 - 1) * kRowW + x0 + src - n2));
                    L.refmap[roff + k] = (uint16_t)addr;
                }
            }
        }
        HPVCD_PROF(0)
        // neighbour samples from the published borders: the row above
        // (corner + two unit widths) and the column to the left
        for (int e = tid; e < 2 * ctbw + 1; e += 256) {
            int px = ux0 - 1 + e, py = uy0 - 1;
            uint16_t v = 0;
            if (py >= 0 && px >= 0 && px < pw) {
                int ncx = px / ctbw;
                uint32_t nu = ui - (uint32_t)U.cx + (uint32_t)ncx - (uint32_t)P.ctb_cols;  // unit (ncx, cy - 1)
                v = border_wait(border + (size_t)nu * kBorderWords + (px - ncx * ctbw), tag, err);
            }
            L.tile[e] = v;
        }
        for (int e = tid; e < ctbh; e += 256) {
            int px = ux0 - 1, py = uy0 + e;
            uint16_t v = 0;
            if (px >= 0 && py < ph) v = border_wait(border + (size_t)(ui - 1) * kBorderWords + kMaxCtb + e, tag, err);
            L.tile[kTopW + e * kRowW] = v;
        }
        __syncthreads();

        HPVCD_PROF(1)
        const uint32_t pflags = uni(P.flags);
        const bool filt_luma = plane == 0 || uni(P.chroma_idc) == 3;
        const bool smoothing = !(pflags & HPVCD_PF_INTRA_SMOOTHING_DISABLED);
        const bool strong_on = plane == 0 && (pflags & HPVCD_PF_STRONG_INTRA_SMOOTHING);
        const bool edge = plane == 0 && !(pflags & HPVCD_PF_INTRA_BOUNDARY_FILTER_OFF);
        const int32_t neutral = 1 << (bd - 1);
        const int lgv = 2 - sh, lgh = 2 - sw;                  // log2 of the availability group sizes
        for (int lv = 0; lv < nlev; lv++) {
            const int lo = uni((int)L.lev[lv]), hi = uni((int)L.lev[lv + 1]);
            for (int j = lo + wave; j < hi; j += 4) {
                const hpvcd_tb tb = L.tb[j];
                // TB fields are the same in every lane: keep them scalar
                const int log2n = uni((int)tb.log2n), mode = uni((int)tb.mode);
                const int x0 = uni((int)tb.x) - ux0, y0 = uni((int)tb.y) - uy0;
                const bool has_res = (uni((int)tb.flags) & TBX_HAS_RES) != 0;
                const int n = 1 << log2n, n2 = 2 * n, nref = 4 * n + 1;
                const int roff = uni((int)L.refoff[j]);
                // substituted reference at s-index k (map built in the prologue)
                auto sval = [&](int k) -> int32_t {
                    const int a = L.refmap[roff + k];
                    return a == 0xffff ? neutral : (int32_t)L.tile[a];
                };
                HPVCD_PROF(5)
                // reference filtering decision (src/intra.rs:151)
                bool do_filter = false;
                if (smoothing && filt_luma && n != 4 && mode != 1) {
                    int d26 = mode > 26 ? mode - 26 : 26 - mode, d10 = mode > 10 ? mode - 10 : 10 - mode;
                    int min_dist = d26 < d10 ? d26 : d10;
                    int thresh = n == 8 ? 7 : (n == 16 ? 1 : (n == 32 ? 0 : 8));
                    do_filter = mode == 0 || min_dist > thresh;
                }
                // angular modes other than pure H/V read an extended main
                // reference ref[k], k in [-n, 2n] (stored at k + n):
                // main[k] for k >= 0, the projected side reference below
                const bool ext = mode >= 2 && mode != 10 && mode != 26;
                const bool vertical = mode >= 18;
                const int angle = mode >= 2 ? L.angle[mode - 2] : 0;
                const int inv = (mode >= 11 && mode <= 25) ? L.inv_angle[mode - 11] : 0;
                auto ext_sidx = [&](int k) -> int {
                    int off = k;
                    if (k < 0) {
                        int idx = (k * inv + 128) >> 8;
                        off = -(idx > n2 ? n2 : (idx < 0 ? 0 : idx));
                    }
                    return vertical ? n2 + off : n2 - off;
                };
                int32_t* S = L.raw[wave];
                int32_t* F = L.ref[wave];
                int32_t dcsum = 0;
                if (!do_filter) {
                    // one pass: substituted references straight into the
                    // layout the predictor reads
                    if (ext) {
                        for (int k = lane; k <= 3 * n; k += 64) F[k] = sval(ext_sidx(k - n));
                    } else {
                        for (int i = lane; i < nref; i += 64) {
                            const int32_t v = sval(i);
                            F[i] = v;
                            if ((i >= n2 - n && i < n2) || (i > n2 && i <= n2 + n)) dcsum += v;
                        }
                    }
                    wave_sync();
                } else {
                    for (int i = lane; i < nref; i += 64) S[i] = sval(i);
                    wave_sync();
                    bool strong = false;
                    const int32_t bl = S[0], tl = S[n2], tr = S[4 * n];
                    if (strong_on && n == 32) {
                        const int thr = 1 << (bd - 5);
                        const int a = bl + tl - 2 * S[n], b = tr + tl - 2 * S[3 * n];
                        strong = (a < 0 ? -a : a) < thr && (b < 0 ? -b : b) < thr;
                    }
                    auto fval = [&](int i) -> int32_t {
                        if (i == 0 || i == 4 * n) return S[i];
                        if (strong) {
                            if (i == n2) return tl;
                            if (i > n2) { const int k = i - n2; return ((n2 - k) * tl + k * tr + n) >> 6; }
                            const int k = n2 - i;
                            return ((n2 - k) * tl + k * bl + n) >> 6;
                        }
                        return (S[i - 1] + 2 * S[i] + S[i + 1] + 2) >> 2;
                    };
                    if (ext) {
                        for (int k = lane; k <= 3 * n; k += 64) F[k] = fval(ext_sidx(k - n));
                    } else {
                        for (int i = lane; i < nref; i += 64) F[i] = fval(i);
                    }
                    wave_sync();
                }
                HPVCD_PROF(2)
                // prediction (src/intra.rs:257); four outputs per lane per
                // pass so their LDS reads overlap
                int32_t dc = 0;
                if (mode == 1) {
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) dcsum += __shfl_xor(dcsum, off);
                    dc = (dcsum + n) >> (log2n + 1);
                }
                const int nn = n * n;
                for (int base = 0; base < nn; base += 256) {
                    int32_t vals[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        vals[q] = 0;
                        if (base + q * 64 >= nn) continue;  // wave-uniform: small TBs use one pass
                        const int p = base + q * 64 + lane;
                        const int x = p & (n - 1), y = (p >> log2n) & (n - 1);
                        int32_t v;
                        if (ext) {
                            const int a = vertical ? y : x, c = vertical ? x : y;
                            const int pos = (a + 1) * angle, ii = pos >> 5, frac = pos & 31;
                            const int32_t r0 = F[c + ii + 1 + n];
                            v = frac == 0 ? r0 : ((32 - frac) * r0 + frac * F[c + ii + 2 + n] + 16) >> 5;
                            v = clampi(v, 0, maxv);
                        } else if (mode == 26) {
                            v = F[n2 + 1 + x];
                            if (edge && n < 32 && x == 0) v = clampi(F[n2 + 1] + ((F[n2 - 1 - y] - F[n2]) >> 1), 0, maxv);
                        } else if (mode == 10) {
                            v = F[n2 - 1 - y];
                            if (edge && n < 32 && y == 0) v = clampi(F[n2 - 1] + ((F[n2 + 1 + x] - F[n2]) >> 1), 0, maxv);
                        } else if (mode == 0) {
                            const int32_t h = (n - 1 - x) * F[n2 - 1 - y] + (x + 1) * F[n2 + n + 1];
                            const int32_t vv = (n - 1 - y) * F[n2 + 1 + x] + (y + 1) * F[n2 - n - 1];
                            v = (h + vv + n) >> (log2n + 1);
                        } else {
                            v = dc;
                            if (edge && n < 32) {
                                if (x == 0 && y == 0) v = (F[n2 - 1] + 2 * dc + F[n2 + 1] + 2) >> 2;
                                else if (y == 0) v = (F[n2 + 1 + x] + 3 * dc + 2) >> 2;
                                else if (x == 0) v = (F[n2 - 1 - y] + 3 * dc + 2) >> 2;
                            }
                        }
                        if (has_res) v = clampi(wadd(v, (int32_t)resb[(y0 + y) * kMaxCtb + x0 + x]), 0, maxv);
                        vals[q] = v;
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        if (base + q * 64 >= nn) continue;
                        const int p = base + q * 64 + lane;
                        const int x = p & (n - 1), y = (p >> log2n) & (n - 1);
                        if (ux0 + x0 + x < pw && uy0 + y0 + y < ph)
                            L.tile[kTopW + (y0 + y) * kRowW + x0 + x + 1] = (uint16_t)vals[q];
                    }
                }
                wave_sync();
                HPVCD_PROF(3)
            }
            __syncthreads();
            HPVCD_PROF(4)
        }
        HPVCD_PROF_FLUSH(ui)
        // publish: the whole CTB to the plane, bottom row and right column
        // as tagged border words
        for (int e = tid; e < ctbw * ctbh; e += 256) {
            int ry = e / ctbw, rx = e - ry * ctbw;
            int px = ux0 + rx, py = uy0 + ry;
            if (px < pw && py < ph) pl[(size_t)py * pw + px] = (T)L.tile[kTopW + ry * kRowW + rx + 1];
        }
        uint32_t* bw = border + (size_t)ui * kBorderWords;
        for (int e = tid; e < ctbw; e += 256)
            border_store(bw + e, (tag << 16) | L.tile[kTopW + (ctbh - 1) * kRowW + e + 1]);
        for (int e = tid; e < ctbh; e += 256)
            border_store(bw + kMaxCtb + e, (tag << 16) | L.tile[kTopW + e * kRowW + ctbw]);
        __syncthreads();
    }
}

