// walker.h — one picture's coding-tree syntax, walked once per picture and
// turned into reconstruction records.
//
// The walker owns the parse-side derivations the reference decoder runs while
// it parses (src/decode.rs): CTB and coding-quadtree order (:1319, :2495),
// the coding unit (:2585), PCM (:2768), intra mode derivation (:3204 MPM,
// :3248 chroma), the transform tree and unit (:3315, :3478, :3619 ACT),
// chroma placement (:5007), QP groups and prediction (:2509, :3084, :5768),
// deblocking edge marking (:3018) and the per-4x4 side maps, SAO merge
// (:1507).  Every syntax decision comes from a Source: the seeded generator
// (random draws, optionally CABAC-encoded) or the CABAC parser.  Both build
// records through this same code, so a parse of a generated stream yields
// the generator's records exactly when every decision is recovered.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <functional>
#include <type_traits>
#include <vector>

#include "../../include/hpvcd.h"
#include "cabac.h"
#include "motion.h"

namespace hpvcd {

// Picture-level syntax parameters (from the generator's configuration or
// from the active SPS / PPS / slice header).
struct PicParams {
    int w = 0, h = 0, chroma_idc = 1, bd = 8, bd_c = 8;
    int log2_ctb = 6, log2_min_cb = 3, log2_min_tb = 2, log2_max_tb = 5;
    int max_depth_intra = 1, max_depth_inter = 1;
    bool tqb_enabled = false, transform_skip = false;
    int log2_max_ts = 2;
    bool implicit_rdpcm = false, explicit_rdpcm = false, cross_component = false, extended_precision = false;
    bool persistent_rice = false, ts_context = false, bypass_alignment = false;
    bool sign_hiding = false, act = false;
    int act_qp[3] = {-5, -5, -3};
    bool pcm = false;
    int pcm_bd = 8, pcm_bd_c = 8, log2_min_pcm = 3, log2_max_pcm = 5;
    bool cu_qp_delta = false;
    int diff_cu_qp_delta_depth = 0;
    int chroma_qp_off[2] = {0, 0};   // PPS + slice offsets of Cb, Cr (dequantisation)
    bool sao = false, sao_luma = false, sao_chroma = false;
    bool amp = false, constrained_intra = false;
    bool wpp = false;                // entropy_coding_sync_enabled_flag: a substream per CTB row of a tile
};

// The inter parameters of one P or B slice: reference lists, the record
// rows of its entries (list 0, then list 1: picture and weights), and the
// merge / AMVP / temporal parameters of its header.
struct InterSlice {
    int slice_type = 1;               // 0 B, 1 P
    bool cabac_init = false;          // cabac_init_flag
    std::vector<RefPic> list[2];
    std::vector<hpvcd_ref> rows;      // list 0 entries, then list 1 entries
    bool tmvp = false;
    bool col_from_l0 = true;
    int col_ref_idx = 0;
    const ColField* col = nullptr;    // the collocated picture's motion (resolved by the caller)
    int max_merge = 5;
    int par_mrg_level = 2;
    bool mvd_l1_zero = false;
};

// inter part modes (H.265 Table 7-10 order)
enum PartMode { PART_2Nx2N = 0, PART_2NxN, PART_Nx2N, PART_NxN, PART_2NxnU, PART_2NxnD, PART_nLx2N, PART_nRx2N };

constexpr uint8_t kModePlanar = 0, kModeDC = 1;

inline uint64_t splitmix64_step(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    uint64_t next() { return splitmix64_step(s); }
    bool bern(float p) { return (double)(next() >> 11) * (1.0 / 9007199254740992.0) < (double)p; }
    uint32_t uni(uint32_t n) { return (uint32_t)(next() % n); }
    int sint(int lo, int hi) { return lo + (int)uni((uint32_t)(hi - lo + 1)); }
};

// Decisions of one picture's syntax.  Each is asked only where the syntax
// element is present; `ctx` arguments are the context increments the
// walker derived.
// A non-owning reference to a callable (a lambda at a call site), called
// through one function pointer: no allocation per call, unlike std::function.
template <typename Sig> class FnRef;
template <typename R, typename... A> class FnRef<R(A...)> {
    void* obj_;
    R (*call_)(void*, A...);

public:
    template <typename F>
    FnRef(F&& f)  // NOLINT: implicit from a callable that outlives the call
        : obj_((void*)&f), call_([](void* o, A... a) -> R { return (*(std::remove_reference_t<F>*)o)(a...); }) {}
    R operator()(A... a) const { return call_(obj_, a...); }
};

struct Source {
    virtual ~Source() = default;
    // init_type: 0 for I slices, 1 or 2 for P / B (9.3.2.2 with cabac_init_flag)
    virtual void begin_slice(int slice_qp, int init_type) = 0;
    // A new substream that is not the slice's first: a tile, or a CTB row of
    // a tile with WPP.  The contexts start from their initial values, or,
    // when sync_row >= 0, from the ones wpp_save stored for that CTB row
    // (9.3.1: the state after the second CTB of the row above).
    virtual void begin_substream(int sync_row) = 0;
    // WPP storage after the second CTB of a CTB row of a tile (9.3.2.4)
    virtual void wpp_save(int row) { (void)row; }
    // end_of_slice_segment_flag / end_of_subset_one_bit after a CTB
    virtual int end_ctb(bool last_of_slice, bool last_of_tile) = 0;
    virtual void sao(hpvcd_sao& s, bool left_avail, bool up_avail, const hpvcd_sao* left, const hpvcd_sao* up) = 0;
    virtual bool split_cu(int ctx_inc) = 0;
    virtual bool tqb() = 0;
    virtual bool intra_nxn() = 0;
    virtual bool pcm_flag() = 0;
    // PCM samples of one CU, luma then Cb then Cr: counts and bit depths
    virtual void pcm_samples(int n_luma, int n_chroma, int bd_y, int bd_c, uint32_t* out) = 0;
    // luma modes of npu PUs: mpm(k, list) fills PU k's candidates from the
    // modes committed so far; commit(k, mode) records PU k's mode
    virtual void luma_modes(int npu, FnRef<void(int, uint8_t*)> mpm, FnRef<void(int, uint8_t)> commit) = 0;
    virtual int chroma_mode_syntax() = 0;   // intra_chroma_pred_mode 0..4
    virtual bool split_tf(int log2) = 0;
    virtual bool cbf_chroma(int depth) = 0;
    virtual bool cbf_luma(int depth) = 0;
    virtual bool act_flag() = 0;
    virtual int qp_delta() = 0;             // CuQpDeltaVal
    virtual int res_scale(int c) = 0;       // ResScaleVal of Cb (0) / Cr (1)
    virtual void residual(const ResidualParams& p, int32_t* levels, ResidualOut& r) = 0;
    // inter syntax (P and B slices)
    virtual bool cu_skip(int ctx_inc) = 0;
    virtual bool pred_mode_intra() = 0;                       // pred_mode_flag
    // part_mode of an inter CU (PartMode); min_size: log2 == MinCbLog2SizeY,
    // amp: AMP partitions are allowed
    virtual int part_mode_inter(int log2, bool min_size, bool amp) = 0;
    virtual bool merge_flag() = 0;
    virtual int merge_idx(int max_cand) = 0;                  // coded when MaxNumMergeCand > 1
    virtual int inter_pred_idc(int ct_depth, bool no_bi) = 0; // 0 PRED_L0, 1 PRED_L1, 2 PRED_BI
    virtual int ref_idx(int num_active) = 0;                  // coded when more than one entry
    // mvd_coding (unless mvd_zero: MvdL1 inferred zero) and mvp_lX_flag of
    // one list of a PU; cand are its two AMVP predictors
    virtual void mv_syntax(const PuGeom& pu, int list, const int16_t cand[2][2], bool mvd_zero, int16_t mvd[2],
                           int& mvp_flag) = 0;
    virtual bool rqt_root_cbf() = 0;
};

// One picture's records as the walker emits them (TBs in decode order, CTBs
// in tile-scan order until finish() reorders them).
struct PicRecords {
    std::vector<hpvcd_tb> tbs;
    std::vector<uint32_t> coeffs, cells;
    std::vector<hpvcd_sao> sao;
    std::vector<hpvcd_pu> pus;
    std::vector<hpvcd_ref> refs;
    int num_slices = 1, num_tiles = 1;
    ColField col;   // the picture's motion as a later picture's collocated motion
};

class PicWalker : private Neighbours {
public:
    PicWalker(const PicParams& P, Source& src, PicRecords& out, int32_t poc = 0);
    // CTB layout: per raster CTB its slice and tile; the walk visits CTBs in
    // ts_to_rs order.  slice_qp per slice.
    void set_layout(const std::vector<uint8_t>& ctb_slice, const std::vector<uint8_t>& ctb_tile);
    // chroma QP offsets (PPS + slice) of the slice that follows
    void set_chroma_offsets(int cb, int cr) {
        chroma_off[0] = cb;
        chroma_off[1] = cr;
    }
    // the inter parameters of the slice that follows (null: an I slice); the
    // slice's reference rows join the picture's unless equal to the last ones
    void set_inter(const InterSlice* s);
    // one CTB (SAO, coding quadtree, end flags); first_in_slice / first_in_tile
    // (the first CTB of a substream: a tile, or with WPP a CTB row of a tile)
    // reset the QP predictor; last_in_tile ends a substream; returns the
    // end_of_slice_segment_flag
    int ctb(int rx, int ry, bool first_in_slice, bool first_in_tile, bool last_in_slice, bool last_in_tile,
            int slice_qp);
    // cells (QP, slice, tile packed), TB order, inter bS, collocated motion
    void finish();
    bool failed() const { return fail_; }

private:
    PicParams P;
    Source& src;
    PicRecords& out;
    int32_t poc_;
    int w, h, gw, gh, sub_w, sub_h, ctb_cols, ctb_rows;
    int qp_y, slice_qp = 0, qp_prev = 0, qp_bd_off = 0, log2_qg = 6;
    int chroma_off[2] = {0, 0};
    bool qp_delta_coded = false;
    bool fail_ = false;
    std::vector<uint32_t> cell;
    std::vector<int8_t> cell_qp;
    std::vector<uint8_t> mode_y, ct_depth;
    std::vector<uint8_t> ctb_slice, ctb_tile;
    int cur_slice = 0, cur_tile = 0;
    uint32_t coeff_base = 0;
    // current CU
    uint8_t cu_chroma_modes[4];
    bool cu_chroma_dm[4];
    int cu_x0 = 0, cu_y0 = 0, cu_pu = 1;
    bool cu_tqb = false, cu_inter = false;
    // inter slices
    const InterSlice* inter_ = nullptr;   // the current slice's (null: I slice)
    const InterSlice* last_rows_ = nullptr;
    MotionParams mp_;
    int ref_base_ = 0;                    // record row of the slice's first list-0 entry
    bool have_motion_ = false;            // an inter slice was seen: the maps below exist
    std::vector<MotionCell> motion;
    std::vector<uint8_t> decoded, tu_v, tu_h, pu_v, pu_h, nz;
    std::vector<int32_t> lv;       // one TB's levels (all zero between TBs)
    std::vector<uint16_t> nzpos;   // the positions a source lists as nonzero

    size_t g_of(int x, int y) const { return (size_t)(y >> 2) * gw + (x >> 2); }
    bool same_region(int x, int y) const;
    // Neighbours
    bool available(int x, int y) const override;
    const MotionCell& motion_at(int x, int y) const override { return motion[g_of(x, y)]; }
    void mark_tu_edges(int x0, int y0, int size);
    void finalize_bs();
    int skip_ctx(int x0, int y0) const;
    MotionCell cell_of(const MergeCand& c) const;
    MotionCell merge_motion(const PuGeom& g, int merge_idx) const;
    MotionCell prediction_unit(const PuGeom& g, int log2_cb, bool& merge);
    void store_pu(const PuGeom& g, const MotionCell& m);
    void inter_tiles(int x0, int y0, int log2);
    void inter_cu(int x0, int y0, int log2, bool skip);
    void mark_block_edges(int x0, int y0, int size, int bs);
    void fill_cells(int x0, int y0, int size, uint32_t bit);
    void set_mode(int x0, int y0, int size, uint8_t mode);
    uint8_t neighbour_mode(int x, int y) const;
    void mpm_list(int x0, int y0, uint8_t* mpm) const;
    int split_ctx(int x0, int y0, int depth) const;
    int predict_qp(int xqg, int yqg);
    int qpy_from_delta(int pred, int delta) const;
    void maybe_qp_delta(bool residual);
    int qp_prime(int qp, int bitd) const;
    static int qpc_table(int qpi, int chroma_idc, int bitd);
    int tb_qp(int comp, bool act);
    hpvcd_tb make_tb(int x, int y, int log2n, int comp, uint8_t mode, bool cbf, bool act = false);
    bool chroma_dm_at(int x0, int y0) const;
    uint8_t chroma_mode_at(int x0, int y0) const;
    void do_chroma(int lx, int ly, int luma_log2, bool cbf_cb, bool cbf_cr, bool cross_comp);
    void act_unit(int x0, int y0, int log2, uint8_t luma_mode, bool cbf_luma, bool cbf_cb, bool cbf_cr);
    void transform_unit(int x0, int y0, int xbase, int ybase, int log2, int depth, int blk, bool cbf_cb, bool cbf_cr);
    void transform_tree(int x0, int y0, int xbase, int ybase, int log2, int depth, int blk, bool intra_split,
                        int max_depth, bool pcb, bool pcr);
    void pcm_unit(int x0, int y0, int log2);
    void coding_unit(int x0, int y0, int log2);
    void coding_quadtree(int x0, int y0, int log2, int depth);
};

// Random-access coding structure with hierarchical B pictures (GOPs of
// `gop` pictures after the intra picture): display-order numbers (POCs) in
// decode order.  Each GOP's last picture comes first, then its interval is
// split at the midpoint, recursively; a shorter last GOP is split the same way.
// `layer` (by POC) receives the temporal layer: 0 for the intra picture and
// the GOP ends, k for the midpoints of the k-th split.
std::vector<int> gop_decode_order(int frames, int gop, std::vector<int>* layer = nullptr);
// Reference lists of the picture at decode index `index` (batch indices):
// a GOP's last picture refers to the earlier GOP ends, the others to the
// pictures of their own GOP in lower temporal layers decoded before them;
// at most num_refs,
// nearest in display order.  List 0 holds the earlier pictures first (nearest
// first), list 1 the later ones; without a later picture list 1 repeats
// list 0.
void gop_ref_lists(int frames, int gop, int num_refs, uint32_t index, std::vector<uint32_t>& l0,
                   std::vector<uint32_t>& l1);

}  // namespace hpvcd
