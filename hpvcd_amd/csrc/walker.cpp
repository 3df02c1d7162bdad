// walker.cpp — the coding-tree walk of walker.h.
#include "walker.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace hpvcd {

PicWalker::PicWalker(const PicParams& p, Source& s, PicRecords& o, int32_t poc)
    : P(p), src(s), out(o), poc_(poc) {
    w = P.w;
    h = P.h;
    gw = (w + 3) / 4;
    gh = (h + 3) / 4;
    sub_w = (P.chroma_idc == 1 || P.chroma_idc == 2) ? 2 : 1;
    sub_h = P.chroma_idc == 1 ? 2 : 1;
    ctb_cols = (w + (1 << P.log2_ctb) - 1) >> P.log2_ctb;
    ctb_rows = (h + (1 << P.log2_ctb) - 1) >> P.log2_ctb;
    qp_bd_off = 6 * (P.bd - 8);
    qp_y = slice_qp = qp_prev = 0;
    log2_qg = std::max(P.log2_min_cb, P.log2_ctb - (P.cu_qp_delta ? P.diff_cu_qp_delta_depth : 0));
    cell.assign((size_t)gw * gh, HPVCD_CELL_INTRA);
    cell_qp.assign((size_t)gw * gh, 0);
    mode_y.assign((size_t)gw * gh, kModeDC);
    ct_depth.assign((size_t)gw * gh, 0);
    ctb_slice.assign((size_t)ctb_cols * ctb_rows, 0);
    ctb_tile.assign((size_t)ctb_cols * ctb_rows, 0);
    out.sao.assign((size_t)ctb_cols * ctb_rows, hpvcd_sao());
    coeff_base = (uint32_t)out.coeffs.size();
    chroma_off[0] = P.chroma_qp_off[0];
    chroma_off[1] = P.chroma_qp_off[1];
    lv.assign(32 * 32, 0);
    nzpos.resize(32 * 32);
    for (int i = 0; i < 4; i++) {
        cu_chroma_modes[i] = kModeDC;
        cu_chroma_dm[i] = false;
    }
    mp_.width = w;
    mp_.height = h;
    mp_.cur_poc = poc;
    mp_.log2_ctb = P.log2_ctb;
}

void PicWalker::set_inter(const InterSlice* s) {
    inter_ = s;
    if (!s) return;
    if (!have_motion_) {
        const size_t ncell = (size_t)gw * gh;
        motion.assign(ncell, MotionCell());
        decoded.assign(ncell, 0);
        tu_v.assign(ncell, 0);
        tu_h.assign(ncell, 0);
        pu_v.assign(ncell, 0);
        pu_h.assign(ncell, 0);
        nz.assign(ncell, 0);
        have_motion_ = true;
    }
    // the slice's rows: shared with the previous inter slice when equal
    bool same = last_rows_ && last_rows_->rows.size() == s->rows.size();
    if (same)
        same = s->rows.empty() ||
               std::memcmp(last_rows_->rows.data(), s->rows.data(), s->rows.size() * sizeof(hpvcd_ref)) == 0;
    if (!same) {
        ref_base_ = (int)out.refs.size();
        out.refs.insert(out.refs.end(), s->rows.begin(), s->rows.end());
        last_rows_ = s;
    }
    mp_.is_b = s->slice_type == 0;
    mp_.par_mrg_level = s->par_mrg_level;
    mp_.max_merge = s->max_merge;
    mp_.tmvp = s->tmvp;
    mp_.col_from_l0 = s->col_from_l0;
    mp_.col = s->col;
    for (int l = 0; l < 2; l++) {
        mp_.list[l] = s->list[l].data();
        mp_.num[l] = (int)s->list[l].size();
    }
}

// 6.4.1 availability for motion: decoded, same slice and tile
// (src/decode.rs:6149)
bool PicWalker::available(int x, int y) const {
    if (x < 0 || y < 0 || x >= w || y >= h) return false;
    return decoded[g_of(x, y)] && same_region(x, y);
}

void PicWalker::set_layout(const std::vector<uint8_t>& cs, const std::vector<uint8_t>& ct) {
    ctb_slice = cs;
    ctb_tile = ct;
}

bool PicWalker::same_region(int x, int y) const {
    const size_t c = (size_t)(y >> P.log2_ctb) * ctb_cols + (x >> P.log2_ctb);
    return ctb_slice[c] == cur_slice && ctb_tile[c] == cur_tile;
}

int PicWalker::ctb(int rx, int ry, bool first_in_slice, bool first_in_tile, bool last_in_slice, bool last_in_tile,
                   int sqp) {
    const size_t rs = (size_t)ry * ctb_cols + rx;
    cur_slice = ctb_slice[rs];
    cur_tile = ctb_tile[rs];
    if (first_in_slice) {
        slice_qp = std::clamp(sqp, -qp_bd_off, 51);
        qp_y = slice_qp;
        // initType (9.3.2.2): 0 for I, P: 1 (2 with cabac_init_flag), B: 2 (1)
        int init_type = 0;
        if (inter_) init_type = inter_->slice_type == 1 ? (inter_->cabac_init ? 2 : 1) : (inter_->cabac_init ? 1 : 2);
        src.begin_slice(slice_qp, init_type);
    } else if (first_in_tile) {
        // WPP: a CTB row that starts inside its tile takes the contexts the
        // row above stored, when that row's upper-right CTB is in the same
        // tile and slice (src/decode.rs:1230-1262)
        int sync = -1;
        const bool row_start = rx == 0 || ctb_tile[rs - 1] != cur_tile;
        if (P.wpp && row_start && ry > 0 && ctb_tile[rs - ctb_cols] == cur_tile && rx + 1 < ctb_cols &&
            ctb_tile[rs - ctb_cols + 1] == cur_tile && ctb_slice[rs - ctb_cols + 1] == cur_slice)
            sync = ry - 1;
        src.begin_substream(sync);
    }
    // src/decode.rs:776 slice start, :1262 tile start: qPY_PREV = SliceQpY
    if (first_in_slice || first_in_tile) qp_prev = slice_qp;
    if (P.sao && (P.sao_luma || P.sao_chroma)) {
        // parse_sao (src/decode.rs:1507): merge candidates in the same slice and tile
        const bool l = rx > 0 && ctb_slice[rs - 1] == cur_slice && ctb_tile[rs - 1] == cur_tile;
        const bool u = ry > 0 && ctb_slice[rs - ctb_cols] == cur_slice && ctb_tile[rs - ctb_cols] == cur_tile;
        src.sao(out.sao[rs], l, u, l ? &out.sao[rs - 1] : nullptr, u ? &out.sao[rs - ctb_cols] : nullptr);
    }
    coding_quadtree(rx << P.log2_ctb, ry << P.log2_ctb, P.log2_ctb, 0);
    // WPP storage after the second CTB of the row in its tile (src/decode.rs:1269)
    if (P.wpp && rx >= 1 && ctb_tile[rs - 1] == cur_tile && (rx == 1 || ctb_tile[rs - 2] != cur_tile))
        src.wpp_save(ry);
    return src.end_ctb(last_in_slice, last_in_tile);
}

void PicWalker::finish() {
    // collocated motion: the cell at each 16x16 block's corner
    out.col.poc = poc_;
    out.col.cw = out.col.ch = 0;
    out.col.cells.clear();
    if (have_motion_) {
        finalize_bs();
        out.col.cw = (w + 15) >> 4;
        out.col.ch = (h + 15) >> 4;
        out.col.cells.resize((size_t)out.col.cw * out.col.ch);
        for (int by = 0; by < out.col.ch; by++)
            for (int bx = 0; bx < out.col.cw; bx++)
                out.col.cells[(size_t)by * out.col.cw + bx] = motion[g_of(bx << 4, by << 4)];
    }
    out.cells.resize(cell.size());
    for (size_t i = 0; i < cell.size(); i++) {
        const int gx = (int)(i % (size_t)gw), gy = (int)(i / (size_t)gw);
        const size_t c = (size_t)((gy * 4) >> P.log2_ctb) * ctb_cols + ((gx * 4) >> P.log2_ctb);
        out.cells[i] = (cell[i] & 0xffu) | ((uint32_t)(uint8_t)cell_qp[i] << 8) | ((uint32_t)ctb_slice[c] << 16) |
                       ((uint32_t)ctb_tile[c] << 24);
    }
    // records list TBs in CTB raster order (decode order inside each CTB)
    bool tiled = false;
    for (uint8_t t : ctb_tile) tiled |= t != 0;
    if (tiled) {
        auto ctb_of = [&](const hpvcd_tb& t) {
            const int lx = t.comp ? t.x * sub_w : t.x, ly = t.comp ? t.y * sub_h : t.y;
            return (ly >> P.log2_ctb) * ctb_cols + (lx >> P.log2_ctb);
        };
        std::stable_sort(out.tbs.begin(), out.tbs.end(),
                         [&](const hpvcd_tb& a, const hpvcd_tb& b) { return ctb_of(a) < ctb_of(b); });
    }
}

// ---------------------------------------------------------------------------
// GOP structure of the generator's random-access sequences
// ---------------------------------------------------------------------------
std::vector<int> gop_decode_order(int frames, int gop, std::vector<int>* layer) {
    std::vector<int> order;
    if (layer) layer->assign((size_t)std::max(frames, 0), 0);
    if (frames <= 0) return order;
    order.push_back(0);
    // the open interval (a, b) after its ends: midpoint, then each half
    struct Split {
        static void run(int a, int b, int depth, std::vector<int>& o, std::vector<int>* layer) {
            const int m = (a + b) / 2;
            if (m <= a || m >= b) return;
            o.push_back(m);
            if (layer) (*layer)[(size_t)m] = depth;
            run(a, m, depth + 1, o, layer);
            run(m, b, depth + 1, o, layer);
        }
    };
    for (int a = 0; a < frames - 1; a += gop) {
        const int b = std::min(a + gop, frames - 1);
        order.push_back(b);
        Split::run(a, b, 1, order, layer);
    }
    return order;
}

void gop_ref_lists(int frames, int gop, int num_refs, uint32_t index, std::vector<uint32_t>& l0,
                   std::vector<uint32_t>& l1) {
    l0.clear();
    l1.clear();
    std::vector<int> layer;
    const std::vector<int> order = gop_decode_order(frames, gop, &layer);
    if (index == 0 || index >= order.size()) return;
    std::vector<uint32_t> idx_of(order.size());
    for (size_t i = 0; i < order.size(); i++) idx_of[(size_t)order[i]] = (uint32_t)i;
    const int p = order[index];
    const int a = (p - 1) / gop * gop, b = std::min(a + gop, frames - 1);
    // candidates: for a GOP's last picture the earlier GOP ends (and the intra
    // picture), for the others the pictures of its own GOP in lower temporal
    // layers (the GOP's ends and the midpoints of coarser splits) decoded
    // before it
    std::vector<int> cand;
    if (p == b) {
        for (int q = a; q >= 0; q -= gop) cand.push_back(q);
    } else {
        for (int q = a; q <= b; q++)
            if (idx_of[(size_t)q] < index &&
                (q == a || q == b || (layer[(size_t)q] > 0 && layer[(size_t)q] < layer[(size_t)p])))
                cand.push_back(q);
    }
    // the nearest in display order, the earlier one first on a tie
    std::stable_sort(cand.begin(), cand.end(), [&](int x, int y) {
        const int dx = std::abs(x - p), dy = std::abs(y - p);
        return dx != dy ? dx < dy : x < y;
    });
    if ((int)cand.size() > num_refs) cand.resize((size_t)num_refs);
    std::vector<int> past, future;
    for (int q : cand) (q < p ? past : future).push_back(q);
    std::sort(past.rbegin(), past.rend());
    std::sort(future.begin(), future.end());
    for (int q : past) l0.push_back(idx_of[(size_t)q]);
    for (int q : future) l0.push_back(idx_of[(size_t)q]);
    if (future.empty()) {
        l1 = l0;  // no later picture: list 1 repeats list 0 (generalised P/B)
    } else {
        for (int q : future) l1.push_back(idx_of[(size_t)q]);
        for (int q : past) l1.push_back(idx_of[(size_t)q]);
    }
}

// ---------------------------------------------------------------------------
// inter CUs
// ---------------------------------------------------------------------------
// transform-block boundary of an inter CU or TU (edge flags only; the
// strength is derived once the picture is complete)
void PicWalker::mark_tu_edges(int x0, int y0, int size) {
    if (x0 > 0 && x0 < w)
        for (int yy = y0; yy < std::min(y0 + size, h); yy += 4) tu_v[g_of(x0, yy)] = 1;
    if (y0 > 0 && y0 < h)
        for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) tu_h[g_of(xx, y0)] = 1;
}

// src/decode.rs:2955 finalize_coeff_bs over the whole picture: an intra
// side gives 2, coefficients at a transform edge or differing motion 1
// (the per-PU marking of src/decode.rs:6329 gives the same strengths on
// the 8x8 grid)
void PicWalker::finalize_bs() {
    for (int gy = 0; gy < gh; gy++)
        for (int gx = 0; gx < gw; gx++) {
            size_t g = (size_t)gy * gw + gx;
            uint32_t& c = cell[g];
            if (gx > 0 && (tu_v[g] || pu_v[g])) {
                uint32_t bs = c & 3u;
                if (bs < 2 && (motion[g].intra || motion[g - 1].intra)) bs = 2;
                else if (bs < 1 && tu_v[g] && (nz[g] || nz[g - 1])) bs = 1;
                else if (bs < 1 && motion_differs(motion[g], motion[g - 1])) bs = 1;
                c = (c & ~3u) | bs;
            }
            if (gy > 0 && (tu_h[g] || pu_h[g])) {
                uint32_t bs = (c >> 2) & 3u;
                if (bs < 2 && (motion[g].intra || motion[g - gw].intra)) bs = 2;
                else if (bs < 1 && tu_h[g] && (nz[g] || nz[g - gw])) bs = 1;
                else if (bs < 1 && motion_differs(motion[g], motion[g - gw])) bs = 1;
                c = (c & ~0xcu) | (bs << 2);
            }
        }
}

// cu_skip_flag context (src/decode.rs:6290): skipped left / above CUs that
// are decoded and in the same slice and tile
int PicWalker::skip_ctx(int x0, int y0) const {
    int inc = 0;
    if (available(x0 - 1, y0) && motion[g_of(x0 - 1, y0)].skip) inc++;
    if (available(x0, y0 - 1) && motion[g_of(x0, y0 - 1)].skip) inc++;
    return inc;
}

// a candidate's motion with the pictures of its references
// (src/decode.rs:6699 cand_to_motion)
MotionCell PicWalker::cell_of(const MergeCand& c) const {
    MotionCell m;
    m.intra = 0;
    m.pred = c.pred;
    for (int l = 0; l < 2; l++) {
        m.ref_idx[l] = c.ref_idx[l];
        m.mv[l][0] = c.mv[l][0];
        m.mv[l][1] = c.mv[l][1];
        if (((c.pred >> l) & 1) && c.ref_idx[l] >= 0 && c.ref_idx[l] < mp_.num[l]) {
            const RefPic& r = mp_.list[l][c.ref_idx[l]];
            m.poc[l] = r.poc;
            m.pic[l] = r.pic;
            if (r.lt) m.lt |= (uint8_t)(1 << l);
        }
    }
    return m;
}

// the merge motion of a PU (src/decode.rs:6626 derive_pu_merge)
MotionCell PicWalker::merge_motion(const PuGeom& g, int merge_idx) const {
    return cell_of(derive_merge(*this, mp_, g, merge_idx));
}

// prediction_unit of a CU that is not skipped (src/decode.rs:6738):
// merge_flag, then merge_idx or inter_pred_idc, ref_idx, mvd and mvp flag
// per list, the vector the predictor plus the difference (16-bit wrap,
// src/decode.rs:7521)
MotionCell PicWalker::prediction_unit(const PuGeom& g, int log2_cb, bool& merge) {
    merge = src.merge_flag();
    if (merge) return merge_motion(g, mp_.max_merge > 1 ? src.merge_idx(mp_.max_merge) : 0);
    int idc = 0;
    if (mp_.is_b) idc = src.inter_pred_idc(std::min(P.log2_ctb - log2_cb, 4), g.w + g.h == 12);
    MotionCell m;
    m.intra = 0;
    for (int l = 0; l < 2; l++) {
        if ((l == 0 && idc == 1) || (l == 1 && idc == 0)) continue;
        const int n = mp_.num[l];
        const int ridx = n > 1 ? src.ref_idx(n) : 0;
        const int32_t ref_poc = ridx < n ? mp_.list[l][ridx].poc : mp_.cur_poc;
        int16_t cand[2][2], mvd[2] = {0, 0};
        derive_amvp(*this, mp_, g, l, ref_poc, cand);
        int flag = 0;
        src.mv_syntax(g, l, cand, l == 1 && idc == 2 && inter_->mvd_l1_zero, mvd, flag);
        m.pred |= (uint8_t)(1 << l);
        m.ref_idx[l] = (int8_t)ridx;
        m.mv[l][0] = (int16_t)(uint16_t)((uint32_t)(uint16_t)cand[flag][0] + (uint32_t)(uint16_t)mvd[0]);
        m.mv[l][1] = (int16_t)(uint16_t)((uint32_t)(uint16_t)cand[flag][1] + (uint32_t)(uint16_t)mvd[1]);
        m.poc[l] = ref_poc;
        if (ridx < n) {
            m.pic[l] = mp_.list[l][ridx].pic;
            if (mp_.list[l][ridx].lt) m.lt |= (uint8_t)(1 << l);
        }
    }
    return m;
}

// the PU's motion over its cells and its record
void PicWalker::store_pu(const PuGeom& g, const MotionCell& m) {
    for (int yy = g.y; yy < std::min(g.y + g.h, h); yy += 4)
        for (int xx = g.x; xx < std::min(g.x + g.w, w); xx += 4) motion[g_of(xx, yy)] = m;
    hpvcd_pu pu;
    std::memset(&pu, 0, sizeof(pu));
    pu.x = (uint16_t)g.x;
    pu.y = (uint16_t)g.y;
    pu.w = (uint8_t)g.w;
    pu.h = (uint8_t)g.h;
    pu.flags = m.pred;
    for (int l = 0; l < 2; l++) {
        if (!((m.pred >> l) & 1)) continue;
        const int idx = m.ref_idx[l];
        if (idx < 0 || idx >= mp_.num[l]) {
            fail_ = true;  // a candidate beyond the list (not produced by conforming streams)
            continue;
        }
        pu.ref[l] = (uint16_t)(ref_base_ + (l == 0 ? idx : mp_.num[0] + idx));
        pu.mv[l][0] = m.mv[l][0];
        pu.mv[l][1] = m.mv[l][1];
    }
    out.pus.push_back(pu);
}

// TBs without residual tiling an inter CU that codes none (skip, or
// rqt_root_cbf = 0): the MC prediction is the reconstruction.
void PicWalker::inter_tiles(int x0, int y0, int log2) {
    if (log2 > P.log2_max_tb) {
        int half = 1 << (log2 - 1);
        for (int k = 0; k < 4; k++) inter_tiles(x0 + (k & 1) * half, y0 + (k >> 1) * half, log2 - 1);
        return;
    }
    out.tbs.push_back(make_tb(x0, y0, log2, 0, kModeDC, false));
    if (!P.chroma_idc) return;
    int cl = P.chroma_idc == 3 ? log2 : log2 - 1;
    for (int comp = 1; comp <= 2; comp++) out.tbs.push_back(make_tb(x0 / sub_w, y0 / sub_h, cl, comp, kModeDC, false));
}

// src/decode.rs:6380 decode_inter_cu: part_mode, PUs in coding order (each
// PU's motion visible to the next), then rqt_root_cbf (inferred for a
// 2Nx2N merge) and the residual tree, split at depth 0 by interSplitFlag
void PicWalker::inter_cu(int x0, int y0, int log2, bool skip) {
    const int cb = 1 << log2;
    cu_inter = true;
    for (int yy = y0; yy < std::min(y0 + cb, h); yy += 4)
        for (int xx = x0; xx < std::min(x0 + cb, w); xx += 4) cell[g_of(xx, yy)] &= ~HPVCD_CELL_INTRA;
    mark_tu_edges(x0, y0, cb);
    int part = PART_2Nx2N;
    if (!skip) part = src.part_mode_inter(log2, log2 == P.log2_min_cb, P.amp && log2 > P.log2_min_cb);
    const int q = cb / 4, hf = cb / 2;
    int rects[4][4], nr = 0;
    auto add = [&](int x, int y, int rw, int rh) {
        rects[nr][0] = x; rects[nr][1] = y; rects[nr][2] = rw; rects[nr][3] = rh; nr++;
    };
    switch (part) {
        case PART_2Nx2N: add(x0, y0, cb, cb); break;
        case PART_2NxN: add(x0, y0, cb, hf); add(x0, y0 + hf, cb, hf); break;
        case PART_Nx2N: add(x0, y0, hf, cb); add(x0 + hf, y0, hf, cb); break;
        case PART_NxN: add(x0, y0, hf, hf); add(x0 + hf, y0, hf, hf); add(x0, y0 + hf, hf, hf); add(x0 + hf, y0 + hf, hf, hf); break;
        case PART_2NxnU: add(x0, y0, cb, q); add(x0, y0 + q, cb, cb - q); break;
        case PART_2NxnD: add(x0, y0, cb, cb - q); add(x0, y0 + cb - q, cb, q); break;
        case PART_nLx2N: add(x0, y0, q, cb); add(x0 + q, y0, cb - q, cb); break;
        default: add(x0, y0, cb - q, cb); add(x0 + cb - q, y0, q, cb); break;
    }
    bool last_merge = false;
    for (int i = 0; i < nr; i++) {
        PuGeom g;
        g.x = rects[i][0]; g.y = rects[i][1]; g.w = rects[i][2]; g.h = rects[i][3];
        g.cu_x = x0; g.cu_y = y0; g.cu_w = cb; g.cu_h = cb; g.part_idx = i;
        if (g.x > x0)
            for (int yy = g.y; yy < std::min(g.y + g.h, h); yy += 4) pu_v[g_of(g.x, yy)] = 1;
        if (g.y > y0)
            for (int xx = g.x; xx < std::min(g.x + g.w, w); xx += 4) pu_h[g_of(xx, g.y)] = 1;
        MotionCell m;
        if (skip) {
            m = merge_motion(g, mp_.max_merge > 1 ? src.merge_idx(mp_.max_merge) : 0);
            last_merge = true;
        } else {
            m = prediction_unit(g, log2, last_merge);
        }
        m.skip = skip ? 1 : 0;
        store_pu(g, m);
    }
    bool root_cbf = false;
    if (!skip) root_cbf = (part == PART_2Nx2N && last_merge) ? true : src.rqt_root_cbf();
    if (root_cbf) {
        const bool inter_split = P.max_depth_inter == 0 && part != PART_2Nx2N;
        transform_tree(x0, y0, x0, y0, log2, 0, 0, inter_split, P.max_depth_inter + (inter_split ? 1 : 0), false, false);
    } else {
        inter_tiles(x0, y0, log2);
    }
    for (int yy = y0; yy < std::min(y0 + cb, h); yy += 4)
        for (int xx = x0; xx < std::min(x0 + cb, w); xx += 4) cell_qp[g_of(xx, yy)] = (int8_t)qp_y;
    qp_prev = qp_y;
    cu_inter = false;
}

// ---------------------------------------------------------------------------
// side maps
// ---------------------------------------------------------------------------
// src/decode.rs:3018 mark_block_edges (bS kept as the max of the marks)
void PicWalker::mark_block_edges(int x0, int y0, int size, int bs) {
    if (x0 > 0 && x0 < w) {
        for (int yy = y0; yy < std::min(y0 + size, h); yy += 4) {
            uint32_t& c = cell[g_of(x0, yy)];
            if ((uint32_t)bs > (c & 3u)) c = (c & ~3u) | (uint32_t)bs;
        }
    }
    if (y0 > 0 && y0 < h) {
        for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) {
            uint32_t& c = cell[g_of(xx, y0)];
            if ((uint32_t)bs > ((c >> 2) & 3u)) c = (c & ~0xcu) | ((uint32_t)bs << 2);
        }
    }
}

void PicWalker::fill_cells(int x0, int y0, int size, uint32_t bit) {
    for (int yy = y0; yy < std::min(y0 + size, h); yy += 4)
        for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) cell[g_of(xx, yy)] |= bit;
}

void PicWalker::set_mode(int x0, int y0, int size, uint8_t mode) {
    for (int yy = y0; yy < std::min(y0 + size, h); yy += 4)
        for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) mode_y[g_of(xx, yy)] = mode;
}

// candIntraPredModeX (src/decode.rs:3226 neighbor_mode): out of the
// picture, another tile or slice, or not intra -> DC (PCM CUs hold DC)
uint8_t PicWalker::neighbour_mode(int x, int y) const {
    if (x < 0 || y < 0 || x >= w || y >= h) return kModeDC;
    if (!same_region(x, y)) return kModeDC;
    if (!(cell[g_of(x, y)] & HPVCD_CELL_INTRA)) return kModeDC;
    return mode_y[g_of(x, y)];
}

// MPM candidate list (src/decode.rs:3282 mpm_list)
void PicWalker::mpm_list(int x0, int y0, uint8_t* mpm) const {
    const uint8_t a = neighbour_mode(x0 - 1, y0);
    uint8_t b = neighbour_mode(x0, y0 - 1);
    if (y0 > 0 && ((y0 - 1) >> P.log2_ctb) != (y0 >> P.log2_ctb)) b = kModeDC;
    if (a == b) {
        if (a < 2) {
            mpm[0] = kModePlanar; mpm[1] = kModeDC; mpm[2] = 26;
        } else {
            mpm[0] = a;
            mpm[1] = (uint8_t)(2 + ((a + 29) % 32));
            mpm[2] = (uint8_t)(2 + ((a - 2 + 1) % 32));
        }
    } else {
        mpm[0] = a;
        mpm[1] = b;
        mpm[2] = (a != kModePlanar && b != kModePlanar) ? kModePlanar : ((a != kModeDC && b != kModeDC) ? kModeDC : 26);
    }
}

// split_cu_flag context (src/decode.rs:2540 split_cu_ctx)
int PicWalker::split_ctx(int x0, int y0, int depth) const {
    int inc = 0;
    if (x0 > 0 && same_region(x0 - 1, y0) && ct_depth[g_of(x0 - 1, y0)] > depth) inc++;
    if (y0 > 0 && same_region(x0, y0 - 1) && ct_depth[g_of(x0, y0 - 1)] > depth) inc++;
    return inc;
}

// ---------------------------------------------------------------------------
// QP
// ---------------------------------------------------------------------------
// qPY_PRED of the quantisation group at (xqg, yqg): the left and above
// QpY inside the same CTB, else the QpY of the previous CU in decode
// order (src/decode.rs:3084 predict_qp; no WPP)
int PicWalker::predict_qp(int xqg, int yqg) {
    const int ctb = 1 << P.log2_ctb;
    const int cx = (xqg / ctb) * ctb, cy = (yqg / ctb) * ctb;
    const int a = (xqg >= 1 && xqg - 1 >= cx) ? cell_qp[g_of(xqg - 1, yqg)] : qp_prev;
    const int b = (yqg >= 1 && yqg - 1 >= cy) ? cell_qp[g_of(xqg, yqg - 1)] : qp_prev;
    return std::clamp((a + b + 1) >> 1, -qp_bd_off, 51);
}

// src/decode.rs:5768 derive_qpy_from_delta
int PicWalker::qpy_from_delta(int pred, int delta) const {
    const int m = 52 + qp_bd_off;
    int q = (pred + delta + 52 + 2 * qp_bd_off) % m;
    if (q < 0) q += m;
    return std::clamp(q - qp_bd_off, -qp_bd_off, 51);
}

// cu_qp_delta of the first TU with a coded residual in a quantisation
// group (src/decode.rs:3528)
void PicWalker::maybe_qp_delta(bool residual) {
    if (!P.cu_qp_delta || !residual || qp_delta_coded) return;
    qp_delta_coded = true;
    qp_y = qpy_from_delta(qp_y, src.qp_delta());
}

int PicWalker::qp_prime(int qp, int bitd) const {
    int off = 6 * (bitd - 8);
    return std::clamp(qp + off, 0, 51 + off);
}

int PicWalker::qpc_table(int qpi, int chroma_idc, int bitd) {
    static const int T[14] = {29, 30, 31, 32, 33, 33, 34, 34, 35, 35, 36, 36, 37, 37};
    qpi = std::clamp(qpi, -6 * (bitd - 8), 57);
    if (chroma_idc != 1) return std::min(qpi, 51);
    if (qpi < 30) return qpi;
    if (qpi > 43) return qpi - 6;
    return T[qpi - 30];
}

// Qp' of a TB.  Under ACT the component offsets apply to QpY before the
// chroma mapping and no PPS chroma offsets (src/decode.rs:3788);
// otherwise chroma adds the PPS and slice offsets (src/decode.rs:5045).
int PicWalker::tb_qp(int comp, bool act) {
    if (!act) {
        if (comp == 0) return qp_prime(qp_y, P.bd);
        return qp_prime(qpc_table(qp_y + chroma_off[comp - 1], P.chroma_idc, P.bd_c), P.bd_c);
    }
    const int qpi = qp_y + P.act_qp[comp];
    if (comp == 0) return qp_prime(std::clamp(qpi, -qp_bd_off, 51), P.bd);
    return qp_prime(qpc_table(qpi, P.chroma_idc, P.bd_c), P.bd_c);
}

// ---------------------------------------------------------------------------
// transform blocks
// ---------------------------------------------------------------------------
hpvcd_tb PicWalker::make_tb(int x, int y, int log2n, int comp, uint8_t mode, bool cbf, bool act) {
    hpvcd_tb tb;
    std::memset(&tb, 0, sizeof(tb));
    tb.x = (uint16_t)x;
    tb.y = (uint16_t)y;
    tb.log2n = (uint8_t)log2n;
    tb.comp = (uint8_t)comp;
    tb.mode = mode;
    tb.qp = (uint8_t)tb_qp(comp, act);
    if (act) tb.flags |= HPVCD_TB_ACT;
    if (cu_tqb) tb.flags |= HPVCD_TB_TQB;
    if (cu_inter) tb.flags |= HPVCD_TB_INTER;
    if (!cbf) return tb;
    tb.flags |= HPVCD_TB_CBF;
    ResidualParams rp;
    rp.log2n = log2n;
    rp.luma = comp == 0;
    rp.scan = cu_inter ? SCAN_DIAG : scan_for(mode, log2n, comp == 0, P.chroma_idc == 3);
    rp.ts_allowed = P.transform_skip && !cu_tqb && log2n <= P.log2_max_ts;
    rp.tqb = cu_tqb;
    rp.explicit_rdpcm = P.explicit_rdpcm && cu_inter;
    rp.implicit_rdpcm = P.implicit_rdpcm && !cu_inter && (mode == 10 || mode == 26);
    rp.sign_hiding = P.sign_hiding;
    rp.persistent_rice = P.persistent_rice;
    rp.ts_context = P.ts_context;
    rp.bypass_alignment = P.bypass_alignment;
    rp.extended_precision = P.extended_precision;
    rp.bit_depth = comp ? P.bd_c : P.bd;
    ResidualOut ro;
    ro.nz = nzpos.data();  // lv is all zero between TBs
    const int n = 1 << log2n;
    src.residual(rp, lv.data(), ro);
    if (ro.ts) tb.flags |= HPVCD_TB_TS;
    if (ro.rdpcm == 0) tb.flags |= HPVCD_TB_RDPCM_H;
    if (ro.rdpcm == 1) tb.flags |= HPVCD_TB_RDPCM_V;
    const uint32_t start = (uint32_t)out.coeffs.size();
    int max_x = 0;
    if (ro.nnz >= 0) {
        // the source listed its nonzero positions: out in raster order (a
        // bit per position of the TB), then cleared
        uint64_t bits[16] = {};  // n * n <= 1024
        for (int i = 0; i < ro.nnz; i++) bits[nzpos[i] >> 6] |= 1ull << (nzpos[i] & 63);
        for (int wi = 0; wi < (n * n + 63) >> 6; wi++)
            for (uint64_t b = bits[wi]; b; b &= b - 1) {
                const int pos = wi * 64 + __builtin_ctzll(b);
                out.coeffs.push_back(HPVCD_COEFF_PACK(pos, lv[pos]));
                max_x = std::max(max_x, pos & (n - 1));
                lv[pos] = 0;
            }
    } else {
        for (int pos = 0; pos < n * n; pos++) {
            if (!lv[pos]) continue;
            out.coeffs.push_back(HPVCD_COEFF_PACK(pos, lv[pos]));
            max_x = std::max(max_x, pos & (n - 1));
            lv[pos] = 0;
        }
    }
    tb.coeff_off = start - coeff_base;
    tb.coeff_cnt = (uint32_t)out.coeffs.size() - start;
    tb.nx = (uint8_t)(max_x + 1);
    return tb;
}

bool PicWalker::chroma_dm_at(int x0, int y0) const {
    int col = (x0 - cu_x0) >= cu_pu, row = (y0 - cu_y0) >= cu_pu;
    return cu_chroma_dm[row * 2 + col];
}

uint8_t PicWalker::chroma_mode_at(int x0, int y0) const {
    int col = (x0 - cu_x0) >= cu_pu, row = (y0 - cu_y0) >= cu_pu;
    return cu_chroma_modes[row * 2 + col];
}

// src/decode.rs:5007 do_chroma: cross_comp_pred(Cb), Cb TB, cross_comp_pred(Cr), Cr TB
void PicWalker::do_chroma(int lx, int ly, int luma_log2, bool cbf_cb, bool cbf_cr, bool cross_comp) {
    uint8_t mode = cu_inter ? kModeDC : chroma_mode_at(lx, ly);
    int clog2 = P.chroma_idc == 3 ? luma_log2 : luma_log2 - 1;
    int cx0 = lx / sub_w, cy0 = ly / sub_h;
    for (int comp = 1; comp <= 2; comp++) {
        int scale = cross_comp ? src.res_scale(comp - 1) : 0;
        bool cbf = comp == 1 ? cbf_cb : cbf_cr;
        hpvcd_tb tb = make_tb(cx0, cy0, clog2, comp, mode, cbf);
        tb.res_scale = (int8_t)scale;
        out.tbs.push_back(tb);
    }
}

// A co-located luma/Cb/Cr TB triple whose residuals go through the
// inverse colour transform (src/decode.rs:3619 transform_unit_act).
void PicWalker::act_unit(int x0, int y0, int log2, uint8_t luma_mode, bool cbf_luma, bool cbf_cb, bool cbf_cr) {
    bool cross_comp = P.cross_component && cbf_luma && (cu_inter || chroma_dm_at(x0, y0));
    out.tbs.push_back(make_tb(x0, y0, log2, 0, luma_mode, cbf_luma, true));
    uint8_t mode = cu_inter ? kModeDC : chroma_mode_at(x0, y0);
    for (int comp = 1; comp <= 2; comp++) {
        int scale = cross_comp ? src.res_scale(comp - 1) : 0;
        hpvcd_tb tb = make_tb(x0, y0, log2, comp, mode, comp == 1 ? cbf_cb : cbf_cr, true);
        tb.res_scale = (int8_t)scale;
        out.tbs.push_back(tb);
    }
}

void PicWalker::transform_unit(int x0, int y0, int xbase, int ybase, int log2, int depth, int blk, bool cbf_cb,
                               bool cbf_cr) {
    // cbf_luma of an inter CU's root TU without chroma residual is inferred
    // (src/decode.rs:3459)
    bool cbf_luma;
    if (cu_inter && depth == 0 && !cbf_cb && !cbf_cr) cbf_luma = true;
    else cbf_luma = src.cbf_luma(depth);
    int size = 1 << log2;
    if (cu_inter) {
        mark_tu_edges(x0, y0, size);
        if (cbf_luma)
            for (int yy = y0; yy < std::min(y0 + size, h); yy += 4)
                for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) nz[g_of(xx, yy)] = 1;
    } else {
        mark_block_edges(x0, y0, size, 2);
    }
    uint8_t luma_mode = cu_inter ? kModeDC : mode_y[g_of(x0, y0)];
    const bool residual = cbf_luma || cbf_cb || cbf_cr;
    // tu_residual_act_flag is present for a TU with residual of an inter
    // CU, or of an intra CU whose chroma modes are all the derived mode
    // (src/decode.rs:4229)
    bool dm_all = cu_chroma_dm[0] && cu_chroma_dm[1] && cu_chroma_dm[2] && cu_chroma_dm[3];
    const bool act = P.act && P.chroma_idc == 3 && residual && (cu_inter || dm_all) && src.act_flag();
    maybe_qp_delta(residual);
    if (act) {
        act_unit(x0, y0, log2, luma_mode, cbf_luma, cbf_cb, cbf_cr);
        return;
    }
    bool cross_comp = P.cross_component && cbf_luma && (cu_inter || chroma_dm_at(x0, y0));
    out.tbs.push_back(make_tb(x0, y0, log2, 0, luma_mode, cbf_luma));
    if (P.chroma_idc == 0) return;
    if (log2 > 2 || P.chroma_idc == 3) do_chroma(x0, y0, log2, cbf_cb, cbf_cr, cross_comp);
    else if (blk == 3) do_chroma(xbase, ybase, 3, cbf_cb, cbf_cr, false);
}

void PicWalker::transform_tree(int x0, int y0, int xbase, int ybase, int log2, int depth, int blk, bool intra_split,
                               int max_depth, bool pcb, bool pcr) {
    bool allowed = log2 <= P.log2_max_tb && log2 > P.log2_min_tb && depth < max_depth && !(intra_split && depth == 0);
    bool split = allowed ? src.split_tf(log2) : (log2 > P.log2_max_tb || (intra_split && depth == 0));
    bool cbf_cb = pcb, cbf_cr = pcr;
    if (P.chroma_idc && (log2 > 2 || P.chroma_idc == 3)) {
        if (depth == 0 || pcb) cbf_cb = src.cbf_chroma(depth);
        if (depth == 0 || pcr) cbf_cr = src.cbf_chroma(depth);
    }
    if (split) {
        int half = 1 << (log2 - 1);
        transform_tree(x0, y0, x0, y0, log2 - 1, depth + 1, 0, intra_split, max_depth, cbf_cb, cbf_cr);
        transform_tree(x0 + half, y0, x0, y0, log2 - 1, depth + 1, 1, intra_split, max_depth, cbf_cb, cbf_cr);
        transform_tree(x0, y0 + half, x0, y0, log2 - 1, depth + 1, 2, intra_split, max_depth, cbf_cb, cbf_cr);
        transform_tree(x0 + half, y0 + half, x0, y0, log2 - 1, depth + 1, 3, intra_split, max_depth, cbf_cb, cbf_cr);
    } else {
        transform_unit(x0, y0, xbase, ybase, log2, depth, blk, cbf_cb, cbf_cr);
    }
}

// I_PCM CU (src/decode.rs:2768 decode_pcm_cu): luma then Cb then Cr,
// every sample a raw PcmBitDepth-bit value, scaled by the TB's shift.
void PicWalker::pcm_unit(int x0, int y0, int log2) {
    int size = 1 << log2;
    fill_cells(x0, y0, size, HPVCD_CELL_PCM);
    set_mode(x0, y0, size, kModeDC);
    mark_block_edges(x0, y0, size, 2);
    const int ncomp = P.chroma_idc ? 3 : 1;
    const int cl2 = P.chroma_idc == 3 ? log2 : log2 - 1;
    const int nl = size * size, nc = P.chroma_idc ? (1 << cl2) * (1 << cl2) : 0;
    std::vector<uint32_t> samples((size_t)nl + 2 * (size_t)nc);
    src.pcm_samples(nl, nc, P.pcm_bd, P.pcm_bd_c, samples.data());
    size_t k = 0;
    for (int comp = 0; comp < ncomp; comp++) {
        int l2 = comp == 0 ? log2 : cl2;
        int n = 1 << l2;
        hpvcd_tb tb;
        std::memset(&tb, 0, sizeof(tb));
        tb.x = (uint16_t)(comp ? x0 / sub_w : x0);
        tb.y = (uint16_t)(comp ? y0 / sub_h : y0);
        tb.log2n = (uint8_t)l2;
        tb.comp = (uint8_t)comp;
        tb.mode = kModeDC;
        tb.qp = (uint8_t)((comp ? P.bd_c - P.pcm_bd_c : P.bd - P.pcm_bd));
        tb.flags = HPVCD_TB_PCM | HPVCD_TB_CBF;
        tb.coeff_off = (uint32_t)out.coeffs.size() - coeff_base;
        for (int pos = 0; pos < n * n; pos++) out.coeffs.push_back(HPVCD_COEFF_PACK(pos, (int)samples[k++]));
        tb.coeff_cnt = (uint32_t)(n * n);
        tb.nx = (uint8_t)n;
        out.tbs.push_back(tb);
    }
    for (int yy = y0; yy < std::min(y0 + size, h); yy += 4)
        for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) cell_qp[g_of(xx, yy)] = (int8_t)qp_y;
    qp_prev = qp_y;
}

// src/decode.rs:2585 coding_unit: cu_transquant_bypass_flag, then in P and
// B slices cu_skip_flag and pred_mode_flag
void PicWalker::coding_unit(int x0, int y0, int log2) {
    int size = 1 << log2;
    const int qgm = ~((1 << log2_qg) - 1);
    if ((x0 & qgm) == x0 && (y0 & qgm) == y0) qp_y = predict_qp(x0, y0);
    cu_tqb = P.tqb_enabled && src.tqb();
    if (cu_tqb) fill_cells(x0, y0, size, HPVCD_CELL_TQB);
    if (inter_) {
        if (src.cu_skip(skip_ctx(x0, y0))) {
            inter_cu(x0, y0, log2, true);
            return;
        }
        if (!src.pred_mode_intra()) {
            inter_cu(x0, y0, log2, false);
            return;
        }
    }
    bool nxn = log2 == P.log2_min_cb && src.intra_nxn();
    // pcm_flag follows part_mode for 2Nx2N CUs of the PCM size range (src/decode.rs:2647)
    if (P.pcm && !nxn && log2 >= P.log2_min_pcm && log2 <= P.log2_max_pcm && src.pcm_flag()) {
        pcm_unit(x0, y0, log2);
        return;
    }
    int npu = nxn ? 2 : 1, pu = size / npu;
    uint8_t luma_modes[4] = {kModeDC, kModeDC, kModeDC, kModeDC};
    src.luma_modes(
        npu * npu, [&](int k, uint8_t* mpm) { mpm_list(x0 + (k % npu) * pu, y0 + (k / npu) * pu, mpm); },
        [&](int k, uint8_t mode) {
            luma_modes[k] = mode;
            set_mode(x0 + (k % npu) * pu, y0 + (k / npu) * pu, pu, mode);
        });
    int n_chroma = (P.chroma_idc == 3 && nxn) ? 4 : 1;
    for (int i = 0; i < 4; i++) { cu_chroma_modes[i] = kModeDC; cu_chroma_dm[i] = false; }
    if (P.chroma_idc) {
        static const uint8_t cand[4] = {0, 26, 10, 1};
        for (int i = 0; i < n_chroma; i++) {
            // Table 8-2 chroma mode derivation (src/decode.rs:3248)
            const int syn = src.chroma_mode_syntax();
            const uint8_t m = syn == 4 ? luma_modes[i] : (cand[syn] == luma_modes[i] ? (uint8_t)34 : cand[syn]);
            cu_chroma_modes[i] = m;
            cu_chroma_dm[i] = syn == 4;
        }
        if (n_chroma == 1)
            for (int i = 1; i < 4; i++) { cu_chroma_modes[i] = cu_chroma_modes[0]; cu_chroma_dm[i] = cu_chroma_dm[0]; }
    }
    cu_x0 = x0; cu_y0 = y0; cu_pu = std::max(pu, 1);
    int max_depth = P.max_depth_intra + (nxn ? 1 : 0);
    transform_tree(x0, y0, x0, y0, log2, 0, 0, nxn, max_depth, false, false);
    if (nxn) {
        int half = size / 2;
        mark_block_edges(x0 + half, y0, half, 2);
        mark_block_edges(x0, y0 + half, half, 2);
    }
    for (int yy = y0; yy < std::min(y0 + size, h); yy += 4)
        for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) cell_qp[g_of(xx, yy)] = (int8_t)qp_y;
    qp_prev = qp_y;
}

void PicWalker::coding_quadtree(int x0, int y0, int log2, int depth) {
    int size = 1 << log2;
    if (x0 >= w || y0 >= h) return;
    bool split;
    if (x0 + size > w || y0 + size > h) split = log2 > P.log2_min_cb;
    else split = log2 > P.log2_min_cb && src.split_cu(split_ctx(x0, y0, depth));
    if (P.cu_qp_delta && log2 >= log2_qg) qp_delta_coded = false;
    if (split) {
        int half = size / 2;
        coding_quadtree(x0, y0, log2 - 1, depth + 1);
        coding_quadtree(x0 + half, y0, log2 - 1, depth + 1);
        coding_quadtree(x0, y0 + half, log2 - 1, depth + 1);
        coding_quadtree(x0 + half, y0 + half, log2 - 1, depth + 1);
    } else {
        for (int yy = y0; yy < std::min(y0 + size, h); yy += 4)
            for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) ct_depth[g_of(xx, yy)] = (uint8_t)depth;
        coding_unit(x0, y0, log2);
        // the CU is decoded: its cells become available to later CUs' motion
        if (have_motion_)
            for (int yy = y0; yy < std::min(y0 + size, h); yy += 4)
                for (int xx = x0; xx < std::min(x0 + size, w); xx += 4) decoded[g_of(xx, yy)] = 1;
    }
}

}  // namespace hpvcd
