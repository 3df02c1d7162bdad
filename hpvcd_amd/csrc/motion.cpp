// motion.cpp — merge / AMVP / temporal derivations of motion.h.
#include "motion.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace hpvcd {

namespace {

enum Pos { A0, A1, B0, B1, B2 };

bool same_motion(const MergeCand& a, const MergeCand& b) {
    return a.pred == b.pred && a.ref_idx[0] == b.ref_idx[0] && a.ref_idx[1] == b.ref_idx[1] &&
           a.mv[0][0] == b.mv[0][0] && a.mv[0][1] == b.mv[0][1] && a.mv[1][0] == b.mv[1][0] &&
           a.mv[1][1] == b.mv[1][1];
}

MergeCand from_cell(const MotionCell& m) {
    MergeCand c;
    c.pred = m.pred;
    for (int l = 0; l < 2; l++) {
        c.ref_idx[l] = m.ref_idx[l];
        c.mv[l][0] = m.mv[l][0];
        c.mv[l][1] = m.mv[l][1];
    }
    return c;
}

bool in_picture(const MotionParams& mp, int x, int y) { return x >= 0 && y >= 0 && x < mp.width && y < mp.height; }

// 6.4.2 prediction block availability: a neighbour inside the current
// coding block is available except the lower-left region of the second
// NxN PU; outside it the z-scan / slice / tile process decides
// (src/motion.rs:131)
bool pb_available(const Neighbours& nb, const MotionParams& mp, const PuGeom& pu, int px, int py) {
    if (px < 0 || py < 0) return false;
    const bool same_cb = px >= pu.cu_x && py >= pu.cu_y && px < pu.cu_x + pu.cu_w && py < pu.cu_y + pu.cu_h;
    if (!same_cb) return in_picture(mp, px, py) && nb.available(px, py);
    return !(pu.w * 2 == pu.cu_w && pu.h * 2 == pu.cu_h && pu.part_idx == 1 && pu.cu_y + pu.h <= py &&
             pu.cu_x + pu.w > px);
}

// an available inter neighbour's cell, or null
const MotionCell* inter_at(const Neighbours& nb, const MotionParams& mp, const PuGeom& pu, int px, int py) {
    if (!pb_available(nb, mp, pu, px, py) || !in_picture(mp, px, py)) return nullptr;
    const MotionCell& m = nb.motion_at(px, py);
    return m.intra ? nullptr : &m;
}

// a spatial merge candidate (src/motion.rs:238): the second PU of a
// vertical (horizontal) split drops A1 (B1); a neighbour in the PU's merge
// estimation region is unavailable
bool spatial(const Neighbours& nb, const MotionParams& mp, const PuGeom& pu, int px, int py, Pos pos, MergeCand& out) {
    if (pu.part_idx == 1) {
        const bool vertical = pu.w < pu.cu_w && pu.h == pu.cu_h;
        const bool horizontal = pu.h < pu.cu_h && pu.w == pu.cu_w;
        if ((pos == A1 && vertical) || (pos == B1 && horizontal)) return false;
    }
    const int lvl = mp.par_mrg_level;
    if (px >= 0 && py >= 0 && (pu.x >> lvl) == (px >> lvl) && (pu.y >> lvl) == (py >> lvl)) return false;
    const MotionCell* m = inter_at(nb, mp, pu, px, py);
    if (!m) return false;
    out = from_cell(*m);
    return true;
}

bool entry_poc(const MotionParams& mp, int list, int idx, int32_t& poc) {
    if (idx < 0 || idx >= mp.num[list]) return false;
    poc = mp.list[list][idx].poc;
    return true;
}

// bottom-right collocated position, when it stays in the PU's CTB row
// (8.5.3.2.8); the centre otherwise or when it yields nothing
bool temporal_at_pu(const MotionParams& mp, const PuGeom& pu, int list, int32_t ref_poc, int16_t out[2]) {
    const int br_y = pu.y + pu.h;
    if ((br_y >> mp.log2_ctb) == (pu.y >> mp.log2_ctb) && temporal_mv(mp, pu.x + pu.w, br_y, list, ref_poc, out))
        return true;
    return temporal_mv(mp, pu.x + pu.w / 2, pu.y + pu.h / 2, list, ref_poc, out);
}

// src/motion.rs:282 temporal_merge: reference index 0 of each list
bool temporal_merge(const MotionParams& mp, const PuGeom& pu, MergeCand& out) {
    out = MergeCand();
    for (int l = 0; l < (mp.is_b ? 2 : 1); l++) {
        const int32_t ref_poc = mp.num[l] ? mp.list[l][0].poc : mp.cur_poc;
        int16_t mv[2];
        if (temporal_at_pu(mp, pu, l, ref_poc, mv)) {
            out.pred |= (uint8_t)(1 << l);
            out.mv[l][0] = mv[0];
            out.mv[l][1] = mv[1];
            out.ref_idx[l] = 0;
        }
    }
    return out.pred != 0;
}

// src/motion.rs:333 combined bi-predictive candidates (8.5.3.2.4)
void combined(const MotionParams& mp, MergeCand* c, int& n, int max) {
    if (n < 2 || n >= max) return;
    static const int L0[12] = {0, 1, 0, 2, 1, 2, 0, 3, 1, 3, 2, 3};
    static const int L1[12] = {1, 0, 2, 0, 2, 1, 3, 0, 3, 1, 3, 2};
    const int orig = n;
    for (int k = 0; n < max && k < orig * (orig - 1); k++) {
        const int i0 = L0[k % 12], i1 = L1[k % 12];
        if (i0 >= orig || i1 >= orig) continue;
        const MergeCand a = c[i0], b = c[i1];
        if (!(a.pred & 1) || !(b.pred & 2)) continue;
        int32_t p0 = 0, p1 = 0;
        const bool h0 = entry_poc(mp, 0, a.ref_idx[0], p0), h1 = entry_poc(mp, 1, b.ref_idx[1], p1);
        const bool same_pic = h0 == h1 && (!h0 || p0 == p1);
        if (!same_pic || a.mv[0][0] != b.mv[1][0] || a.mv[0][1] != b.mv[1][1]) {
            MergeCand m;
            m.pred = 3;
            m.ref_idx[0] = a.ref_idx[0];
            m.ref_idx[1] = b.ref_idx[1];
            m.mv[0][0] = a.mv[0][0];
            m.mv[0][1] = a.mv[0][1];
            m.mv[1][0] = b.mv[1][0];
            m.mv[1][1] = b.mv[1][1];
            c[n++] = m;
        }
    }
}

// src/motion.rs:375 zero candidates (8.5.3.2.5)
void zeros(const MotionParams& mp, MergeCand* c, int& n, int max) {
    const int num_ref = mp.is_b ? std::min(mp.num[0], mp.num[1]) : mp.num[0];
    if (num_ref == 0) return;
    for (int zero_idx = 0; n < max && zero_idx <= 32; zero_idx++) {
        const int r = zero_idx < num_ref ? zero_idx : 0;
        MergeCand m;
        m.pred = mp.is_b ? 3 : 1;
        m.ref_idx[0] = (int8_t)r;
        m.ref_idx[1] = (int8_t)(mp.is_b ? r : -1);
        c[n++] = m;
    }
}

bool target_is_lt(const MotionParams& mp, int list_first, int32_t ref_poc) {
    for (int k = 0; k < 2; k++) {
        const int l = k == 0 ? list_first : 1 - list_first;
        for (int i = 0; i < mp.num[l]; i++)
            if (mp.list[l][i].poc == ref_poc) return mp.list[l][i].lt;
    }
    return false;
}

// src/motion.rs:515 amvp_spatial: a same-POC pass over the positions,
// then (when allowed) the first neighbour with a reference of the target's
// long-term class, scaled by POC distance
bool amvp_spatial(const Neighbours& nb, const MotionParams& mp, const PuGeom& pu, const int (*pos)[2], int npos,
                  int list, int32_t ref_poc, bool allow_scaled, bool same_ref, bool target_lt, int16_t out[2]) {
    if (same_ref) {
        for (int i = 0; i < npos; i++) {
            const MotionCell* m = inter_at(nb, mp, pu, pos[i][0], pos[i][1]);
            if (!m) continue;
            for (int k = 0; k < 2; k++) {
                const int l = k == 0 ? list : 1 - list;
                if (((m->pred >> l) & 1) && m->poc[l] == ref_poc) {
                    out[0] = m->mv[l][0];
                    out[1] = m->mv[l][1];
                    return true;
                }
            }
        }
    }
    if (!allow_scaled) return false;
    for (int i = 0; i < npos; i++) {
        const MotionCell* m = inter_at(nb, mp, pu, pos[i][0], pos[i][1]);
        if (!m) continue;
        for (int k = 0; k < 2; k++) {
            const int l = k == 0 ? list : 1 - list;
            if (((m->pred >> l) & 1) && (((m->lt >> l) & 1) != 0) == target_lt) {
                if (target_lt) {
                    out[0] = m->mv[l][0];
                    out[1] = m->mv[l][1];
                } else {
                    scale_mv(m->mv[l], mp.cur_poc - m->poc[l], mp.cur_poc - ref_poc, out);
                }
                return true;
            }
        }
    }
    return false;
}

}  // namespace

void scale_mv(const int16_t in[2], int col_dist, int cur_dist, int16_t out[2]) {
    const int td = std::clamp(col_dist, -128, 127), tb = std::clamp(cur_dist, -128, 127);
    if (td == 0) {
        out[0] = in[0];
        out[1] = in[1];
        return;
    }
    const int tx = (16384 + (std::abs(td) >> 1)) / td;
    const int dsf = std::clamp((tb * tx + 32) >> 6, -4096, 4095);
    for (int c = 0; c < 2; c++) {
        const int p = dsf * in[c];
        const int v = (p > 0 ? 1 : (p < 0 ? -1 : 0)) * ((std::abs(p) + 127) >> 8);
        out[c] = (int16_t)std::clamp(v, -32768, 32767);
    }
}

bool temporal_mv(const MotionParams& mp, int x, int y, int list, int32_t ref_poc, int16_t out[2]) {
    const ColField* col = mp.col;
    if (!col || col->poc == mp.cur_poc) return false;
    if (x >= mp.width || y >= mp.height) return false;
    const MotionCell* m = col->at(x, y);
    if (!m || m->intra) return false;
    int sel;
    if (!(m->pred & 1)) {
        sel = 1;
    } else if (!(m->pred & 2)) {
        sel = 0;
    } else {
        // bi-predicted collocated block: the target list when every active
        // reference precedes the current picture, else list collocated_from_l0
        bool all_before = true;
        for (int l = 0; l < 2; l++)
            for (int i = 0; i < mp.num[l]; i++) all_before &= mp.list[l][i].poc < mp.cur_poc;
        sel = all_before ? list : (mp.col_from_l0 ? 1 : 0);
    }
    const int32_t col_ref_poc = m->poc[sel];
    const bool col_ref_lt = (m->lt >> sel) & 1;
    bool cur_ref_lt = false;
    for (int i = 0; i < mp.num[list]; i++)
        if (mp.list[list][i].poc == ref_poc) {
            cur_ref_lt = mp.list[list][i].lt;
            break;
        }
    if (cur_ref_lt != col_ref_lt) return false;
    const int col_dist = col->poc - col_ref_poc, cur_dist = mp.cur_poc - ref_poc;
    if (cur_ref_lt || col_dist == cur_dist) {
        out[0] = m->mv[sel][0];
        out[1] = m->mv[sel][1];
    } else {
        scale_mv(m->mv[sel], col_dist, cur_dist, out);
    }
    return true;
}

int derive_merge_list(const Neighbours& nb, const MotionParams& mp, const PuGeom& pu, MergeCand* c) {
    const int max = std::clamp(mp.max_merge, 1, 5);
    int n = 0;
    auto push = [&](const MergeCand& m) {
        if (n < max) c[n++] = m;
    };
    const int x = pu.x, y = pu.y, w = pu.w, h = pu.h;
    MergeCand a1, b1, t;
    const bool ha1 = spatial(nb, mp, pu, x - 1, y + h - 1, A1, a1);
    if (ha1) push(a1);
    const bool hb1 = spatial(nb, mp, pu, x + w - 1, y - 1, B1, b1);
    if (hb1 && !(ha1 && same_motion(a1, b1))) push(b1);
    if (spatial(nb, mp, pu, x + w, y - 1, B0, t) && !(hb1 && same_motion(b1, t))) push(t);
    if (spatial(nb, mp, pu, x - 1, y + h, A0, t) && !(ha1 && same_motion(a1, t))) push(t);
    if (n < 4 && spatial(nb, mp, pu, x - 1, y - 1, B2, t) && !(ha1 && same_motion(a1, t)) &&
        !(hb1 && same_motion(b1, t)))
        push(t);
    if (mp.tmvp && n < max && temporal_merge(mp, pu, t)) c[n++] = t;
    if (mp.is_b) combined(mp, c, n, max);
    zeros(mp, c, n, max);
    return n;
}

MergeCand derive_merge(const Neighbours& nb, const MotionParams& mp, const PuGeom& g, int merge_idx) {
    PuGeom q = g;
    if (mp.par_mrg_level > 2 && g.cu_w == 8) {
        q.x = g.cu_x;
        q.y = g.cu_y;
        q.w = g.cu_w;
        q.h = g.cu_h;
        q.part_idx = 0;
    }
    MergeCand c[5];
    const int n = derive_merge_list(nb, mp, q, c);
    MergeCand pick;
    if (merge_idx >= 0 && merge_idx < n) {
        pick = c[merge_idx];
    } else {  // src/motion.rs:401 zero_cand
        pick.pred = mp.is_b ? 3 : 1;
        pick.ref_idx[0] = 0;
        pick.ref_idx[1] = (int8_t)(mp.is_b ? 0 : -1);
    }
    if (g.w + g.h == 12 && pick.pred == 3) {
        pick.pred = 1;
        pick.ref_idx[1] = -1;
        pick.mv[1][0] = pick.mv[1][1] = 0;
    }
    return pick;
}

void derive_amvp(const Neighbours& nb, const MotionParams& mp, const PuGeom& pu, int list, int32_t ref_poc,
                 int16_t out[2][2]) {
    const int x = pu.x, y = pu.y, w = pu.w, h = pu.h;
    const int a_pos[2][2] = {{x - 1, y + h}, {x - 1, y + h - 1}};
    const int b_pos[3][2] = {{x + w, y - 1}, {x + w - 1, y - 1}, {x - 1, y - 1}};
    const bool is_scaled = inter_at(nb, mp, pu, a_pos[0][0], a_pos[0][1]) != nullptr ||
                           inter_at(nb, mp, pu, a_pos[1][0], a_pos[1][1]) != nullptr;
    // the target's long-term class: its entry in list 0, then list 1
    // (src/motion.rs:445 searches both lists for the POC)
    const bool target_lt = target_is_lt(mp, 0, ref_poc);
    int16_t a[2] = {0, 0}, b[2] = {0, 0};
    bool ha = amvp_spatial(nb, mp, pu, a_pos, 2, list, ref_poc, true, true, target_lt, a);
    bool hb = amvp_spatial(nb, mp, pu, b_pos, 3, list, ref_poc, false, true, target_lt, b);
    if (!is_scaled) {
        if (!ha && hb) {
            ha = true;
            a[0] = b[0];
            a[1] = b[1];
        }
        hb = amvp_spatial(nb, mp, pu, b_pos, 3, list, ref_poc, true, false, target_lt, b);
    }
    int n = 0;
    if (ha) {
        out[n][0] = a[0];
        out[n][1] = a[1];
        n++;
    }
    if (hb && !(ha && a[0] == b[0] && a[1] == b[1])) {
        out[n][0] = b[0];
        out[n][1] = b[1];
        n++;
    }
    if (n < 2 && mp.tmvp) {
        int16_t t[2];
        if (temporal_at_pu(mp, pu, list, ref_poc, t)) {
            out[n][0] = t[0];
            out[n][1] = t[1];
            n++;
        }
    }
    for (; n < 2; n++) out[n][0] = out[n][1] = 0;
}

bool motion_differs(const MotionCell& p, const MotionCell& q) {
    // references compared by POC, an unused list as "none"
    auto ref = [](const MotionCell& m, int l, int32_t& poc) {
        if (!((m.pred >> l) & 1)) return false;
        poc = m.poc[l];
        return true;
    };
    int32_t pp[2] = {0, 0}, qp[2] = {0, 0};
    const bool pu[2] = {ref(p, 0, pp[0]), ref(p, 1, pp[1])};
    const bool qu[2] = {ref(q, 0, qp[0]), ref(q, 1, qp[1])};
    auto eq = [](bool ua, int32_t a, bool ub, int32_t b) { return ua == ub && (!ua || a == b); };
    const bool direct = eq(pu[0], pp[0], qu[0], qp[0]) && eq(pu[1], pp[1], qu[1], qp[1]);
    const bool crossed = eq(pu[0], pp[0], qu[1], qp[1]) && eq(pu[1], pp[1], qu[0], qp[0]);
    if (!direct && !crossed) return true;
    auto mv = [](const MotionCell& m, int l, int c) -> int { return ((m.pred >> l) & 1) ? m.mv[l][c] : 0; };
    auto far = [&](int pl, int ql) {
        return std::abs(mv(p, pl, 0) - mv(q, ql, 0)) >= 4 || std::abs(mv(p, pl, 1) - mv(q, ql, 1)) >= 4;
    };
    if (!eq(pu[0], pp[0], pu[1], pp[1])) return direct ? (far(0, 0) || far(1, 1)) : (far(0, 1) || far(1, 0));
    return (far(0, 0) || far(1, 1)) && (far(0, 1) || far(1, 0));
}

}  // namespace hpvcd

// ---------------------------------------------------------------------------
// probe: the derivations on a caller-built neighbourhood (unit parity tests)
// ---------------------------------------------------------------------------
#include "../../include/hpvcd.h"
#include "internal.h"

namespace {

struct ProbeNb : hpvcd::Neighbours {
    const hpvcd_motion_case* c;
    std::vector<hpvcd::MotionCell> cells;
    int gw = 0;
    bool available(int x, int y) const override {
        if (x < 0 || y < 0 || x >= c->width || y >= c->height) return false;
        return c->cur[(size_t)(y >> 2) * gw + (x >> 2)].avail != 0;
    }
    const hpvcd::MotionCell& motion_at(int x, int y) const override { return cells[(size_t)(y >> 2) * gw + (x >> 2)]; }
};

hpvcd::MotionCell cell_from(const hpvcd_mv_cell& m) {
    hpvcd::MotionCell o;
    o.intra = m.inter ? 0 : 1;
    o.pred = m.pred;
    o.lt = m.lt;
    for (int l = 0; l < 2; l++) {
        o.ref_idx[l] = m.ref_idx[l];
        o.mv[l][0] = m.mv[l][0];
        o.mv[l][1] = m.mv[l][1];
        o.poc[l] = m.ref_poc[l];
    }
    return o;
}

}  // namespace

extern "C" hpvcd_status hpvcd_motion_probe(const hpvcd_motion_case* c, hpvcd_motion_result* r) {
    using namespace hpvcd;
    if (!c || !r || !c->cur) return set_error(HPVCD_E_INVALID, "null argument");
    if (c->width <= 0 || c->height <= 0 || c->width % 4 || c->height % 4 || c->width > 4096 || c->height > 4096 ||
        c->num_ref[0] < 0 || c->num_ref[0] > 16 || c->num_ref[1] < 0 || c->num_ref[1] > 16 || c->log2_ctb < 4 ||
        c->log2_ctb > 6 || c->cu_size < 8 || c->cu_x < 0 || c->cu_y < 0 || c->cu_x + c->cu_size > c->width ||
        c->cu_y + c->cu_size > c->height || c->pu_x < c->cu_x || c->pu_y < c->cu_y || c->pu_w < 4 || c->pu_h < 4 ||
        c->pu_x + c->pu_w > c->cu_x + c->cu_size || c->pu_y + c->pu_h > c->cu_y + c->cu_size)
        return set_error(HPVCD_E_INVALID, "motion probe: geometry");
    ProbeNb nb;
    nb.c = c;
    nb.gw = c->width / 4;
    const size_t ncell = (size_t)nb.gw * (c->height / 4);
    nb.cells.resize(ncell);
    for (size_t i = 0; i < ncell; i++) nb.cells[i] = cell_from(c->cur[i]);
    ColField col;
    if (c->col) {
        col.poc = c->col_poc;
        col.cw = (c->width + 15) >> 4;
        col.ch = (c->height + 15) >> 4;
        for (int by = 0; by < col.ch; by++)
            for (int bx = 0; bx < col.cw; bx++) col.cells.push_back(cell_from(c->col[(size_t)(by * 4) * nb.gw + bx * 4]));
    }
    std::vector<RefPic> lists[2];
    for (int l = 0; l < 2; l++)
        for (int i = 0; i < c->num_ref[l]; i++) {
            RefPic p;
            p.poc = c->ref_poc[l][i];
            p.lt = c->ref_lt[l][i] != 0;
            lists[l].push_back(p);
        }
    MotionParams mp;
    mp.width = c->width;
    mp.height = c->height;
    mp.is_b = c->is_b != 0;
    mp.cur_poc = c->cur_poc;
    mp.log2_ctb = c->log2_ctb;
    mp.par_mrg_level = c->par_mrg_level;
    mp.max_merge = c->max_merge;
    mp.tmvp = c->tmvp != 0;
    mp.col_from_l0 = c->col_from_l0 != 0;
    mp.col = c->col ? &col : nullptr;
    for (int l = 0; l < 2; l++) {
        mp.list[l] = lists[l].data();
        mp.num[l] = (int)lists[l].size();
    }
    PuGeom g;
    g.x = c->pu_x;
    g.y = c->pu_y;
    g.w = c->pu_w;
    g.h = c->pu_h;
    g.cu_x = c->cu_x;
    g.cu_y = c->cu_y;
    g.cu_w = g.cu_h = c->cu_size;
    g.part_idx = c->part_idx;
    std::memset(r, 0, sizeof(*r));
    MergeCand list[5];
    r->num_merge = derive_merge_list(nb, mp, g, list);
    for (int k = 0; k < 5; k++) {
        const MergeCand m = derive_merge(nb, mp, g, k);
        hpvcd_mv_cell& o = r->merge[k];
        o.inter = 1;
        o.pred = m.pred;
        for (int l = 0; l < 2; l++) {
            o.ref_idx[l] = m.ref_idx[l];
            o.mv[l][0] = m.mv[l][0];
            o.mv[l][1] = m.mv[l][1];
        }
    }
    for (int l = 0; l < 2; l++)
        for (int i = 0; i < mp.num[l]; i++) derive_amvp(nb, mp, g, l, lists[l][(size_t)i].poc, r->amvp[l][i]);
    return HPVCD_OK;
}
