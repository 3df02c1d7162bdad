"""Pure-Python restatement of the reference's motion-vector derivations.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the C++
parser's merge / AMVP / temporal derivations (hpvcd_motion_probe), never by
the product.  It follows, function by function:

  scale_mv                 src/motion.rs:41
  prediction_block_available  src/motion.rs:131
  derive_merge             src/motion.rs:155 (spatial :238, temporal_merge :282,
                           derive_combined :333, derive_zero :375, zero_cand :401)
  derive_amvp              src/motion.rs:417 (amvp_spatial :515)
  temporal                 src/decode.rs:6180 (DecoderNeighbors::temporal)
  derive_pu_merge          src/decode.rs:6626 (singleMCLFlag, 8x4 / 4x8 bi -> L0)

A scene is a dict: width, height, log2_ctb, par_mrg_level, max_merge, is_b,
tmvp, cur_poc, col_poc, col_from_l0, lists (two lists of (poc, lt)), cur and
col (cell grids: row-major lists of dicts with keys inter, pred, lt, avail,
ref_idx, mv, ref_poc), and the PU: pu (x, y, w, h), cu (x, y, size), part_idx.
Motion is returned as (pred, (ref0, ref1), ((x0, y0), (x1, y1))).
"""
from __future__ import annotations


def _clip(lo, hi, v):
    return max(lo, min(hi, v))


def _trunc_div(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def scale_mv(mv, col_dist, cur_dist):
    td = _clip(-128, 127, col_dist)
    tb = _clip(-128, 127, cur_dist)
    if td == 0:
        return tuple(mv)
    tx = _trunc_div(16384 + (abs(td) >> 1), td)
    dsf = _clip(-4096, 4095, (tb * tx + 32) >> 6)

    def sc(c):
        p = dsf * c
        sign = (p > 0) - (p < 0)
        return _clip(-32768, 32767, sign * ((abs(p) + 127) >> 8))

    return (sc(mv[0]), sc(mv[1]))


class Scene:
    def __init__(self, s):
        self.s = s
        self.w, self.h = s["width"], s["height"]
        self.gw = self.w // 4

    def cell(self, grid, x, y):
        return grid[(y >> 2) * self.gw + (x >> 2)]

    # DecoderNeighbors::motion_at / available
    def motion_at(self, x, y):
        if x < 0 or y < 0 or x >= self.w or y >= self.h:
            return None
        return self.cell(self.s["cur"], x, y)

    def available(self, x, y):
        if x < 0 or y < 0 or x >= self.w or y >= self.h:
            return False
        return bool(self.cell(self.s["cur"], x, y)["avail"])

    def lists(self):
        return self.s["lists"]

    # DecoderNeighbors::temporal (the collocated picture given directly)
    def temporal(self, x, y, lst, ref_poc):
        s = self.s
        if s.get("col") is None:
            return None
        cur_poc = s["cur_poc"]
        if s["col_poc"] == cur_poc:
            return None
        if x >= self.w or y >= self.h:
            return None
        m = self.cell(s["col"], (x >> 4) << 4, (y >> 4) << 4)
        if not m["inter"]:
            return None
        l0, l1 = bool(m["pred"] & 1), bool(m["pred"] & 2)
        if not l0:
            sel = 1
        elif not l1:
            sel = 0
        else:
            all_before = all(p < cur_poc for p, _ in self.lists()[0]) and all(p < cur_poc for p, _ in self.lists()[1])
            sel = lst if all_before else (1 if s["col_from_l0"] else 0)
        mv_col = m["mv"][sel]
        col_ref_poc = m["ref_poc"][sel]
        col_ref_lt = bool((m["lt"] >> sel) & 1)
        cur_ref_lt = False
        for p, lt in self.lists()[lst]:
            if p == ref_poc:
                cur_ref_lt = bool(lt)
                break
        if cur_ref_lt != col_ref_lt:
            return None
        col_dist = s["col_poc"] - col_ref_poc
        cur_dist = cur_poc - ref_poc
        if cur_ref_lt or col_dist == cur_dist:
            return tuple(mv_col)
        return scale_mv(mv_col, col_dist, cur_dist)


def _pb_available(nb, pu, px, py):
    if px < 0 or py < 0:
        return False
    x, y, w, h, cux, cuy, cuw, cuh, part = pu
    same_cb = cux <= px < cux + cuw and cuy <= py < cuy + cuh
    if not same_cb:
        return nb.available(px, py)
    return not (w * 2 == cuw and h * 2 == cuh and part == 1 and cuy + h <= py and cux + w > px)


def _cand_of(m):
    return (m["pred"], tuple(m["ref_idx"]), (tuple(m["mv"][0]), tuple(m["mv"][1])))


def _spatial(nb, pu, px, py, pos, lvl):
    x, y, w, h, cux, cuy, cuw, cuh, part = pu
    if part == 1:
        vertical = w < cuw and h == cuh
        horizontal = h < cuh and w == cuw
        if (pos == "A1" and vertical) or (pos == "B1" and horizontal):
            return None
    if px >= 0 and py >= 0 and (x >> lvl) == (px >> lvl) and (y >> lvl) == (py >> lvl):
        return None
    if not _pb_available(nb, pu, px, py):
        return None
    m = nb.motion_at(px, py)
    if m is None or not m["inter"]:
        return None
    return _cand_of(m)


def _temporal_pos(nb, pu, lst, ref_poc):
    x, y, w, h = pu[:4]
    ctb = nb.s["log2_ctb"]
    br_y = y + h
    if (br_y >> ctb) == (y >> ctb):
        v = nb.temporal(x + w, br_y, lst, ref_poc)
        if v is not None:
            return v
    return nb.temporal(x + w // 2, y + h // 2, lst, ref_poc)


def merge_list(nb, pu):
    s = nb.s
    max_cand = s["max_merge"]
    is_b = bool(s["is_b"])
    lvl = s["par_mrg_level"]
    lists = nb.lists()
    cands = []

    def push(c):
        if len(cands) < max_cand:
            cands.append(c)

    x, y, w, h = pu[:4]
    a1 = _spatial(nb, pu, x - 1, y + h - 1, "A1", lvl)
    if a1 is not None:
        push(a1)
    b1 = _spatial(nb, pu, x + w - 1, y - 1, "B1", lvl)
    if b1 is not None and (a1 is None or a1 != b1):
        push(b1)
    c = _spatial(nb, pu, x + w, y - 1, "B0", lvl)
    if c is not None and (b1 is None or b1 != c):
        push(c)
    c = _spatial(nb, pu, x - 1, y + h, "A0", lvl)
    if c is not None and (a1 is None or a1 != c):
        push(c)
    if len(cands) < 4:
        c = _spatial(nb, pu, x - 1, y - 1, "B2", lvl)
        if c is not None and not ((a1 is not None and a1 == c) or (b1 is not None and b1 == c)):
            push(c)
    if s["tmvp"] and len(cands) < max_cand:
        cur = s["cur_poc"]
        r0 = lists[0][0][0] if lists[0] else cur
        mv0 = _temporal_pos(nb, pu, 0, r0)
        mv1 = None
        if is_b:
            r1 = lists[1][0][0] if lists[1] else cur
            mv1 = _temporal_pos(nb, pu, 1, r1)
        if mv0 is not None or mv1 is not None:
            pred = (1 if mv0 is not None else 0) | (2 if mv1 is not None else 0)
            cands.append((pred, (0 if mv0 is not None else -1, 0 if mv1 is not None else -1),
                          (mv0 if mv0 is not None else (0, 0), mv1 if mv1 is not None else (0, 0))))
    if is_b and 2 <= len(cands) < max_cand:
        L0 = [0, 1, 0, 2, 1, 2, 0, 3, 1, 3, 2, 3]
        L1 = [1, 0, 2, 0, 2, 1, 3, 0, 3, 1, 3, 2]
        n = len(cands)
        k = 0
        while len(cands) < max_cand and k < n * (n - 1):
            i0, i1 = L0[k % 12], L1[k % 12]
            k += 1
            if i0 >= n or i1 >= n:
                continue
            a, b = cands[i0], cands[i1]
            if (a[0] & 1) and (b[0] & 2):
                ra, rb = max(a[1][0], 0), max(b[1][1], 0)
                poc0 = lists[0][ra][0] if ra < len(lists[0]) else None
                poc1 = lists[1][rb][0] if rb < len(lists[1]) else None
                if poc0 != poc1 or a[2][0] != b[2][1]:
                    cands.append((3, (a[1][0], b[1][1]), (a[2][0], b[2][1])))
    n0, n1 = len(lists[0]), len(lists[1])
    num_ref = min(n0, n1) if is_b else n0
    if num_ref:
        z = 0
        while len(cands) < max_cand:
            r = z if z < num_ref else 0
            cands.append((3 if is_b else 1, (r, r if is_b else -1), ((0, 0), (0, 0))))
            z += 1
            if z > 32:
                break
    return cands


def pu_merge(nb, pu, merge_idx):
    """derive_pu_merge: the PU's merge motion for merge_idx."""
    x, y, w, h, cux, cuy, cuw, cuh, part = pu
    geom = pu
    if nb.s["par_mrg_level"] > 2 and cuw == 8:
        geom = (cux, cuy, cuw, cuh, cux, cuy, cuw, cuh, 0)
    cands = merge_list(nb, geom)
    is_b = bool(nb.s["is_b"])
    c = cands[merge_idx] if merge_idx < len(cands) else (3 if is_b else 1, (0, 0 if is_b else -1), ((0, 0), (0, 0)))
    if w + h == 12 and c[0] == 3:
        c = (1, (c[1][0], -1), (c[2][0], (0, 0)))
    return c


def _amvp_spatial(nb, pu, positions, lst, ref_poc, allow_scaled, same_ref, target_lt):
    cur_poc = nb.s["cur_poc"]
    if same_ref:
        for px, py in positions:
            if not _pb_available(nb, pu, px, py):
                continue
            m = nb.motion_at(px, py)
            if m is None or not m["inter"]:
                continue
            for l in (lst, 1 - lst):
                if (m["pred"] >> l) & 1 and m["ref_poc"][l] == ref_poc:
                    return tuple(m["mv"][l])
    if not allow_scaled:
        return None
    for px, py in positions:
        if not _pb_available(nb, pu, px, py):
            continue
        m = nb.motion_at(px, py)
        if m is None or not m["inter"]:
            continue
        for l in (lst, 1 - lst):
            if (m["pred"] >> l) & 1 and bool((m["lt"] >> l) & 1) == target_lt:
                if target_lt:
                    return tuple(m["mv"][l])
                return scale_mv(m["mv"][l], cur_poc - m["ref_poc"][l], cur_poc - ref_poc)
    return None


def amvp(nb, pu, lst, ref_poc):
    x, y, w, h = pu[:4]
    a_pos = [(x - 1, y + h), (x - 1, y + h - 1)]
    b_pos = [(x + w, y - 1), (x + w - 1, y - 1), (x - 1, y - 1)]

    def a_inter(px, py):
        if not _pb_available(nb, pu, px, py):
            return False
        m = nb.motion_at(px, py)
        return m is not None and bool(m["inter"])

    is_scaled = a_inter(*a_pos[0]) or a_inter(*a_pos[1])
    target_lt = False
    for p, lt in list(nb.lists()[0]) + list(nb.lists()[1]):
        if p == ref_poc:
            target_lt = bool(lt)
            break
    a = _amvp_spatial(nb, pu, a_pos, lst, ref_poc, True, True, target_lt)
    b = _amvp_spatial(nb, pu, b_pos, lst, ref_poc, False, True, target_lt)
    if not is_scaled:
        if a is None:
            a = b
        b = _amvp_spatial(nb, pu, b_pos, lst, ref_poc, True, False, target_lt)
    preds = []
    if a is not None:
        preds.append(a)
    if b is not None and a != b:
        preds.append(b)
    if len(preds) < 2 and nb.s["tmvp"]:
        t = _temporal_pos(nb, pu, lst, ref_poc)
        if t is not None:
            preds.append(t)
    while len(preds) < 2:
        preds.append((0, 0))
    return preds[:2]
