"""Pins of the parser's motion derivations (CPU).

The generator and the parser share one implementation of the merge, AMVP
and temporal derivations (hpvcd_amd/csrc/motion.cpp), so the stream round
trip cannot catch a mistake in them.  They are pinned here, through
hpvcd_motion_probe, against:

* the reference's own unit tests of src/motion.rs:616-896, restated as
  scenes (the probe has no stub neighbours, so each test's condition is
  built from real cells);
* closed forms of the POC-distance scaling (8.5.3.2.9 at known distances);
* oracle/motion.py, an independent pure-Python restatement of
  src/motion.rs:41-575, src/decode.rs:6180 (temporal) and :6626
  (derive_pu_merge), on a few thousand random scenes: random motion,
  availability, long-term flags, reference lists and collocated fields,
  PUs of every partition shape, merge levels 0..6, list sizes 1..5, P and
  B slices, temporal prediction on and off.
"""
import ctypes as C
import os
import random
import sys

import pytest

import hpvcd_amd as H
from hpvcd_amd import N

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import motion as M  # noqa: E402


def _cell(inter=0, pred=0, lt=0, avail=0, ref_idx=(-1, -1), mv=((0, 0), (0, 0)), ref_poc=(0, 0)):
    return dict(inter=inter, pred=pred, lt=lt, avail=avail, ref_idx=tuple(ref_idx),
                mv=(tuple(mv[0]), tuple(mv[1])), ref_poc=tuple(ref_poc))


def _scene(width=64, height=64, **kw):
    gw, gh = width // 4, height // 4
    s = dict(width=width, height=height, log2_ctb=6, par_mrg_level=2, max_merge=5, is_b=0, tmvp=0, cur_poc=8,
             col_poc=0, col_from_l0=1, lists=([], []), cur=[_cell() for _ in range(gw * gh)], col=None,
             pu=(0, 0, 16, 16), cu=(0, 0, 16), part_idx=0)
    s.update(kw)
    return s


def _set(s, grid, x, y, **kw):
    s[grid][(y >> 2) * (s["width"] // 4) + (x >> 2)] = _cell(**kw)


def _mv_cells(cells):
    arr = (N.MvCell * len(cells))()
    for i, c in enumerate(cells):
        a = arr[i]
        a.inter, a.pred, a.lt, a.avail = c["inter"], c["pred"], c["lt"], c["avail"]
        for l in range(2):
            a.ref_idx[l] = c["ref_idx"][l]
            a.mv[l][0], a.mv[l][1] = c["mv"][l]
            a.ref_poc[l] = c["ref_poc"][l]
    return arr


def probe(s):
    """The C++ derivations on scene s: (merge list length, merge motion for
    merge_idx 0..4, AMVP predictors per list and reference index)."""
    c = N.MotionCase()
    for k in ("width", "height", "log2_ctb", "par_mrg_level", "max_merge", "is_b", "tmvp", "cur_poc", "col_poc",
              "col_from_l0"):
        setattr(c, k, int(s[k]))
    for l in range(2):
        c.num_ref[l] = len(s["lists"][l])
        for i, (poc, lt) in enumerate(s["lists"][l]):
            c.ref_poc[l][i] = poc
            c.ref_lt[l][i] = lt
    c.pu_x, c.pu_y, c.pu_w, c.pu_h = s["pu"]
    c.cu_x, c.cu_y, c.cu_size = s["cu"]
    c.part_idx = s["part_idx"]
    cur = _mv_cells(s["cur"])
    c.cur = C.cast(cur, C.POINTER(N.MvCell))
    col = None
    if s["col"] is not None:
        col = _mv_cells(s["col"])
        c.col = C.cast(col, C.POINTER(N.MvCell))
    r = N.MotionResult()
    N.check(N.lib().hpvcd_motion_probe(C.byref(c), C.byref(r)))
    merge = []
    for k in range(5):
        m = r.merge[k]
        merge.append((m.pred, (m.ref_idx[0], m.ref_idx[1]),
                      ((m.mv[0][0], m.mv[0][1]), (m.mv[1][0], m.mv[1][1]))))
    amvp = [[[(r.amvp[l][i][p][0], r.amvp[l][i][p][1]) for p in range(2)] for i in range(len(s["lists"][l]))]
            for l in range(2)]
    return r.num_merge, merge, amvp


def oracle(s):
    nb = M.Scene(s)
    x, y, w, h = s["pu"]
    cx, cy, cs = s["cu"]
    pu = (x, y, w, h, cx, cy, cs, cs, s["part_idx"])
    n = len(M.merge_list(nb, pu))
    merge = [M.pu_merge(nb, pu, k) for k in range(5)]
    amvp = [[[tuple(v) for v in M.amvp(nb, pu, l, poc)] for poc, _ in s["lists"][l]] for l in range(2)]
    return n, merge, amvp


# ---------------------------------------------------------------------------
# the reference's unit tests (src/motion.rs:616-896) as scenes
# ---------------------------------------------------------------------------
def test_scale_identity_and_halving_closed_forms():
    """scale_identity_when_equal_dist / scale_halves_on_double_dist
    (src/motion.rs:616-631) through the temporal candidate: with td = tb = 4
    the factor is (4 * ((16384 + 2) / 4) + 32) >> 6 = 256, the identity; with
    td = 8, tb = 4 it is 128, and 16 * 128 = 2048 rounds to 8."""
    for col_dist, cur_dist, mv, want in ((4, 4, (10, -6), (10, -6)), (8, 4, (16, 0), (8, 0)),
                                         (-2, 6, (5, -7), (-15, 21)), (3, 1, (9, -9), (3, -3))):
        s = _scene(tmvp=1, cur_poc=20, lists=([(20 - cur_dist, 0)], []), col_poc=40)
        s["col"] = [_cell() for _ in range(16 * 16)]
        _set(s, "col", 0, 0, inter=1, pred=1, ref_idx=(0, -1), mv=(mv, (0, 0)), ref_poc=(40 - col_dist, 0))
        n, merge, amvp = probe(s)
        assert merge[0] == (1, (0, -1), (want, (0, 0))), (col_dist, cur_dist)
        assert amvp[0][0][0] == want
        assert oracle(s)[1][0] == merge[0]
    # M.scale_mv restates the same equations: check it on the closed forms too
    assert M.scale_mv((10, -6), 4, 4) == (10, -6) and M.scale_mv((16, 0), 8, 4) == (8, 0)


def test_temporal_merge_keeps_l1_only_candidate():
    """temporal_merge_keeps_l1_only_candidate (src/motion.rs:664): a temporal
    candidate available for list 1 only keeps ref_idx -1 and a zero vector
    for list 0.  Here list 0's target is long-term and the collocated
    reference short-term, so list 0 has no temporal vector."""
    s = _scene(is_b=1, tmvp=1, cur_poc=8, lists=([(4, 1)], [(12, 0)]), col_poc=16)
    s["col"] = [_cell() for _ in range(16 * 16)]
    # the centre (8, 8) of the 16x16 PU reads the 16x16 block at (0, 0); the
    # collocated reference is 4 after it, as
    # list 1's target is 4 after the current picture: unscaled
    _set(s, "col", 0, 0, inter=1, pred=2, ref_idx=(-1, 0), mv=((0, 0), (12, -4)), ref_poc=(0, 20))
    n, merge, _ = probe(s)
    assert merge[0] == (2, (-1, 0), ((0, 0), (12, -4)))
    assert oracle(s)[1][0] == merge[0]


def test_parallel_merge_level_filters_same_region():
    """default_parallel_merge_level_filters_same_region (src/motion.rs:703):
    a spatial neighbour in the PU's merge estimation region is unavailable."""
    s = _scene(par_mrg_level=4, lists=([(4, 0)], []), pu=(8, 8, 8, 8), cu=(8, 8, 8))
    for x in range(0, 64, 4):
        for y in range(0, 64, 4):
            _set(s, "cur", x, y, inter=1, avail=1, pred=1, ref_idx=(0, -1), mv=((x, y), (0, 0)), ref_poc=(4, 0))
    n, merge, _ = probe(s)
    # A1 (7, 15) and B1 (15, 7), A0 (7, 16), B0 (16, 7), B2 (7, 7): all in the
    # 16x16 region of (8, 8) except A0 and B0
    found = {m[2][0] for m in merge}
    assert (4, 12) not in found and (12, 4) not in found and (4, 4) not in found
    assert (4, 16) in found and (16, 4) in found
    assert oracle(s)[1] == merge
    s["par_mrg_level"] = 2
    found = {m[2][0] for m in probe(s)[1]}
    assert (4, 12) in found and (12, 4) in found


def test_previous_pu_in_same_cu_is_available():
    """previous_pu_in_same_cu_does_not_require_decoded_map (src/motion.rs:759):
    a neighbour inside the current coding block is available although the
    picture-level availability says no."""
    s = _scene(par_mrg_level=0, lists=([(4, 0)], []), pu=(0, 8, 8, 8), cu=(0, 0, 16), part_idx=2)
    _set(s, "cur", 7, 7, inter=1, avail=0, pred=1, ref_idx=(0, -1), mv=((5, 6), (0, 0)), ref_poc=(4, 0))
    n, merge, amvp = probe(s)
    assert merge[0] == (1, (0, -1), ((5, 6), (0, 0)))
    assert amvp[0][0][0] == (5, 6)
    assert oracle(s)[1] == merge


def test_second_nxn_pu_rejects_lower_left_of_the_cu():
    """second_nxn_pu_rejects_unavailable_same_cu_region (src/motion.rs:816):
    for the second NxN PU the lower-left region inside the CU (7, 8) is not
    yet decoded and unavailable; (8, 8) is available."""
    s = _scene(par_mrg_level=0, lists=([(4, 0)], []), pu=(8, 0, 8, 8), cu=(0, 0, 16), part_idx=1)
    for x in range(0, 64, 4):
        for y in range(0, 64, 4):
            _set(s, "cur", x, y, inter=1, avail=1, pred=1, ref_idx=(0, -1), mv=((x, y), (0, 0)), ref_poc=(4, 0))
    n, merge, amvp = probe(s)
    found = {m[2][0] for m in merge}
    assert (4, 8) not in found          # A0 (7, 8)
    assert (4, 4) in found              # A1 (7, 7) inside PU 0
    assert oracle(s)[1] == merge and oracle(s)[2] == amvp


def test_zero_merge_candidates_cycle_the_reference_indices():
    """zero_merge_candidates_include_current_picture_reference_slot
    (src/motion.rs:866) without current-picture referencing: with no
    neighbours, merge_idx 1 of a P slice with two references is the zero
    vector on reference 1; AMVP pads with zero vectors."""
    s = _scene(max_merge=2, lists=([(4, 0), (2, 0)], []))
    n, merge, amvp = probe(s)
    assert n == 2
    assert merge[1] == (1, (1, -1), ((0, 0), (0, 0)))
    assert merge[3] == (1, (0, -1), ((0, 0), (0, 0)))   # past the list: the zero candidate
    assert amvp[0][0] == [(0, 0), (0, 0)]


def test_8x4_pu_keeps_list0_of_a_bi_candidate():
    """src/decode.rs:6641-6663: an 8x4 PU's bi-predictive merge candidate is
    used as list 0 only."""
    s = _scene(is_b=1, lists=([(4, 0)], [(12, 0)]), pu=(0, 8, 8, 4), cu=(0, 0, 8 * 2), part_idx=1)
    s["cu"] = (0, 4, 8)
    s["pu"] = (0, 8, 8, 4)
    _set(s, "cur", 0, 0, inter=1, avail=1, pred=3, ref_idx=(0, 0), mv=((3, 4), (5, 6)), ref_poc=(4, 12))
    for x in range(0, 64, 4):
        _set(s, "cur", x, 0, inter=1, avail=1, pred=3, ref_idx=(0, 0), mv=((3, 4), (5, 6)), ref_poc=(4, 12))
    n, merge, _ = probe(s)
    assert all(m[0] == 1 and m[1][1] == -1 and m[2][1] == (0, 0) for m in merge)
    assert oracle(s)[1] == merge


# ---------------------------------------------------------------------------
# random scenes against the Python restatement
# ---------------------------------------------------------------------------
PARTS = {  # partitions of a CU of size n: list of (x, y, w, h) offsets
    "2Nx2N": lambda n: [(0, 0, n, n)],
    "2NxN": lambda n: [(0, 0, n, n // 2), (0, n // 2, n, n // 2)],
    "Nx2N": lambda n: [(0, 0, n // 2, n), (n // 2, 0, n // 2, n)],
    "NxN": lambda n: [(0, 0, n // 2, n // 2), (n // 2, 0, n // 2, n // 2), (0, n // 2, n // 2, n // 2),
                      (n // 2, n // 2, n // 2, n // 2)],
    "2NxnU": lambda n: [(0, 0, n, n // 4), (0, n // 4, n, 3 * n // 4)],
    "2NxnD": lambda n: [(0, 0, n, 3 * n // 4), (0, 3 * n // 4, n, n // 4)],
    "nLx2N": lambda n: [(0, 0, n // 4, n), (n // 4, 0, 3 * n // 4, n)],
    "nRx2N": lambda n: [(0, 0, 3 * n // 4, n), (3 * n // 4, 0, n // 4, n)],
}


def _random_scene(rnd):
    width, height = rnd.choice([(64, 64), (96, 64), (128, 96), (80, 48)])
    is_b = rnd.random() < 0.6
    cur_poc = rnd.randrange(-20, 60)
    n0 = rnd.randrange(1, 5)
    n1 = rnd.randrange(1, 5) if is_b else 0
    pool = [cur_poc + d for d in range(-8, 9) if d]

    def entry():
        return (rnd.choice(pool), 1 if rnd.random() < 0.1 else 0)

    lists = ([entry() for _ in range(n0)], [entry() for _ in range(n1)])
    mv_range = rnd.choice([2, 6, 40])

    def rand_cell(avail_p, grid_lists):
        if rnd.random() < 0.25:
            return _cell(avail=int(rnd.random() < avail_p))
        pred = rnd.choice([1, 2, 3]) if is_b or rnd.random() < 0.3 else 1
        ref_idx, ref_poc, lt = [-1, -1], [0, 0], 0
        mv = [[0, 0], [0, 0]]
        for l in range(2):
            if pred >> l & 1:
                lst = grid_lists[l] if grid_lists[l] else grid_lists[0]
                i = rnd.randrange(len(lst)) if lst else 0
                ref_idx[l] = i
                if lst and rnd.random() < 0.8:
                    ref_poc[l], ltf = lst[i]
                else:
                    ref_poc[l], ltf = rnd.choice(pool), int(rnd.random() < 0.1)
                lt |= ltf << l
                mv[l] = [rnd.randint(-mv_range, mv_range), rnd.randint(-mv_range, mv_range)]
        return _cell(inter=1, pred=pred, lt=lt, avail=int(rnd.random() < avail_p), ref_idx=ref_idx,
                     mv=mv, ref_poc=ref_poc)

    gw, gh = width // 4, height // 4
    avail_p = rnd.choice([0.3, 0.8, 1.0])
    cur = [rand_cell(avail_p, lists) for _ in range(gw * gh)]
    # a few repeated motions so that pruning and combination matter
    if rnd.random() < 0.5:
        proto = rand_cell(1.0, lists)
        for i in rnd.sample(range(gw * gh), gw * gh // 3):
            cur[i] = dict(proto, avail=cur[i]["avail"])
    col = None
    col_poc = rnd.choice(pool + [cur_poc])
    if rnd.random() < 0.8:
        col_lists = ([(rnd.choice(pool), int(rnd.random() < 0.1)) for _ in range(3)],) * 2
        col = [rand_cell(1.0, col_lists) for _ in range(gw * gh)]
    cu_size = rnd.choice([8, 16, 32, 64])
    while cu_size > min(width, height):
        cu_size //= 2
    cx = rnd.randrange(0, width - cu_size + 1, cu_size)
    cy = rnd.randrange(0, height - cu_size + 1, cu_size)
    shapes = ["2Nx2N", "2NxN", "Nx2N"] + (["NxN"] if cu_size > 8 else []) + \
             (["2NxnU", "2NxnD", "nLx2N", "nRx2N"] if cu_size > 8 else [])
    parts = PARTS[rnd.choice(shapes)](cu_size)
    part_idx = rnd.randrange(len(parts))
    px, py, pw, ph = parts[part_idx]
    return dict(width=width, height=height, log2_ctb=rnd.choice([4, 5, 6]), par_mrg_level=rnd.choice([0, 2, 2, 3, 4, 6]),
                max_merge=rnd.randint(1, 5), is_b=int(is_b), tmvp=int(rnd.random() < 0.7), cur_poc=cur_poc,
                col_poc=col_poc, col_from_l0=rnd.randint(0, 1), lists=lists, cur=cur, col=col,
                pu=(cx + px, cy + py, pw, ph), cu=(cx, cy, cu_size), part_idx=part_idx)


@pytest.mark.parametrize("seed", range(6))
def test_random_scenes_match_the_restatement(seed):
    rnd = random.Random(1000 + seed)
    hits = dict(temporal=0, scaled=0, combined=0)
    for trial in range(400):
        s = _random_scene(rnd)
        got, want = probe(s), oracle(s)
        assert got == want, (seed, trial, s["pu"], s["cu"], s["part_idx"])
        hits["combined"] += sum(1 for m in got[1] if m[0] == 3)
        hits["temporal"] += int(s["tmvp"] and s["col"] is not None)
    assert hits["combined"] > 0 and hits["temporal"] > 0
