"""Pins of the oracle's normative tables and boundary rules through closed forms.

Each table the oracle (and the GPU back end) restates is checked against the
reference's own table (tests/golden/ref_tables.json, extracted from the
reference source by tests/golden/make_tables_fixture.py) through an output
whose value follows from the H.265 equations in closed form, so a wrong entry,
a shifted index or a transposed operand changes the result:

* intra angles (src/intra.rs:31): a ramp reference of slope 32 makes every
  angular prediction exact, pred = 32 (x + 1) + c + (y + 1) * angle;
* inverse angles (src/intra.rs:37): the left column is laid out so that the
  projected reference row is the same ramp, which only holds when the
  projection uses the reference's inverse angle;
* beta / tC (src/deblock.rs:35,44): filtered-or-not thresholds of luma
  segments whose second derivative sweeps d, and the weak-filter clip at
  |q0 - p0| = 5 tC, at every QP, bS 1 and 2, slice offsets and 10 bits;
* chroma deblocking (src/decode.rs:2026-2142): the chroma 8-sample grid, bS 2
  only, the clip of Delta to tC at QpC = qpc(qPL + pps offset) (only the PPS
  offsets count, hazard A4);
* slice and tile gates of deblocking and SAO (src/decode.rs:1712,
  src/sao.rs:380);
* MC taps (src/mc.rs:30,37): the response to a single-sample impulse at 12
  bits reads back tap_x * tap_y for every fraction.

The pictures are I_PCM pictures (tests/handmade.py), so their samples are
chosen exactly; the same record sets run through the GPU in
tests/test_gpu_pins.py.
"""
import json
import os

import numpy as np
import pytest

import handmade as M
from hpvcd_amd import _native as N
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
TABLES = json.load(open(os.path.join(HERE, "golden", "ref_tables.json")))
ANGLE = TABLES["intra_pred_angle"]            # modes 2..34
INV_ANGLE = TABLES["inv_angle"]               # modes 11..25
BETA = TABLES["beta"]
TC = TABLES["tc"]


# --------------------------------------------------------------------------------------
# intra angles and inverse angles
# --------------------------------------------------------------------------------------
def _angular(mode, n, corner, above, left, bd=12):
    return O.intra_predict(corner, np.asarray(above), np.asarray(left), n, mode, bd=bd, smoothing=False,
                           edge_filters=False).astype(np.int64)


def _ramp_refs(n, c, main_is_above, angle, inv):
    """corner, above[0..2n), left[0..2n): the main reference row holds the ramp
    ref[k] = 32 k + c (k = 0 is the corner), and for a negative angle the side
    column holds, at each sample the projection ref[x] = side[-1 + ((x * inv +
    128) >> 8)] reads (x = (n * angle) >> 5 .. -1), the ramp value 32 x + c."""
    main = [32 * (k + 1) + c for k in range(2 * n)]
    side = [0] * (2 * n)
    if angle < 0 and (n * angle) >> 5 < -1:  # the projection exists only then (H.265 8.4.4.2.6)
        for x in range((n * angle) >> 5, 0):
            j = -1 + ((x * inv + 128) >> 8)
            assert 0 <= j < 2 * n
            side[j] = 32 * x + c
    return (c, main, side) if main_is_above else (c, side, main)


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_angular_modes_match_reference_angle_tables(n):
    c = 32 * 33 + 7  # keeps 32 x + c > 0 for the projected part of the row
    for mode in range(2, 35):
        angle = ANGLE[mode - 2]
        inv = INV_ANGLE[mode - 11] if 11 <= mode <= 25 else 0
        vertical = mode >= 18
        corner, above, left = _ramp_refs(n, c, vertical, angle, inv)
        p = _angular(mode, n, corner, above, left)
        y, x = np.mgrid[0:n, 0:n]
        if vertical:
            want = 32 * (x + 1) + c + (y + 1) * angle
        else:
            want = 32 * (y + 1) + c + (x + 1) * angle
        assert np.array_equal(p, want), f"mode {mode} n {n}"


def test_inverse_angle_changes_the_prediction():
    """The projection is not vacuous: a wrong inverse angle breaks the ramp."""
    n, c = 32, 32 * 33 + 7
    for mode in (11, 15, 18, 21, 25):
        angle, inv = ANGLE[mode - 2], INV_ANGLE[mode - 11]
        wrong = (inv * 7) // 8 if inv < -256 else inv - 64
        corner, above, left = _ramp_refs(n, c, mode >= 18, angle, wrong)
        p = _angular(mode, n, corner, above, left)
        y, x = np.mgrid[0:n, 0:n]
        want = (32 * (x + 1) + c + (y + 1) * angle) if mode >= 18 else (32 * (y + 1) + c + (x + 1) * angle)
        assert not np.array_equal(p, want), mode


# --------------------------------------------------------------------------------------
# beta / tC through luma segment decisions
# --------------------------------------------------------------------------------------
def _beta_tc_picture(qp, bd=8, bs=2, slices=((0, 0, 0, 1),), base=100, tc_steps=()):
    """A 16-wide 4:0:0 PCM picture with one vertical edge at x = 8 (bS bs on
    every cell of it, no horizontal edges).  Segment k (rows 4k..4k+3) for
    k < 66 has second derivative dp0 = k // 2 on line 0 and dp3 = k - k // 2
    on line 3 (d = k) on the p side and a flat q side 4 << (bd - 8) higher;
    then one segment per entry of tc_steps with flat sides that step by it."""
    s = 1 << (bd - 8)
    nseg = 66 + len(tc_steps)
    h = 8 * ((nseg + 1) // 2)
    y = np.full((h, 16), base * s, np.int64)
    for k in range(nseg):
        r = 4 * k
        if k < 66:
            for line, curv in ((0, k // 2), (1, k // 2), (2, k // 2), (3, k - k // 2)):
                y[r + line, 5] = (base + curv) * s  # p2: |p2 - 2 p1 + p0| = curv
            y[r:r + 4, 8:] = (base + 4) * s
        else:
            y[r:r + 4, 8:] = base * s + tc_steps[k - 66]
    gw, gh = 4, h // 4
    cw = np.array([M.cell_word(bsv=bs if gx == 2 else 0, extra=N.CELL_PCM, qp=qp) for gy in range(gh)
                   for gx in range(gw)], np.uint32)
    return M.pcm_picture([y], 0, bd, cw, slices=slices), y


def _segment_changes(rs, p, y_in):
    out = O.picture_planes(rs.batch, p, N.STAGE_DEBLOCK)[0].astype(np.int64)
    q0 = out[0::4, 8] - y_in[0::4, 8]
    p0 = out[0::4, 7] - y_in[0::4, 7]
    return p0, q0


def test_beta_and_tc_tables_at_every_qp():
    qps = list(range(0, 52))
    pics, ins = [], []
    for q in qps:
        tc = TC[min(q + 2, 53)]
        pic, y = _beta_tc_picture(q, tc_steps=(5 * max(tc, 1),))
        pics.append(pic)
        ins.append(y)
    rs = M.batch_of(pics)
    for i, q in enumerate(qps):
        p0, q0 = _segment_changes(rs, i, ins[i])
        beta, tc = BETA[q], TC[min(q + 2, 53)]
        filtered = q0[:66] != 0
        if tc == 0:
            assert not filtered.any(), q
        else:
            assert np.array_equal(filtered, np.arange(66) < beta), f"qp {q}: beta {beta}"
        # weak filter at a step of 5 tC: p0 and q0 move by exactly tC
        assert q0[66] == -tc and p0[66] == tc, f"qp {q}: tc {tc}"


def test_tc_index_of_bs1_and_slice_offsets():
    cases = [(30, 1, (0, 0, 0, 1)), (40, 1, (0, 0, 0, 1)), (24, 2, (0, 2, -1, 1)), (45, 2, (0, -3, 3, 1)),
             (20, 2, (0, 6, 6, 1)), (50, 1, (0, -6, 2, 1))]
    pics, ins, want = [], [], []
    for q, bs, sl in cases:
        beta = BETA[min(max(q + 2 * sl[1], 0), 51)]
        tc = TC[min(max(q + 2 * (bs - 1) + 2 * sl[2], 0), 53)]
        pic, y = _beta_tc_picture(q, bs=bs, slices=(sl,), tc_steps=(5 * max(tc, 1),))
        pics.append(pic)
        ins.append(y)
        want.append((beta, tc))
    rs = M.batch_of(pics)
    for i, (beta, tc) in enumerate(want):
        p0, q0 = _segment_changes(rs, i, ins[i])
        if tc:
            assert np.array_equal(q0[:66] != 0, np.arange(66) < beta), cases[i]
        assert q0[66] == -tc and p0[66] == tc, cases[i]


def test_beta_tc_scale_with_bit_depth():
    """10 bits: beta' = beta << 2 and tC' = tC << 2 (d thresholds 4x)."""
    for q in (22, 37, 51):
        beta, tc = BETA[q] << 2, TC[q + 2] << 2
        pic, y = _beta_tc_picture(q, bd=10, tc_steps=(5 * tc,))
        rs = M.batch_of([pic])
        p0, q0 = _segment_changes(rs, 0, y)
        # the d sweep is in 10-bit units of 4: d = 4k
        assert np.array_equal(q0[:66] != 0, 4 * np.arange(66) < beta), q
        assert q0[66] == -tc and p0[66] == tc, q


# --------------------------------------------------------------------------------------
# chroma deblocking
# --------------------------------------------------------------------------------------
def _chroma_picture(qp_p, qp_q, pps_cb, pps_cr, bs=2, step=100):
    """4:2:0 32x16 PCM picture: luma flat, Cb/Cr flat 60 left of chroma x = 8
    (luma 16) and 60 + step right of it, plus steps at chroma x = 4 and 12
    (luma 8, 24: luma edges but not on the chroma 8-sample grid).  Luma bS bs
    at x = 8, 16, 24; the cells left of x = 16 carry QP qp_p, the others qp_q."""
    w, h = 32, 16
    y = np.full((h, w), 100, np.int64)
    cb = np.full((h // 2, w // 2), 60, np.int64)
    cb[:, 8:] += step
    cb[:, 4:8] += 7
    cb[:, 12:] += 9
    cr = cb.copy()
    cw = np.array([M.cell_word(bsv=bs if gx in (2, 4, 6) else 0, extra=N.CELL_PCM, qp=qp_p if gx < 4 else qp_q)
                   for gy in range(h // 4) for gx in range(w // 4)], np.uint32)
    return M.pcm_picture([y, cb, cr], 1, 8, cw, pps_cb_qp_offset=pps_cb, pps_cr_qp_offset=pps_cr), (y, cb, cr)


def test_chroma_deblocking_grid_qp_and_clip():
    cases = [(30, 30, 0, 0), (37, 41, 6, -4), (45, 45, -12, 12), (27, 35, 3, 5), (51, 51, 0, -7)]
    pics, ins = [], []
    for c in cases:
        pic, planes = _chroma_picture(*c)
        pics.append(pic)
        ins.append(planes)
    pics.append(_chroma_picture(40, 40, 0, 0, bs=1)[0])
    ins.append(_chroma_picture(40, 40, 0, 0, bs=1)[1])
    rs = M.batch_of(pics)
    for i, (qp, qq, ocb, ocr) in enumerate(cases + [(40, 40, 0, 0)]):
        out = O.picture_planes(rs.batch, i, N.STAGE_DEBLOCK)
        qpl = (qp + qq + 1) >> 1
        for comp, off in ((1, ocb), (2, ocr)):
            d = out[comp].astype(np.int64) - ins[i][comp]
            if i == len(cases):  # bS 1: chroma is never filtered
                assert not d.any()
                continue
            tc = TC[min(O.lib().oracle_qpc(qpl + off, 1, 8) + 2, 53)]
            delta = min(tc, (3 * 100 + 4) >> 3)
            assert np.all(d[:, 7] == delta) and np.all(d[:, 8] == -delta), (cases[i], comp)
            # only p0 and q0 move, and only at the chroma 8-grid edge
            keep = np.ones(16, bool)
            keep[[7, 8]] = False
            assert not d[:, keep].any(), (cases[i], comp)


# --------------------------------------------------------------------------------------
# slice and tile gates of deblocking and SAO
# --------------------------------------------------------------------------------------
def _gate_picture(axis, gate, sao=False, lf_across=(1, 1)):
    """32x32 4:0:0 PCM picture, CTB 16, 8x8 CUs whose flat levels step along
    one axis (rows for axis 0, columns for axis 1), bS 2 on every 8-grid edge
    across that axis.  gate 'slice': CTB row 1 is slice 1, lf_across per
    slice; gate 'tile': CTB column 1 is tile 1 (loop filter across tiles off).
    With sao, every CTB takes edge offset along the axis with offsets that
    move every sample that is a local extremum."""
    w = h = 32
    vals = np.array([40, 90, 30, 100])
    if axis == 0:
        y = np.repeat(vals, 8)[:, None] * np.ones((1, w), np.int64)
        y += (np.arange(h) % 2 == 1)[:, None] * 3  # every sample a local extremum along the axis
    else:
        y = np.ones((h, 1), np.int64) * np.repeat(vals, 8)[None, :]
        y += (np.arange(w) % 2 == 1)[None, :] * 3
    cw = []
    for gy in range(h // 4):
        for gx in range(w // 4):
            bsv = 2 if axis == 1 and gx % 2 == 0 and gx else 0
            bsh = 2 if axis == 0 and gy % 2 == 0 and gy else 0
            sl = (1 if gy >= 4 else 0) if gate == "slice" else 0
            tl = (1 if gx >= 4 else 0) if gate == "tile" else 0
            cw.append(M.cell_word(bsv=bsv, bsh=bsh, extra=N.CELL_PCM, qp=37, slice_idx=sl, tile=tl))
    flags = 0
    slices = ((0, 0, 0, 1),)
    if gate == "slice":
        slices = ((0, 0, 0, lf_across[0]), (0, 0, 0, lf_across[1]))
    if gate == "tile":
        flags |= N.PF_TILES
    if gate == "tile_across":
        flags |= N.PF_TILES | N.PF_LF_ACROSS_TILES
    sao_recs = None
    if sao:
        flags |= N.PF_SAO_LUMA
        sao_recs = []
        for _ in range(4):
            s = N.Sao()
            s.type_idx[0] = 2
            s.eo_class[0] = 1 if axis == 0 else 0
            for k, v in enumerate((5, 2, -2, -5)):
                s.offsets[0][k] = v
            sao_recs.append(s)
    return M.pcm_picture([y], 0, 8, np.array(cw, np.uint32), flags=flags, slices=slices, sao=sao_recs,
                         deblock=not sao), y


def _decode(pics, stage):
    rs = M.batch_of([p for p, _ in pics])
    return [O.picture_planes(rs.batch, i, stage)[0].astype(np.int64) for i in range(len(pics))]


def test_deblocking_slice_gate_uses_the_q_side_slice():
    pics = [_gate_picture(0, "slice", lf_across=(1, 0)), _gate_picture(0, "slice", lf_across=(0, 1)),
            _gate_picture(0, "slice", lf_across=(1, 1))]
    gated, open_p_side, control = _decode(pics, N.STAGE_DEBLOCK)
    y_in = pics[0][1]
    band = slice(13, 19)  # p2..q2 of the edge at y = 16
    assert np.array_equal(gated[band], y_in[band])          # q slice forbids: untouched
    assert not np.array_equal(control[band], y_in[band])    # control: filtered
    assert np.array_equal(open_p_side, control)             # only the q side's flag counts
    for edge in (8, 24):                                     # edges inside a slice: filtered in all
        b = slice(edge - 3, edge + 3)
        assert np.array_equal(gated[b], control[b]) and not np.array_equal(gated[b], y_in[b])


def test_deblocking_tile_gate():
    gated, control = _decode([_gate_picture(1, "tile"), _gate_picture(1, "tile_across")], N.STAGE_DEBLOCK)
    y_in = _gate_picture(1, "tile")[1]
    band = slice(13, 19)
    assert np.array_equal(gated[:, band], y_in[:, band])
    assert not np.array_equal(control[:, band], y_in[:, band])
    for edge in (8, 24):
        b = slice(edge - 3, edge + 3)
        assert np.array_equal(gated[:, b], control[:, b]) and not np.array_equal(gated[:, b], y_in[:, b])


def test_sao_slice_and_tile_gates():
    for axis, gate, ctrl in ((0, "slice", "slice_open"), (1, "tile", "tile_across")):
        if ctrl == "slice_open":
            pics = [_gate_picture(axis, gate, sao=True, lf_across=(1, 0)),
                    _gate_picture(axis, gate, sao=True, lf_across=(1, 1))]
        else:
            pics = [_gate_picture(axis, gate, sao=True), _gate_picture(axis, ctrl, sao=True)]
        gated, control = _decode(pics, N.STAGE_SAO)
        y_in = pics[0][1]
        g = gated if axis == 0 else gated.T
        c = control if axis == 0 else control.T
        x = y_in if axis == 0 else y_in.T
        # the two samples whose edge-offset neighbour lies across the boundary
        assert np.array_equal(g[[15, 16]], x[[15, 16]]), gate
        assert not np.array_equal(c[[15, 16]], x[[15, 16]]), gate
        inner = [r for r in range(1, 31) if r not in (15, 16)]
        assert np.array_equal(g[inner], c[inner]) and not np.array_equal(g[inner], x[inner]), gate


# --------------------------------------------------------------------------------------
# MC interpolation taps through impulse responses
# --------------------------------------------------------------------------------------
LUMA_F = TABLES["luma_filter"]
CHROMA_F = TABLES["chroma_filter"]


def _impulse_case(mv, delta, bd=12):
    """A 32x32 4:2:0 reference, flat at 1 << (bd - 1) with one impulse of
    delta at luma (16, 16) and chroma (8, 8); the inter picture is tiled by
    8x8 PUs that all take vector mv (uni-prediction, no weights)."""
    base = 1 << (bd - 1)
    y = np.full((32, 32), base, np.int64)
    cb = np.full((16, 16), base, np.int64)
    y[16, 16] += delta
    cb[8, 8] += delta
    return M.inter_case(y, cb, cb.copy(), mvs=[mv], pu=8, bd=bd), base


def _expect(base, taps, xs, ys, ipos, mv, delta, chroma, bd=12):
    """Uni-predicted samples at (xs, ys) over a flat reference holding one
    impulse: integer position x + (mv >> 2) (luma) or x + (mv >> 3) (4:2:0
    chroma), stage 1 >> (bd - 8), stage 2 >> 6, then the uni rounding
    >> (14 - bd) (src/mc.rs:351-560)."""
    sh, mask = (3, 7) if chroma else (2, 3)
    ntap = len(taps[0])
    half = ntap // 2 - 1
    kx = ipos[0] - (xs + (mv[0] >> sh)) + half
    ky = ipos[1] - (ys + (mv[1] >> sh)) + half
    tx = np.where((kx >= 0) & (kx < ntap), np.take(taps[mv[0] & mask], np.clip(kx, 0, ntap - 1)), 0)
    ty = np.where((ky >= 0) & (ky < ntap), np.take(taps[mv[1] & mask], np.clip(ky, 0, ntap - 1)), 0)
    v = (ty * ((tx * delta) >> (bd - 8))) >> 6
    return base + ((v + (1 << (13 - bd))) >> (14 - bd))


IMPULSE_MVS = [(16 + fx, 16 + fy) for fy in range(8) for fx in range(8)]


@pytest.mark.parametrize("delta", [2000, -2000])
def test_mc_taps_impulse_response_every_fraction(delta):
    """Luma quarter- and chroma eighth-sample taps: every output sample of
    every PU reads back base + tap_x * tap_y * delta through the stages'
    floors.  The PU at (8, 8) sees the impulse under every tap pair of its
    fractions (luma fractions mv & 3 for mv & 7 < 4, all chroma fractions)."""
    for mv in IMPULSE_MVS:
        rs, base = _impulse_case(mv, delta)
        got = O.picture_planes(rs.batch, 1, N.STAGE_RECON)
        ys, xs = np.mgrid[0:32, 0:32]
        assert np.array_equal(got[0], _expect(base, LUMA_F, xs, ys, (16, 16), mv, delta, False)), (mv, "luma")
        cys, cxs = np.mgrid[0:16, 0:16]
        want_c = _expect(base, CHROMA_F, cxs, cys, (8, 8), mv, delta, True)
        assert np.array_equal(got[1], want_c) and np.array_equal(got[2], want_c), (mv, "chroma")


def test_mc_impulse_positions_cover_every_tap():
    """The impulse case is not vacuous: in the PU at (8, 8) every tap index
    of the 8-tap (luma) and 4-tap (chroma) filters multiplies the impulse."""
    for mv in IMPULSE_MVS:
        kl = {16 - (x + (mv[0] >> 2)) + 3 for x in range(8, 16)}
        kc = {8 - (x + (mv[0] >> 3)) + 1 for x in range(4, 8)}
        if (mv[0] & 7) < 4:
            assert kl == set(range(8)), mv
        assert kc == set(range(4)), mv
