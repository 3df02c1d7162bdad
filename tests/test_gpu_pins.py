"""The closed-form pin record sets of test_oracle_pins.py through the GPU.

Each case runs the HIP back end on the same I_PCM / MC impulse records whose
results follow from the reference's tables in closed form, and compares
every sample with both the closed form and the oracle, so the kernels'
own copies of the beta/tC, chroma QP and interpolation tables are pinned
against the reference's, not only against the oracle's restatement.
"""
import numpy as np
import pytest

import handmade as M
import test_oracle_pins as P
from hpvcd_amd import _native as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _planes(gpu, rs, stage):
    out = []
    with gpu.job(rs) as job:
        job.run(N.MODE_BATCHED, stage)
        for p in range(rs.num_pictures):
            out.append(job.download_planes(p))
    return out


def _same_as_oracle(rs, got, stage):
    for p in range(rs.num_pictures):
        want = O.picture_planes(rs.batch, p, stage)
        for c, (g, w) in enumerate(zip(got[p], want)):
            if w is not None:
                assert np.array_equal(g, w), f"picture {p} plane {c}"


def test_gpu_beta_tc_tables(gpu):
    pics, ins, qps = [], [], list(range(0, 52))
    for q in qps:
        tc = P.TC[min(q + 2, 53)]
        pic, y = P._beta_tc_picture(q, tc_steps=(5 * max(tc, 1),))
        pics.append(pic)
        ins.append(y)
    rs = M.batch_of(pics)
    got = _planes(gpu, rs, N.STAGE_DEBLOCK)
    _same_as_oracle(rs, got, N.STAGE_DEBLOCK)
    for i, q in enumerate(qps):
        out = got[i][0].astype(np.int64)
        q0 = out[0::4, 8] - ins[i][0::4, 8]
        beta, tc = P.BETA[q], P.TC[min(q + 2, 53)]
        if tc:
            assert np.array_equal(q0[:66] != 0, np.arange(66) < beta), q
        assert q0[66] == -tc, q


def test_gpu_chroma_deblocking(gpu):
    cases = [(30, 30, 0, 0), (37, 41, 6, -4), (45, 45, -12, 12), (27, 35, 3, 5), (51, 51, 0, -7)]
    built = [P._chroma_picture(*c) for c in cases]
    rs = M.batch_of([b[0] for b in built])
    got = _planes(gpu, rs, N.STAGE_DEBLOCK)
    _same_as_oracle(rs, got, N.STAGE_DEBLOCK)
    for i, (qp, qq, ocb, ocr) in enumerate(cases):
        qpl = (qp + qq + 1) >> 1
        for comp, off in ((1, ocb), (2, ocr)):
            d = got[i][comp].astype(np.int64) - built[i][1][comp]
            tc = P.TC[min(O.lib().oracle_qpc(qpl + off, 1, 8) + 2, 53)]
            assert np.all(d[:, 7] == min(tc, 38)) and np.all(d[:, 8] == -min(tc, 38)), cases[i]


def test_gpu_slice_tile_gates(gpu):
    pics = [P._gate_picture(0, "slice", lf_across=(1, 0)), P._gate_picture(0, "slice", lf_across=(1, 1)),
            P._gate_picture(1, "tile"), P._gate_picture(1, "tile_across"),
            P._gate_picture(0, "slice", sao=True, lf_across=(1, 0)), P._gate_picture(1, "tile", sao=True)]
    rs = M.batch_of([p for p, _ in pics])
    for stage in (N.STAGE_DEBLOCK, N.STAGE_SAO):
        got = _planes(gpu, rs, stage)
        _same_as_oracle(rs, got, stage)
    got = _planes(gpu, rs, N.STAGE_DEBLOCK)
    y_in = pics[0][1]
    assert np.array_equal(got[0][0][13:19], y_in[13:19]) and not np.array_equal(got[1][0][13:19], y_in[13:19])


@pytest.mark.parametrize("delta", [2000, -2000])
def test_gpu_mc_taps_impulse(gpu, delta):
    for mv in P.IMPULSE_MVS:
        rs, base = P._impulse_case(mv, delta)
        with gpu.job(rs) as job:
            job.run(N.MODE_BATCHED, N.STAGE_RECON)
            got = job.download_planes(1)
        ys, xs = np.mgrid[0:32, 0:32]
        assert np.array_equal(got[0], P._expect(base, P.LUMA_F, xs, ys, (16, 16), mv, delta, False)), mv
        cys, cxs = np.mgrid[0:16, 0:16]
        want_c = P._expect(base, P.CHROMA_F, cxs, cys, (8, 8), mv, delta, True)
        assert np.array_equal(got[1], want_c) and np.array_equal(got[2], want_c), mv
