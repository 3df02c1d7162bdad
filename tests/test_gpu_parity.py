"""GPU parity: the HIP back end against the CPU oracle on the same records.

Every Y/Cb/Cr sample after each stage (reconstruction, deblocking, SAO) and
every output RGB sample must be bit-identical (integer pipeline; no
tolerance).  Runs through the C ABI (libhpvcd.so).
"""
import ctypes as C

import numpy as np
import pytest

import handmade as M
import hpvcd_amd as H
from hpvcd_amd import _native as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SMALL = {
    # 512x512 Main 4:2:0 8-bit, QP 32 (the single-image workload)
    "single512": dict(workload="single512"),
    # smooth content (few levels) so deblocking takes strong/weak paths
    "smooth420": dict(workload="single512", p_sig=0.02, p_cbf=0.35, qp=37, base_seed=11),
    # 2x2 grid of 512x512 placed in a 1000x1000 output (right/bottom clipping)
    "grid2x2_clip": dict(workload="grid4032", grid_cols=2, grid_rows=2, out_width=1000, out_height=1000,
                         base_seed=7),
    # 10-bit 4:2:0 (16-bit sample storage, 32-bit residuals), RGB16
    "main10": dict(workload="single512", bit_depth=10, out_kind=N.OUT_RGB16, qp=30, base_seed=3),
    # RExt 4:4:4 10-bit with TS / implicit RDPCM / CCP / rotation, RGB16
    "rext444": dict(workload="rext4096", grid_cols=2, grid_rows=1, out_width=1024, out_height=512,
                    ts_rotation=1, base_seed=5),
    # 4:4:4 8-bit with transform skip up to 8x8 and CCP, RGB8
    "rext444_8bit": dict(workload="rext4096", bit_depth=8, out_kind=N.OUT_RGB8, grid_cols=1, grid_rows=1,
                         out_width=512, out_height=512, log2_max_ts=3, base_seed=9),
    # lossless CUs (transquant bypass): exempt from deblocking and SAO
    "tqb": dict(workload="single512", p_tqb=0.3, p_sig=0.05, base_seed=13),
    # 10-bit picture decoded to RGB8 (>> (bd - 8))
    "main10_rgb8": dict(workload="single512", bit_depth=10, out_kind=N.OUT_RGB8_FROM16, base_seed=17),
    # CTB 32 and 16
    "ctb32": dict(workload="single512", log2_ctb=5, base_seed=19),
    "ctb16": dict(workload="single512", log2_ctb=4, log2_max_tb=4, base_seed=23, tile_width=256, tile_height=128,
                  out_width=256, out_height=128),
    # adaptive colour transform on 4:4:4 10-bit (with TS / RDPCM / CCP)
    "act444": dict(workload="rext4096", grid_cols=1, grid_rows=1, out_width=512, out_height=512, act=1, p_act=0.6,
                   base_seed=41),
    # ACT at 8 bits (i32 residuals where the plain path wraps to i16), lossless CUs, TS up to 8x8
    "act444_8bit": dict(workload="rext4096", bit_depth=8, out_kind=N.OUT_RGB8, grid_cols=1, grid_rows=1,
                        out_width=512, out_height=512, log2_max_ts=3, act=1, p_act=0.6, p_tqb=0.15, base_seed=45),
    # I_PCM CUs, exempt from the loop filters
    "pcm420": dict(workload="single512", pcm=1, p_pcm=0.15, pcm_bit_depth=6, pcm_lf_disabled=1, base_seed=47),
    # I_PCM CUs in 4:4:4 10-bit, loop-filtered
    "pcm444_10": dict(workload="rext4096", grid_cols=1, grid_rows=1, out_width=512, out_height=512, pcm=1,
                      p_pcm=0.15, pcm_bit_depth=7, base_seed=49),
    # inter pictures: 1 I + 4 low-delay B, 4 references, explicit weighted prediction, AMP
    "inter420_wp": dict(workload="single512", tile_width=256, tile_height=128, grid_cols=1, grid_rows=5,
                        num_outputs=5, out_width=256, out_height=128, inter=1, weighted=1, base_seed=61),
    # random access: 1 I + 12 pictures in hierarchical B GOPs of 8 (outputs in display order)
    "inter420_gop8": dict(workload="single512", tile_width=256, tile_height=128, grid_cols=1, grid_rows=13,
                          num_outputs=13, out_width=256, out_height=128, inter=1, weighted=1, gop=8,
                          base_seed=67),
    # P pictures without weights, larger motion
    "inter420_p": dict(workload="single512", tile_width=192, tile_height=136, grid_cols=1, grid_rows=4,
                       num_outputs=4, out_width=192, out_height=136, inter=1, p_bi=0.0, mv_sigma=40.0,
                       base_seed=63),
    # inter 4:4:4 10-bit: TS, explicit RDPCM, CCP and ACT on inter TUs
    "inter444_10": dict(workload="rext4096", tile_width=256, tile_height=128, grid_cols=1, grid_rows=3,
                        num_outputs=3, out_width=256, out_height=128, inter=1, weighted=1, act=1, p_act=0.4,
                        base_seed=65),
    # scaling lists: random lists (8-bit 4:2:0), default lists on 4:4:4 10-bit (32x32 chroma entries),
    # random lists on inter pictures (inter matrices)
    "scaling_random": dict(workload="single512", scaling_list=2, qp=27, base_seed=81),
    "scaling_default444": dict(workload="rext4096", grid_cols=1, grid_rows=1, out_width=512, out_height=512,
                               scaling_list=1, base_seed=83),
    "scaling_inter": dict(workload="single512", tile_width=256, tile_height=128, grid_cols=1, grid_rows=3,
                          num_outputs=3, out_width=256, out_height=128, inter=1, weighted=1, scaling_list=2,
                          base_seed=85),
    # extended precision at 12 bits 4:4:4 with transform skip up to 8x8 and long level tails
    "ep12": dict(workload="rext4096", grid_cols=1, grid_rows=1, out_width=512, out_height=512, bit_depth=12,
                 extended_precision=1, p_level_geom=0.02, log2_max_ts=3, qp=10, base_seed=87),
    # per-CU QP deltas, PPS/slice chroma offsets and slice deblocking offsets
    "qp_delta": dict(workload="single512", cu_qp_delta=1, diff_cu_qp_delta_depth=2, qp_delta_max=8,
                     pps_cb_qp_offset=4, pps_cr_qp_offset=-5, slice_cb_qp_offset=-2, slice_cr_qp_offset=3,
                     beta_offset_div2=3, tc_offset_div2=-2, p_sig=0.05, qp=32, base_seed=89),
    "qp_delta10": dict(workload="single512", bit_depth=10, out_kind=N.OUT_RGB16, cu_qp_delta=1,
                       diff_cu_qp_delta_depth=1, qp_delta_max=12, beta_offset_div2=-4, tc_offset_div2=5,
                       base_seed=91),
    # slices and tiles: gated deblocking and SAO across their boundaries, per-slice parameters
    "slices_tiles": dict(workload="single512", num_slices=5, tile_cols=3, tile_rows=2, random_slice_params=1,
                         lf_across_tiles=0, cu_qp_delta=1, p_sig=0.05, qp=35, base_seed=93),
    "slices_tiles_across": dict(workload="single512", num_slices=3, tile_cols=2, tile_rows=3,
                                random_slice_params=1, lf_across_tiles=1, p_sao_merge=0.3, base_seed=95),
    "slices_no_across": dict(workload="rext4096", grid_cols=1, grid_rows=1, out_width=512, out_height=512,
                             num_slices=4, lf_across_slices=0, tile_cols=2, base_seed=97),
    # colour conversion: BT.601 full range, BT.2020, GBR and YCgCo 4:4:4, 4:0:0 to LUMA8 / LUMA16
    "bt601_full": dict(workload="single512", matrix=5, full_range=1, base_seed=99),
    "bt601_6_limited": dict(workload="single512", matrix=6, base_seed=101),
    "bt2020_10": dict(workload="single512", matrix=9, bit_depth=10, out_kind=N.OUT_RGB16, base_seed=103),
    "gbr444_full": dict(workload="rext4096", grid_cols=1, grid_rows=1, out_width=512, out_height=512, matrix=0,
                        full_range=1, base_seed=105),
    "gbr444_limited_rgb8": dict(workload="rext4096", grid_cols=1, grid_rows=1, out_width=512, out_height=512,
                                matrix=0, out_kind=N.OUT_RGB8_FROM16, base_seed=107),
    "ycgco444": dict(workload="rext4096", grid_cols=1, grid_rows=1, out_width=512, out_height=512, matrix=8,
                     bit_depth=8, out_kind=N.OUT_RGB8, base_seed=109),
    "mono8": dict(workload="single512", chroma_idc=0, out_kind=N.OUT_LUMA8, base_seed=111),
    "mono12": dict(workload="single512", chroma_idc=0, bit_depth=12, out_kind=N.OUT_LUMA16, base_seed=113),
    # conformance window crop of grid tiles, and oriented outputs
    "crop_grid": dict(workload="grid4032", grid_cols=3, grid_rows=2, tile_width=256, tile_height=128,
                      out_width=3 * 232, out_height=2 * 112, conf_win=(8, 16, 4, 12), base_seed=115),
    "orient_rot90": dict(workload="grid4032", grid_cols=2, grid_rows=2, tile_width=128, tile_height=64,
                         out_width=240, out_height=120, orientation=6, base_seed=117),
    "orient_rot270_mono": dict(workload="grid4032", grid_cols=2, grid_rows=1, tile_width=128, tile_height=64,
                               out_width=250, out_height=64, orientation=8, chroma_idc=0, out_kind=N.OUT_LUMA8,
                               base_seed=119),
    "orient_flips16": dict(workload="rext4096", grid_cols=2, grid_rows=1, out_width=1000, out_height=500,
                           orientation=3, conf_win=(0, 8, 0, 12), base_seed=121),
}

# every other orientation on one small grid (RGB only, against the oracle)
ORIENTATIONS = [2, 4, 5, 7]


def _first_diff(a, b):
    idx = np.argwhere(a != b)
    return tuple(idx[0]) if len(idx) else None


def _referenced(rs):
    refs = set()
    for p in range(rs.num_pictures):
        P = rs.picture(p)
        for k in range(P.ref_count):
            r = rs.batch.refs[P.ref_begin + k].picture
            if r != N.REF_MISSING:
                refs.add(r)
    return refs


def _check_planes(job, rs, stage):
    # a picture that later pictures predict from always completes SAO
    refd = _referenced(rs)
    for p in range(rs.num_pictures):
        want = O.picture_planes(rs.batch, p, N.STAGE_SAO if p in refd else stage)
        got = job.download_planes(p)
        for name, w, g in zip("YUV", want, got):
            if w is None:
                continue
            assert np.array_equal(w, g), (
                f"picture {p} plane {name} stage {stage}: {int(np.count_nonzero(w != g))} samples differ, "
                f"first at {_first_diff(w, g)}")


@pytest.mark.parametrize("case", sorted(SMALL))
def test_stages_bit_exact(gpu, case):
    rs = H.RecordSet.generate(**SMALL[case])
    print(f"case {case}: {rs.num_pictures} pictures, {rs.batch.num_tbs} TBs, {rs.batch.num_pus} PUs", flush=True)
    with gpu.job(rs) as job:
        for stage in (N.STAGE_RECON, N.STAGE_DEBLOCK, N.STAGE_SAO):
            job.run(N.MODE_BATCHED, stage)
            _check_planes(job, rs, stage)
        job.run(N.MODE_BATCHED, N.STAGE_RGB)
        got = job.download()
    want = rs.output_buffers()
    O.decode_batch(rs.batch, want, threads=8)
    for g, w in zip(got, want):
        assert np.array_equal(g, w), f"RGB differs in {int(np.count_nonzero(g != w))} samples"


@pytest.mark.parametrize("mode", [N.MODE_PLAIN, N.MODE_GRAPH])
def test_modes_identical(gpu, mode):
    rs = H.RecordSet.generate("grid4032", grid_cols=3, grid_rows=2, out_width=1400, out_height=1000, base_seed=29)
    with gpu.job(rs) as job:
        job.run(N.MODE_BATCHED)
        ref = job.download()
        for _ in range(2):
            job.run(mode)
            got = job.download()
            for a, b in zip(ref, got):
                assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", [N.MODE_PLAIN, N.MODE_GRAPH])
def test_modes_act_pcm(gpu, mode):
    """Per-picture ACT lists (plain mode) and the replayed graph."""
    rs = H.RecordSet.generate("rext4096", grid_cols=2, grid_rows=2, out_width=1024, out_height=1024, act=1,
                              p_act=0.6, pcm=1, p_pcm=0.1, pcm_bit_depth=8, pcm_lf_disabled=1, base_seed=53)
    want = rs.output_buffers()
    O.decode_batch(rs.batch, want, threads=8)
    with gpu.job(rs) as job:
        for _ in range(2):
            job.run(mode)
            for g, w in zip(job.download(), want):
                assert np.array_equal(g, w), f"{int(np.count_nonzero(g != w))} samples differ"


@pytest.mark.parametrize("n,scale_cb,scale_cr", [(16, 0, 0), (8, 2, -1), (4, -8, 4), (32, 1, 0)])
def test_act_lossless_closed_form(gpu, n, scale_cb, scale_cr):
    rs, r = M.act_lossless_case(seed=n, n=n, scale_cb=scale_cb, scale_cr=scale_cr)
    with gpu.job(rs) as job:
        job.run(N.MODE_BATCHED, N.STAGE_RECON)
        got = job.download_planes(0)
    for g, t, w in zip(got, r, O.picture_planes(rs.batch, 0, N.STAGE_RECON)):
        assert np.array_equal(g[:n, :n], 128 + t) and np.array_equal(g, w)


@pytest.mark.parametrize("chroma_idc,log2_cu,bd,bits", [(1, 3, 8, 5), (1, 5, 8, 8), (3, 4, 10, 7), (0, 4, 12, 1)])
def test_pcm_closed_form(gpu, chroma_idc, log2_cu, bd, bits):
    rs, want = M.pcm_case(seed=log2_cu, chroma_idc=chroma_idc, bd=bd, pcm_bits=bits, log2_cu=log2_cu, size=64)
    with gpu.job(rs) as job:
        job.run(N.MODE_BATCHED, N.STAGE_RECON)
        got = job.download_planes(0)
    for g, w in zip(got, want):
        if w is not None:
            assert np.array_equal(g, w)


def _mc_case_planes(gpu, rs):
    with gpu.job(rs) as job:
        job.run(N.MODE_BATCHED, N.STAGE_RECON)
        return job.download_planes(1)


FULL_MVS = [(8, 16), (-24, 8), (40, -16), (-200, 96), (0, 0), (336, -400), (16, 8), (-8, -8)]


@pytest.mark.parametrize("bi", [False, True])
def test_mc_full_sample_copy(gpu, bi):
    import test_oracle_kat as K

    rs = M.inter_case(K.MC_Y, K.MC_CB, K.MC_CR, mvs=FULL_MVS, bi=bi)
    for g, w in zip(_mc_case_planes(gpu, rs), K.mc_expected_full(K.MC_Y, K.MC_CB, K.MC_CR, FULL_MVS)):
        assert np.array_equal(g, w)


def test_mc_fractional_against_oracle(gpu):
    """Random fractional vectors (uni, bi, weighted) on random references, 8x8 PUs."""
    rng = np.random.default_rng(5)
    y, cb, cr = rng.integers(0, 256, (32, 64)), rng.integers(0, 256, (16, 32)), rng.integers(0, 256, (16, 32))
    mvs = [(int(a), int(b)) for a, b in rng.integers(-300, 300, (32, 2))]
    wts = (((40, 9), (90, -20), (64, 100)), ((80, -4), (33, 7), (64, 30)))
    for bi in (False, True):
        for weights in (None, wts):
            rs = M.inter_case(y, cb, cr, mvs=mvs, bi=bi, weights=weights, pu=8)
            got = _mc_case_planes(gpu, rs)
            want = O.picture_planes(rs.batch, 1, N.STAGE_RECON)
            for g, w in zip(got, want):
                assert np.array_equal(g, w), f"bi={bi} weighted={weights is not None}"


def test_inter_missing_references(gpu):
    """Entries without a picture: a bi PU falls back to its other list, a PU
    with no usable list predicts mid-grey (src/decode.rs:7206)."""
    rs = H.RecordSet.generate(**dict(SMALL["inter420_wp"], grid_rows=3, num_outputs=3))
    refs = (N.Ref * rs.batch.num_refs).from_address(C.addressof(rs.batch.refs.contents))
    for k in range(0, rs.batch.num_refs, 3):
        refs[k].picture = N.REF_MISSING
    assert N.lib().hpvcd_validate_batch(C.byref(rs.batch)) == 0
    want = rs.output_buffers()
    O.decode_batch(rs.batch, want, threads=4)
    got = gpu.decode(rs)
    for g, w in zip(got, want):
        assert np.array_equal(g, w), f"{int(np.count_nonzero(g != w))} samples differ"


@pytest.mark.parametrize("mode", [N.MODE_PLAIN, N.MODE_GRAPH])
def test_modes_inter(gpu, mode):
    rs = H.RecordSet.generate(**dict(SMALL["inter420_wp"], base_seed=67))
    want = rs.output_buffers()
    O.decode_batch(rs.batch, want, threads=8)
    with gpu.job(rs) as job:
        for _ in range(2):
            job.run(mode)
            for g, w in zip(job.download(), want):
                assert np.array_equal(g, w), f"{int(np.count_nonzero(g != w))} samples differ"


def test_mixed_formats_one_batch(gpu):
    """8-bit 4:2:0 and 10-bit 4:4:4 pictures in one batch (two format classes)."""
    a = H.RecordSet.generate("single512", base_seed=31)
    b = H.RecordSet.generate("rext4096", grid_cols=1, grid_rows=1, out_width=512, out_height=512, base_seed=37)
    rs = H.RecordSet.concat([a, b])
    got = gpu.decode(rs)
    want = rs.output_buffers()
    O.decode_batch(rs.batch, want, threads=2)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


@pytest.mark.parametrize("workload", ["grid4032", "rext4096", "video1080"])
def test_full_size_workloads(gpu, workload):
    """The full benchmark workloads, every output sample against the oracle."""
    rs = H.RecordSet.generate(workload)
    with gpu.job(rs) as job:
        job.run(N.MODE_GRAPH)
        got = job.download()
    want = rs.output_buffers()
    O.decode_batch(rs.batch, want, threads=8)
    for g, w in zip(got, want):
        assert np.array_equal(g, w), f"{int(np.count_nonzero(g != w))} samples differ"


@pytest.mark.parametrize("name", ["single420", "single_conf_window", "grid_clip_window", "rext444_10", "mono12"])
def test_raw_plane_outputs(gpu, name):
    """Raw-plane outputs (decode_yuv) of record sets, GPU against oracle."""
    from test_yuv_output import CASES as YUV
    rs = H.RecordSet.generate(**YUV[name])
    got = gpu.decode(rs)
    want = rs.output_buffers()
    O.decode_batch(rs.batch, want, threads=4)
    for g, w in zip(got, want):
        assert np.array_equal(g.reshape(-1), w.reshape(-1)), name


@pytest.mark.parametrize("orientation", ORIENTATIONS)
def test_orientations(gpu, orientation):
    rs = H.RecordSet.generate("grid4032", grid_cols=2, grid_rows=2, tile_width=128, tile_height=64,
                              out_width=230, out_height=100, orientation=orientation, base_seed=123)
    got = gpu.decode(rs)
    want = rs.output_buffers()
    O.decode_batch(rs.batch, want, threads=4)
    for g, w in zip(got, want):
        assert np.array_equal(g, w), f"orientation {orientation}: {int(np.count_nonzero(g != w))} samples differ"


@pytest.mark.parametrize("bd", [8, 10, 12])
@pytest.mark.parametrize("ep", [False, True])
@pytest.mark.parametrize("stress", [False, True])
def test_dequant_stress_levels(gpu, bd, ep, stress):
    """The reference's SAFE/STRESS level tables (src/transform.rs:1374-1382)
    through K1 at every position of 4..32 TBs, with and without extended
    precision, transform skip on the 4x4 TBs in the second half."""
    for ts in (False, True):
        rs = M.stress_case(bd=bd, ep=ep, stress=stress, ts=ts, seed=bd)
        print(f"stress bd {bd} ep {ep} stress {stress} ts {ts}", flush=True)
        with gpu.job(rs) as job:
            job.run(N.MODE_BATCHED, N.STAGE_RECON)
            got = job.download_planes(0)
        want = O.picture_planes(rs.batch, 0, N.STAGE_RECON)
        for c, (g, w) in enumerate(zip(got, want)):
            if w is not None:
                assert np.array_equal(g, w), f"plane {c}: {int(np.count_nonzero(g != w))} samples differ"


def test_mixed_intra_and_inter_waves_stop_at_recon(gpu):
    """An all-intra picture and an inter sequence's I picture share wave 0;
    at STAGE_RECON the unreferenced intra picture returns reconstructed (not
    deblocked) samples while the referenced I picture completes SAO."""
    a = H.RecordSet.generate("single512", tile_width=256, tile_height=128, out_width=256, out_height=128,
                             base_seed=125)
    b = H.RecordSet.generate(**dict(SMALL["inter420_wp"], grid_rows=3, num_outputs=3, base_seed=127))
    rs = H.RecordSet.concat([a, b])
    with gpu.job(rs) as job:
        job.run(N.MODE_BATCHED, N.STAGE_RECON)
        _check_planes(job, rs, N.STAGE_RECON)
        job.run(N.MODE_GRAPH, N.STAGE_DEBLOCK)
        _check_planes(job, rs, N.STAGE_DEBLOCK)


@pytest.mark.parametrize("backend", [N.BACKEND_GPU_PLAIN, N.BACKEND_GPU_BATCHED])
def test_decoder_records_every_backend(backend):
    """hpvcd_decode_records through the drop-in decoder object: library-owned
    images equal the oracle's outputs for both GPU backends, oriented
    outputs report the swapped size, 4:0:0 decodes to gray RGB8."""
    cases = [H.RecordSet.generate("grid4032", grid_cols=2, grid_rows=2, out_width=1000, out_height=1000,
                                  base_seed=131),
             H.RecordSet.generate(**dict(SMALL["orient_rot90"])),
             H.RecordSet.generate("single512", chroma_idc=0, out_kind=N.OUT_RGB8, base_seed=133),
             H.RecordSet.generate(**dict(SMALL["inter420_wp"]))]
    with H.Decoder(backend=backend) as d:
        for rs in cases:
            got = d.decode_records(rs)
            want = rs.output_buffers()
            O.decode_batch(rs.batch, want, threads=4)
            for g, w, o in zip(got, want, rs.outputs()):
                if o.orientation in (5, 6, 7, 8):
                    assert g.shape[:2] == (o.width, o.height)
                assert np.array_equal(g.reshape(-1), w.reshape(-1))


def test_decoder_limits():
    rs = H.RecordSet.generate("single512", base_seed=135)
    with H.Decoder(max_dimension=256) as d:
        with pytest.raises(H.HpvcdError) as e:
            d.decode_records(rs)
        assert e.value.status == 12  # E_LIMIT
