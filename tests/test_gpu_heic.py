"""HEIC bytes in, image out, on the GPU (hpvcd_decode_heic / _rgb8).

The decoder parses the container and the CABAC syntax on the host and
reconstructs on the GPU; every output sample must equal the oracle's image
of the generator's records for the same configuration (bit-exact), for
both GPU backends.  A grid tile that fails to parse stays black while the
other tiles still match.
"""
import numpy as np
import pytest

import heif_build as B
import hpvcd_amd as H
from hpvcd_amd import N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = {
    "single512": (dict(H.WORKLOADS["single512"], base_seed=201), -1),
    "grid3x2_clip": (dict(H.WORKLOADS["grid4032"], grid_cols=3, grid_rows=2, out_width=1400, out_height=1000,
                          base_seed=203), -1),
    "rext444_10_rgb16": (dict(H.WORKLOADS["rext4096"], grid_cols=2, grid_rows=1, out_width=1024, out_height=512,
                              base_seed=205), -1),
    "rext444_10_rgb8": (dict(H.WORKLOADS["rext4096"], grid_cols=2, grid_rows=1, out_width=1024, out_height=512,
                             base_seed=205), -2),
    "mono_luma": (dict(H.WORKLOADS["single512"], chroma_idc=0, base_seed=207), -1),
    "mono12_rgb8": (dict(H.WORKLOADS["single512"], chroma_idc=0, bit_depth=12, base_seed=209), -2),
    "rot90_grid": (dict(H.WORKLOADS["grid4032"], grid_cols=2, grid_rows=2, out_width=1024, out_height=1000,
                        orientation=N.ORIENT_ROTATE_90, base_seed=211), -1),
    "slices_tiles_pcm_qp": (dict(H.WORKLOADS["single512"], num_slices=3, tile_cols=2, tile_rows=2, pcm=1, p_pcm=0.2,
                                 pcm_bit_depth=7, random_slice_params=1, cu_qp_delta=1, sign_hiding=1,
                                 base_seed=213), -1),
    "wpp_grid2x2": (dict(H.WORKLOADS["grid4032"], grid_cols=2, grid_rows=2, out_width=1024, out_height=1024, wpp=1,
                         persistent_rice=1, base_seed=217), -1),
}

KIND = {-1: {(False, False): N.OUT_RGB8, (False, True): N.OUT_RGB16, (True, False): N.OUT_LUMA8,
             (True, True): N.OUT_LUMA16},
        -2: {(False, False): N.OUT_RGB8, (False, True): N.OUT_RGB8_FROM16, (True, False): N.OUT_RGB8,
             (True, True): N.OUT_RGB8_FROM16}}


def _oracle_image(params, want):
    kind = KIND[want][(params.get("chroma_idc", 1) == 0, params.get("bit_depth", 8) > 8)]
    rs = H.RecordSet.generate(**dict(params, out_kind=kind))
    out = rs.output_buffers()
    O.decode_batch(rs.batch, out, threads=8)
    return out[0]


@pytest.mark.parametrize("backend", [N.BACKEND_GPU_BATCHED, N.BACKEND_GPU_PLAIN])
def test_decode_heic_matches_oracle(backend):
    with H.Decoder(backend=backend) as d:
        for name, (params, want) in CASES.items():
            print(f"decode_heic case {name} backend {backend}", flush=True)
            f = H.gen_heic(**params)
            got = d.decode(f) if want == -1 else d.decode_rgb8(f)
            ref = _oracle_image(params, want)
            assert got.size == ref.size, name
            assert np.array_equal(got.reshape(-1), ref.reshape(-1)), name


def test_damaged_grid_tile_is_black():
    params = dict(H.WORKLOADS["grid4032"], tile_width=128, tile_height=128, grid_cols=3, grid_rows=2, out_width=380,
                  out_height=250, base_seed=215)
    s = H.gen_hevc(**params)
    f = B.grid_file(s, 2, 3, 380, 250, damage=(4,))
    ref = _oracle_image(params, -1)
    with H.Decoder() as d:
        print("decode_heic case damaged tile 4", flush=True)
        got = d.decode(f)
    assert got.shape == ref.shape
    assert not got[128:250, 128:256].any()
    mask = np.ones(got.shape[:2], bool)
    mask[128:250, 128:256] = False
    assert np.array_equal(got[mask], ref[mask])


YUV_CASES = {
    "single512": dict(H.WORKLOADS["single512"], base_seed=221),
    "grid3x2_clip": dict(H.WORKLOADS["grid4032"], grid_cols=3, grid_rows=2, out_width=1400, out_height=1000,
                         base_seed=223),
    "rext444_10": dict(H.WORKLOADS["rext4096"], grid_cols=2, grid_rows=1, out_width=1024, out_height=512,
                       base_seed=225),
    "mono12": dict(H.WORKLOADS["single512"], chroma_idc=0, bit_depth=12, base_seed=227),
    "rot90_grid": dict(H.WORKLOADS["grid4032"], grid_cols=2, grid_rows=2, out_width=1024, out_height=1000,
                       orientation=N.ORIENT_ROTATE_90, base_seed=229),
}


def _yuv_kind(params):
    c, deep = params.get("chroma_idc", 1), params.get("bit_depth", 8) > 8
    base = {1: N.OUT_YUV420_8, 3: N.OUT_YUV444_8, 0: N.OUT_YUV400_8}[c]
    return base + (1 if deep else 0)


def test_decode_heic_yuv_matches_oracle():
    """Raw planes (Decoder::decode_yuv) of HEIC files: every sample equals
    the oracle's raw-plane output of the same records; orientation is
    reported, not applied."""
    with H.Decoder() as d:
        for name, params in YUV_CASES.items():
            print(f"decode_yuv case {name}", flush=True)
            f = H.gen_heic(**params)
            y, cb, cr, info = d.decode_yuv(f)
            rs = H.RecordSet.generate(**dict(params, out_kind=_yuv_kind(params), orientation=0))
            out = rs.output_buffers()
            O.decode_batch(rs.batch, out, threads=8)
            ey, ecb, ecr = H.yuv_planes(out[0], rs.outputs()[0])
            assert np.array_equal(y, ey), name
            if ecb is None:
                assert cb is None and cr is None
            else:
                assert np.array_equal(cb, ecb) and np.array_equal(cr, ecr), name
            assert info["orientation"] == params.get("orientation", 0)
