"""HEVC stream in, frames out, on the GPU (hpvcd_decode_hevc).

A coded inter sequence (P / B slices, RPS, reference lists, merge / AMVP /
temporal candidates, weighted prediction) is parsed on the host and
reconstructed on the GPU; every frame, in output order, must equal the
oracle's decode of the generator's records for the same configuration
(bit-exact), for both GPU backends.  The parse equals those records byte for
byte (test_video_parse.py), so this closes the chain from stream bytes to
pixels.
"""
import numpy as np
import pytest

import hpvcd_amd as H
from hpvcd_amd import N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BASE = dict(tile_width=256, tile_height=192, out_width=256, out_height=192, grid_cols=1, inter=1, num_refs=3,
            qp=30, p_bi=0.5)

CASES = {
    "gop8_tmvp_weighted": dict(BASE, grid_rows=12, num_outputs=12, gop=8, tmvp=1, weighted=1, base_seed=301),
    "low_delay_p": dict(BASE, grid_rows=6, num_outputs=6, p_bi=0.0, base_seed=303),
    "low_delay_b_poc_wrap": dict(BASE, grid_rows=20, num_outputs=20, num_refs=2, merge_cand=3, base_seed=305),
    "yuv444_10_skip": dict(BASE, grid_rows=6, num_outputs=6, chroma_idc=3, bit_depth=10, out_kind=N.OUT_RGB16,
                           weighted=1, tmvp=1, p_skip=0.6, gop=4, base_seed=307),
    "slices_tiles_wpp_amp": dict(BASE, grid_rows=6, num_outputs=6, num_slices=3, tile_cols=2, tile_rows=2, wpp=1,
                                 tmvp=1, par_merge_level=3, rps_in_sps=1, gop=4, base_seed=309),
    "video1080_gop8": dict(H.WORKLOADS["video1080"], grid_rows=9, num_outputs=9, base_seed=311),
}


def _oracle_frames(params):
    rs = H.RecordSet.generate(**params)
    out = rs.output_buffers()
    O.decode_batch(rs.batch, out, threads=8)
    return out


@pytest.mark.parametrize("backend", [N.BACKEND_GPU_BATCHED, N.BACKEND_GPU_PLAIN])
def test_decode_hevc_matches_oracle(backend):
    with H.Decoder(backend=backend) as d:
        for name, params in CASES.items():
            print(f"decode_hevc case {name} backend {backend}", flush=True)
            bs = H.gen_hevc(**params)
            got = d.decode_hevc(bs, out_kind=params.get("out_kind", N.OUT_RGB8))
            ref = _oracle_frames(params)
            assert len(got) == len(ref) == params["num_outputs"], name
            for k, (a, b) in enumerate(zip(got, ref)):
                assert a.size == b.size, (name, k)
                assert np.array_equal(a.reshape(-1), b.reshape(-1)), (name, k)
