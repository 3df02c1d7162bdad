"""Allocation-free decode entry point (GPU, through the C ABI).

hpvcd_gpu_decode_batch_noalloc must give the same output images as
hpvcd_gpu_decode_batch, byte for byte, on two calls in a row (the second
with a smaller batch, so stale scratch contents from the first would show),
and refuse a batch larger than its scratch array with HPVCD_E_LIMIT before
any device work.
"""
import ctypes as C

import numpy as np
import pytest

import hpvcd_amd as H
from hpvcd_amd import N

pytestmark = pytest.mark.gpu


def _decode(fn, gpu, rs):
    bufs = rs.output_buffers()
    ptrs = (C.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
    st = fn(gpu._h, C.byref(rs.batch), ptrs)
    return st, bufs


def test_noalloc_matches_decode_batch_on_two_calls(gpu):
    L = N.lib()
    assert L.hpvcd_noalloc_scratch_bytes() == 64 << 20
    # first call: two format classes (8-bit 4:2:0 and 10-bit 4:4:4 with
    # transform skip and cross-component prediction); second: one smaller picture
    first = H.RecordSet.concat([
        H.RecordSet.generate("single512", pcm=1, pcm_bit_depth=7, p_pcm=0.1, num_slices=2, p_sao_merge=0.2),
        H.RecordSet.generate("rext4096", grid_cols=1, grid_rows=1, tile_width=256, tile_height=256,
                             out_width=256, out_height=256, config_id=11),
    ])
    second = H.RecordSet.generate("single512", tile_width=192, tile_height=128, out_width=192, out_height=128,
                                  config_id=12, qp=37)
    for name, rs in (("first", first), ("second", second)):
        print(f"no-alloc case {name}: {rs.num_pictures} pictures", flush=True)
        st0, ref = _decode(L.hpvcd_gpu_decode_batch, gpu, rs)
        assert st0 == 0, L.hpvcd_last_error()
        st1, got = _decode(L.hpvcd_gpu_decode_batch_noalloc, gpu, rs)
        assert st1 == 0, L.hpvcd_last_error()
        for a, b in zip(ref, got):
            assert a.shape == b.shape
            assert np.array_equal(a, b)
        assert any(int(b.max()) > 0 for b in got)


def test_noalloc_refuses_a_batch_beyond_its_scratch(gpu):
    L = N.lib()
    rs = H.RecordSet.generate("grid4032")
    print("no-alloc case: grid4032 (beyond the scratch array)", flush=True)
    st, bufs = _decode(L.hpvcd_gpu_decode_batch_noalloc, gpu, rs)
    assert st == N.E_LIMIT
    assert b"device scratch" in L.hpvcd_last_error()
    assert all(int(b.max()) == 0 for b in bufs)
