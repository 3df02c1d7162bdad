"""Raw-plane outputs (Decoder::decode_yuv) in the oracle, against the
reference's rules restated here with numpy (CPU).

A single item is plane_window (src/lib.rs:143): the display-size window of
the coded planes at the conformance-window offset, rows and columns past the
plane repeating the last one.  A grid is stitch_native_yuv_band
(src/lib.rs:882): tile (row, col) goes to (col * tile_w, row * tile_h) in
luma and (col * ceil(tile_w / sub_w), row * ceil(tile_h / sub_h)) in chroma,
clipped to the output planes; what no tile covers stays 0.  Samples are the
coded values after SAO (u8 for 8-bit, u16 above); no orientation is applied.
"""
import numpy as np
import pytest

import hpvcd_amd as H
from hpvcd_amd import _native as N
from oracle import oracle as O


def _window(plane, top, left, h, w):
    ys = np.minimum(min(top, plane.shape[0] - 1) + np.arange(h), plane.shape[0] - 1)
    xs = np.minimum(min(left, plane.shape[1] - 1) + np.arange(w), plane.shape[1] - 1)
    return plane[np.ix_(ys, xs)]


def _expected(rs):
    o = rs.outputs()[0]
    W, Hh = o.width, o.height
    y_out = np.zeros((Hh, W), np.uint16)
    p0 = rs.picture(0)
    sw = 2 if p0.chroma_idc == 1 else 1
    sh = sw
    cwo, cho = -(-W // sw), -(-Hh // sh)
    c_out = [np.zeros((cho, cwo), np.uint16) for _ in range(2)]
    for i in range(rs.num_pictures):
        p = rs.picture(i)
        planes = O.picture_planes(rs.batch, i, O.STAGE_SAO)
        col, row = p.dst_x // p.win_w, p.dst_y // p.win_h
        y_out[p.dst_y:p.dst_y + p.copy_h, p.dst_x:p.dst_x + p.copy_w] = _window(
            planes[0], p.crop_top, p.crop_left, p.copy_h, p.copy_w)
        if p.chroma_idc:
            tcw, tch = -(-max(p.win_w, p.copy_w) // sw), -(-max(p.win_h, p.copy_h) // sh)
            dx, dy = col * tcw, row * tch
            cw, ch = min(tcw, cwo - dx), min(tch, cho - dy)
            for k in range(2):
                c_out[k][dy:dy + ch, dx:dx + cw] = _window(planes[1 + k], p.crop_top // sh, p.crop_left // sw, ch, cw)
    return y_out, (c_out if p0.chroma_idc else None)


CASES = {
    "single420": dict(workload="single512", out_kind=N.OUT_YUV420_8, base_seed=601),
    "single_conf_window": dict(workload="single512", tile_width=192, tile_height=136, out_width=168, out_height=128,
                               conf_win=(8, 16, 2, 6), out_kind=N.OUT_YUV420_8, base_seed=603),
    "grid_clip_window": dict(workload="grid4032", tile_width=256, tile_height=128, grid_cols=3, grid_rows=2,
                             conf_win=(4, 12, 2, 6), out_width=700, out_height=230, out_kind=N.OUT_YUV420_8,
                             base_seed=605),
    "rext444_10": dict(workload="rext4096", grid_cols=2, grid_rows=1, tile_width=256, tile_height=128,
                       out_width=500, out_height=128, out_kind=N.OUT_YUV444_16, base_seed=607),
    "mono12": dict(workload="single512", chroma_idc=0, bit_depth=12, out_kind=N.OUT_YUV400_16, base_seed=609),
}


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_raw_planes_follow_the_reference_rules(name):
    rs = H.RecordSet.generate(**CASES[name])
    bufs = rs.output_buffers()
    O.decode_batch(rs.batch, bufs, threads=4)
    y, cb, cr = H.yuv_planes(bufs[0], rs.outputs()[0])
    ey, ec = _expected(rs)
    assert np.array_equal(y, ey)
    if ec is None:
        assert cb is None
    else:
        assert np.array_equal(cb, ec[0]) and np.array_equal(cr, ec[1])
    # the samples are the coded values: 8-bit pictures fit u8
    assert y.dtype == (np.uint8 if rs.picture(0).bit_depth_luma == 8 else np.uint16)


def test_raw_plane_kind_must_match_the_picture():
    """A raw-plane output of another chroma format or sample size is refused."""
    import ctypes as C
    for kind in (N.OUT_YUV444_8, N.OUT_YUV420_16, N.OUT_YUV400_8):
        rs = H.RecordSet.generate("single512", out_kind=kind)
        assert N.lib().hpvcd_validate_batch(C.byref(rs.batch)) == 15, kind
