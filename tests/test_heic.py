"""HEIF container (CPU): HEIC file -> records, as the reference reads it.

Parity bar: hpvcd_parse_heic of the generator's HEIC file gives the records
hpvcd_gen_records makes for the same configuration, byte for byte (pictures
with their grid placement, TBs, levels, cells, SAO, scaling sets, outputs).
The container rules are the reference's (src/heif.rs:242-1160,
src/lib.rs:1497-1870): the tests below pin each one with a file built for
it (tests/heif_build.py), mirroring the reference's own container tests
(src/heif.rs tests: non-HEIF input, box and item size limits) and its
documented behaviour (last iloc extent kept, irot/imir mapping and the
dropped second transform, VUI colour over colr over sRGB for single
items, tile colour from the first tile's SPS for grids, failed tiles left
black, tile placement by the filtered tile list).
"""
import pytest

import heif_build as B
import hpvcd_amd as H
from hpvcd_amd import N

ARRAYS = ("pictures", "tbs", "coeffs", "sao", "cell_info", "scaling", "pus", "refs", "outputs")

CASES = {
    "single512": dict(workload="single512"),
    "grid3x2_clip": dict(workload="grid4032", grid_cols=3, grid_rows=2, out_width=1400, out_height=1000),
    "rext444_10": dict(workload="rext4096", grid_cols=2, grid_rows=1, out_width=1024, out_height=512),
    "mono_luma": dict(workload="single512", chroma_idc=0, out_kind=N.OUT_LUMA8),
    "rot90_grid": dict(workload="grid4032", grid_cols=2, grid_rows=2, out_width=1024, out_height=1000,
                       orientation=N.ORIENT_ROTATE_90),
    "flip_v": dict(workload="single512", orientation=N.ORIENT_FLIP_V),
    "crop_window": dict(workload="single512", tile_width=384, tile_height=264, out_width=360, out_height=240,
                        conf_win=(8, 16, 2, 6)),
    "window_single": dict(workload="single512", tile_width=384, tile_height=264, out_width=360, out_height=256,
                          conf_win=(8, 16, 2, 6)),
    "slices_tiles_pcm": dict(workload="single512", num_slices=3, tile_cols=2, tile_rows=2, pcm=1, p_pcm=0.2,
                             pcm_bit_depth=7, random_slice_params=1),
    "full_range_601": dict(workload="grid4032", grid_cols=2, grid_rows=1, out_width=1024, out_height=512,
                           matrix=6, full_range=1),
}


def _params(name):
    p = dict(CASES[name])
    wl = p.pop("workload")
    return dict(H.WORKLOADS[wl], **p)


@pytest.mark.parametrize("name", sorted(CASES))
def test_parse_of_generated_heic_equals_records(name):
    params = _params(name)
    f = H.gen_heic(**params)
    g = H.RecordSet.generate(**params)
    p = H.RecordSet.parse_heic(f, out_kind=H.gen_config(**params).out_kind)
    for a in ARRAYS:
        assert g.raw(a) == p.raw(a), a


def test_default_kinds_follow_decode_heic():
    """-1: luma for 4:0:0, RGB otherwise, 16-bit above 8-bit depth; -2: RGB8
    (>> (bd - 8) for deep pictures) as decode_rgb8 (src/lib.rs:1880-1990)."""
    cases = [(dict(H.WORKLOADS["single512"]), N.OUT_RGB8, N.OUT_RGB8),
             (dict(H.WORKLOADS["single512"], bit_depth=10), N.OUT_RGB16, N.OUT_RGB8_FROM16),
             (dict(H.WORKLOADS["single512"], chroma_idc=0), N.OUT_LUMA8, N.OUT_RGB8),
             (dict(H.WORKLOADS["single512"], chroma_idc=0, bit_depth=12), N.OUT_LUMA16, N.OUT_RGB8_FROM16),
             (dict(H.WORKLOADS["rext4096"], grid_cols=2, grid_rows=1, out_width=1024, out_height=512),
              N.OUT_RGB16, N.OUT_RGB8_FROM16)]
    for params, k_default, k_rgb8 in cases:
        f = H.gen_heic(**params)
        assert H.RecordSet.parse_heic(f, -1).outputs()[0].kind == k_default
        assert H.RecordSet.parse_heic(f, -2).outputs()[0].kind == k_rgb8


SMALL = dict(H.WORKLOADS["single512"], tile_width=128, tile_height=128, out_width=128, out_height=128)


def _stream(**kw):
    return H.gen_hevc(**dict(SMALL, **kw))


def _status(data, **settings):
    s = None
    if settings:
        s = N.Settings()
        N.lib().hpvcd_settings_default(H.C.byref(s))
        for k, v in settings.items():
            setattr(s, k, v)
    try:
        H.RecordSet.parse_heic(data, settings=s)
        return 0
    except H.HpvcdError as e:
        return e.status


def test_non_heif_and_missing_boxes():
    assert _status(B.box(b"ftyp", b"mp42")) == N.E_NOT_HEIF
    assert _status(b"\0\0\0\x10junkjunkjunk") == N.E_NOT_HEIF
    f = B.Heif()
    f.with_meta = False
    assert _status(f.build()) == N.E_MISSING_BOX
    # an item without hvcC, an item without an iloc entry
    vps, sps, pps, pics = B.split_annexb(_stream())
    f = B.Heif()
    f.item(1, b"hvc1", B.sample(pics[0]), [f.prop(b"ispe", B.ispe(128, 128))])
    assert _status(f.build()) == N.E_MISSING_BOX
    f = B.Heif()
    f.item(1, b"hvc1", B.sample(pics[0]), [f.prop(b"hvcC", B.hvcc(vps, sps, pps)), f.prop(b"ispe", B.ispe(128, 128))],
           no_iloc=True)
    assert _status(f.build()) == N.E_MISSING_BOX
    # hvcC without a PPS array
    assert _status(B.single_item_file(_stream(), (128, 128), hvcc_arrays=(32, 33))) == N.E_PARAMSET


def test_box_and_item_size_limits():
    good = B.single_item_file(_stream(), (128, 128))
    assert _status(good) == 0
    # the meta box is larger than the cap: iteration stops before it (src/heif.rs Boxes::with_limit)
    assert _status(good, max_box_size=64) == N.E_MISSING_BOX
    assert _status(good, max_item_size=100) == N.E_LIMIT
    assert _status(B.single_item_file(_stream(), (0, 128))) == N.E_BAD_DIMENSIONS
    assert _status(good, max_dimension=64) == N.E_LIMIT
    assert _status(good, max_pixels=128 * 127) == N.E_LIMIT


@pytest.mark.parametrize("version,cm,bogus", [(0, 0, False), (1, 0, True), (1, 1, False), (2, 0, False),
                                              (2, 1, True)])
def test_iloc_versions_and_last_extent(version, cm, bogus):
    s = _stream()
    ref = H.RecordSet.parse_hevc(s)
    f = B.single_item_file(s, (128, 128), iloc_version=version, cm=cm, bogus_first_extent=bogus)
    p = H.RecordSet.parse_heic(f)
    for a in ("tbs", "coeffs", "sao", "cell_info"):
        assert p.raw(a) == ref.raw(a)


@pytest.mark.parametrize("length_size", [2, 3, 4])
def test_nal_length_sizes(length_size):
    s = _stream()
    ref = H.RecordSet.parse_hevc(s)
    p = H.RecordSet.parse_heic(B.single_item_file(s, (128, 128), length_size=length_size))
    assert p.raw("coeffs") == ref.raw("coeffs")


def test_one_byte_nal_lengths():
    s = H.gen_hevc(**dict(SMALL, tile_width=32, tile_height=32, out_width=32, out_height=32, qp=51, p_cbf=0.02,
                          sao=0))
    _, _, _, pics = B.split_annexb(s)
    assert all(len(n) < 256 for n in pics[0])
    p = H.RecordSet.parse_heic(B.single_item_file(s, (32, 32), length_size=1))
    assert p.raw("tbs") == H.RecordSet.parse_hevc(s).raw("tbs")


@pytest.mark.parametrize("props,want", [
    ([(b"irot", b"\x01")], N.ORIENT_ROTATE_270),
    ([(b"irot", b"\x02")], N.ORIENT_ROTATE_180),
    ([(b"irot", b"\x03")], N.ORIENT_ROTATE_90),
    ([(b"irot", b"\x00")], 0),
    ([(b"imir", b"\x00")], N.ORIENT_FLIP_H),
    ([(b"imir", b"\x01")], N.ORIENT_FLIP_V),
    # a transform after a first non-identity one is dropped (src/heif.rs:1140-1160)
    ([(b"irot", b"\x01"), (b"imir", b"\x00")], N.ORIENT_ROTATE_270),
    ([(b"imir", b"\x01"), (b"irot", b"\x03")], N.ORIENT_FLIP_V),
    ([(b"irot", b"\x00"), (b"imir", b"\x01")], N.ORIENT_FLIP_V),
])
def test_irot_imir_orientation(props, want):
    p = H.RecordSet.parse_heic(B.single_item_file(_stream(), (128, 128), extra_props=props))
    assert p.outputs()[0].orientation == want


def test_single_item_colour_priority():
    """VUI colour unless its matrix is unspecified, then colr nclx, then sRGB
    (full-range BT.709) (src/lib.rs:1536-1546)."""
    vui = _stream(matrix=6, full_range=0)
    p = H.RecordSet.parse_heic(B.single_item_file(vui, (128, 128), colr=B.nclx(9, True)))
    assert p.picture(0).matrix == 6 and not p.picture(0).flags & N.PF_FULL_RANGE
    unspec = _stream(matrix=2, full_range=0)
    p = H.RecordSet.parse_heic(B.single_item_file(unspec, (128, 128), colr=B.nclx(9, True)))
    assert p.picture(0).matrix == 9 and p.picture(0).flags & N.PF_FULL_RANGE
    p = H.RecordSet.parse_heic(B.single_item_file(unspec, (128, 128)))
    assert p.picture(0).matrix == 1 and p.picture(0).flags & N.PF_FULL_RANGE
    # a reserved matrix code reads as unspecified
    p = H.RecordSet.parse_heic(B.single_item_file(_stream(matrix=3), (128, 128), colr=B.nclx(5, False)))
    assert p.picture(0).matrix == 5


def test_grid_colour_from_first_tile_sps():
    """Grids take the first tile's VUI colour, unspecified included, and
    ignore the grid item's colr (src/lib.rs:1801-1817)."""
    s = H.gen_hevc(**dict(SMALL, grid_cols=2, grid_rows=1, matrix=2, full_range=1))
    p = H.RecordSet.parse_heic(B.grid_file(s, 1, 2, 256, 128, grid_props=[(b"colr", B.nclx(9, False))]))
    assert p.num_pictures == 2
    for i in range(2):
        assert p.picture(i).matrix == 2 and p.picture(i).flags & N.PF_FULL_RANGE


def test_grid_placement_and_failed_tiles():
    s = H.gen_hevc(**dict(SMALL, grid_cols=3, grid_rows=2, base_seed=5))
    whole = H.RecordSet.parse_heic(B.grid_file(s, 2, 3, 380, 250))
    assert [(whole.picture(i).dst_x, whole.picture(i).dst_y, whole.picture(i).copy_w, whole.picture(i).copy_h)
            for i in range(6)] == [(0, 0, 128, 128), (128, 0, 128, 128), (256, 0, 124, 128),
                                   (0, 128, 128, 122), (128, 128, 128, 122), (256, 128, 124, 122)]
    # a damaged tile is left out (black in the image); the others keep their places
    dmg = H.RecordSet.parse_heic(B.grid_file(s, 2, 3, 380, 250, damage=(4,)))
    assert dmg.num_pictures == 5
    assert [(dmg.picture(i).dst_x, dmg.picture(i).dst_y) for i in range(5)] == [
        (0, 0), (128, 0), (256, 0), (0, 128), (256, 128)]
    # a tile without an iloc entry drops out of the tile list, and the tiles
    # after it move up one place (the reference's filtered list)
    drop = H.RecordSet.parse_heic(B.grid_file(s, 2, 3, 380, 250, drop_iloc=(1,)))
    assert drop.num_pictures == 5
    assert [(drop.picture(i).dst_x, drop.picture(i).dst_y) for i in range(5)] == [
        (0, 0), (128, 0), (256, 0), (0, 128), (128, 128)]
    assert drop.raw("tbs")[:len(whole.raw("tbs")) // 6] == whole.raw("tbs")[:len(whole.raw("tbs")) // 6]


def test_grid_descriptor_32bit_sizes_and_zero_size():
    s = H.gen_hevc(**dict(SMALL, grid_cols=2, grid_rows=1))
    p = H.RecordSet.parse_heic(B.grid_file(s, 1, 2, 250, 100, wide=True))
    assert (p.outputs()[0].width, p.outputs()[0].height) == (250, 100)
    assert _status(B.grid_file(s, 1, 2, 0, 100)) == N.E_BAD_DIMENSIONS


def test_heic_writer_refuses_transpose():
    with pytest.raises(H.HpvcdError) as e:
        H.gen_heic(**dict(SMALL, orientation=N.ORIENT_TRANSPOSE))
    assert e.value.status == N.E_UNSUPPORTED
