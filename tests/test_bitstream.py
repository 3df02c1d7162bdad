"""Coded HEVC of the record producer (CPU): writer -> parser -> records.

The generator's decisions are CABAC-coded into an Annex B stream
(hpvcd_gen_hevc); the C++ parser (hpvcd_parse_hevc) reads it back through the
same coding-tree walk the generator uses.  Parity bar: the parsed records
equal the generated records byte for byte (every TB, level, cell word, SAO
set, scaling set, and for one-picture configurations the picture and output
records too).  The CABAC tables themselves are pinned against the
reference's in test_cabac_engine.py; the syntax the parser reads follows
src/decode.rs:1319-3560 and src/cabac/residual_full.rs:392 (the reference
ships no bitstream fixtures, so the ordering of syntax elements is pinned by
this round trip and by those citations).

Malformed streams (truncation, flipped bytes, a missing PPS) return a status
and never crash.
"""
import ctypes as C
import random

import pytest

import hpvcd_amd as H
from hpvcd_amd import N

SMALL = dict(tile_width=256, tile_height=256, out_width=256, out_height=256, grid_cols=1, grid_rows=1)

CASES = {
    "main8": {},
    "ctb16": dict(log2_ctb=4, log2_min_cb=3, log2_max_tb=4, max_transform_depth_intra=2),
    "ctb32_cb16_pcm": dict(log2_ctb=5, log2_min_cb=4, pcm=1, pcm_bit_depth=8, p_pcm=0.3),
    "pcm7_lf_off": dict(pcm=1, pcm_bit_depth=7, p_pcm=0.25, pcm_lf_disabled=1),
    "tqb_ts_rdpcm": dict(p_tqb=0.3, transform_skip=1, implicit_rdpcm=1, ts_rotation=1),
    "ts32": dict(transform_skip=1, log2_max_ts=5, p_ts=0.5),
    "qp_delta": dict(cu_qp_delta=1, diff_cu_qp_delta_depth=2, qp_delta_max=25),
    "qp_delta10": dict(cu_qp_delta=1, diff_cu_qp_delta_depth=1, qp_delta_max=12, bit_depth=10, out_kind=N.OUT_RGB16),
    "tiles_slices": dict(tile_cols=2, tile_rows=2, num_slices=3, p_sao_merge=0.3, random_slice_params=1),
    "slices_no_across": dict(num_slices=4, lf_across_slices=0),
    "tiles_no_across": dict(tile_cols=4, tile_rows=1, lf_across_tiles=0),
    "scaling_random": dict(scaling_list=2),
    "scaling_default": dict(scaling_list=1),
    "sign_hiding": dict(sign_hiding=1),
    "rext_entropy": dict(sign_hiding=1, persistent_rice=1, bypass_alignment=1, transform_skip=1, ts_context=1,
                         p_ts=0.5),
    "ep12": dict(bit_depth=12, extended_precision=1, p_level_geom=0.05, out_kind=N.OUT_RGB16),
    "deep10_big_levels": dict(bit_depth=10, p_level_geom=0.02, out_kind=N.OUT_RGB16),
    "mono8": dict(chroma_idc=0, out_kind=N.OUT_LUMA8),
    "mono12": dict(chroma_idc=0, bit_depth=12, out_kind=N.OUT_LUMA16),
    "yuv444_act_ccp": dict(chroma_idc=3, act=1, p_act=0.5, cross_component=1),
    "yuv444_act_qp10": dict(chroma_idc=3, act=1, p_act=0.5, cross_component=1, cu_qp_delta=1, bit_depth=10,
                            out_kind=N.OUT_RGB16),
    "conf_window": dict(tile_width=192, tile_height=136, out_width=168, out_height=128, conf_win=(8, 16, 2, 6)),
    "chroma_offsets": dict(pps_cb_qp_offset=3, pps_cr_qp_offset=-4, slice_cb_qp_offset=-2, slice_cr_qp_offset=5,
                           beta_offset_div2=3, tc_offset_div2=-2),
    "no_loop_filters": dict(deblock=0, sao=0),
    "random_slices_no_filters": dict(deblock=0, sao=0, random_slice_params=1, num_slices=4),
    "qp51": dict(qp=51),
    "qp0": dict(qp=0, p_level_geom=0.01),
    "full_range_601": dict(full_range=1, matrix=6),
    # WPP: a substream per CTB row, contexts (and Rice statistics) synced
    # from the state after the second CTB of the row above
    # (src/decode.rs:1230-1275)
    "wpp": dict(wpp=1),
    "wpp_tiles": dict(wpp=1, tile_cols=3, tile_rows=2),
    "wpp_slices_mid_row": dict(wpp=1, num_slices=3),
    "wpp_slices_tiles": dict(wpp=1, num_slices=5, tile_cols=2, tile_rows=2, random_slice_params=1),
    "wpp_rice_pcm_qp": dict(wpp=1, persistent_rice=1, pcm=1, pcm_bit_depth=7, p_pcm=0.2, cu_qp_delta=1,
                            diff_cu_qp_delta_depth=1, qp_delta_max=10),
    "wpp_one_ctb_tiles": dict(wpp=1, tile_cols=4, tile_rows=1),
}

ARRAYS = ("tbs", "coeffs", "sao", "cell_info", "scaling", "pus", "refs")


def _roundtrip(params):
    bs = H.gen_hevc(**params)
    g = H.RecordSet.generate(**params)
    p = H.RecordSet.parse_hevc(bs, out_kind=H.gen_config(**params).out_kind)
    return bs, g, p


@pytest.mark.parametrize("name", sorted(CASES))
def test_parse_of_generated_stream_equals_records(name):
    params = dict(SMALL, **CASES[name])
    _, g, p = _roundtrip(params)
    assert g.num_pictures == p.num_pictures == 1
    for a in ARRAYS:
        assert g.raw(a) == p.raw(a), a
    assert g.raw("pictures") == p.raw("pictures")
    assert g.raw("outputs") == p.raw("outputs")
    assert N.lib().hpvcd_validate_batch(C.byref(p.batch)) == 0


def test_wpp_entry_points_over_escaped_slice_data():
    """Substreams start at their entry points, which count the emulation
    prevention bytes of the NAL unit (7.3.8.1, src/decode.rs:1184-1196): a
    WPP picture whose slice data holds such a byte parses to its records."""
    import heif_build as B

    params = dict(H.WORKLOADS["single512"], tile_width=1024, tile_height=512, out_width=1024, out_height=512,
                  wpp=1, qp=22, p_sig=0.5, base_seed=26)
    bs, g, p = _roundtrip(params)
    _, _, _, pics = B.split_annexb(bs)
    assert sum(n.count(b"\x00\x00\x03") for n in pics[0]) >= 1
    for a in ARRAYS:
        assert g.raw(a) == p.raw(a), a


def test_grid_pictures_parse_to_the_generated_tiles():
    params = dict(H.WORKLOADS["grid4032"], grid_cols=3, grid_rows=2, out_width=1536, out_height=1024)
    _, g, p = _roundtrip(params)
    assert p.num_pictures == 6
    for a in ARRAYS:
        assert g.raw(a) == p.raw(a), a
    # the coding fields of every picture record match (placement differs:
    # the stream has no grid, so each picture is its own output)
    for i in range(6):
        a, b = g.picture(i), p.picture(i)
        for f, _ in N.Picture._fields_:
            if f in ("dst_x", "dst_y", "copy_w", "copy_h", "out_index"):
                continue
            va, vb = getattr(a, f), getattr(b, f)
            if f == "slices":
                va, vb = bytes(va), bytes(vb)
            assert va == vb, (i, f)


def test_parse_returns_statuses_on_damaged_streams():
    params = dict(SMALL, tile_width=128, tile_height=128, out_width=128, out_height=128, pcm=1, pcm_bit_depth=8,
                  p_pcm=0.2, num_slices=2, tile_cols=2)
    bs = H.gen_hevc(**params)
    rnd = random.Random(7)
    statuses = set()
    for trial in range(120):
        b = bytearray(bs)
        if trial % 3 == 0:
            b = b[:rnd.randrange(4, len(b))]
        else:
            for _ in range(1 + trial % 4):
                i = rnd.randrange(40, len(b))
                b[i] ^= 1 << rnd.randrange(8)
        try:
            H.RecordSet.parse_hevc(bytes(b))
            statuses.add(0)
        except H.HpvcdError as e:
            statuses.add(e.status)
    # damage is reported as a bitstream, desync or parameter-set error
    assert statuses - {0} <= {N.E_BITSTREAM, N.E_CABAC_DESYNC, N.E_PARAMSET, N.E_UNSUPPORTED, N.E_BAD_DIMENSIONS,
                             N.E_UNSUPPORTED_BITDEPTH, N.E_UNSUPPORTED_CHROMA, N.E_LIMIT}
    assert N.E_CABAC_DESYNC in statuses


def test_parse_refuses_missing_parameter_sets_and_heic_inter_configs():
    bs = H.gen_hevc(**SMALL)
    # drop the PPS NAL unit (third start code)
    parts = bs.split(b"\x00\x00\x00\x01")
    no_pps = b"".join(b"\x00\x00\x00\x01" + x for x in parts[1:] if x and ((x[0] >> 1) & 63) != 34)
    with pytest.raises(H.HpvcdError) as e:
        H.RecordSet.parse_hevc(no_pps)
    assert e.value.status == N.E_PARAMSET
    # image items are intra pictures: the HEIC writer refuses inter sequences
    with pytest.raises(H.HpvcdError) as e:
        H.gen_heic(**dict(SMALL, grid_rows=2, inter=1))
    assert e.value.status == N.E_UNSUPPORTED
    with pytest.raises(H.HpvcdError):
        H.RecordSet.parse_hevc(b"\x01\x02\x03")
