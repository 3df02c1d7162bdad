"""Coded inter sequences (CPU): writer -> video parse -> records.

The generator codes P and B pictures (hpvcd_gen_hevc): slice headers with
POC LSBs, short-term RPSs (inline, or in the SPS with inter RPS prediction),
reference list sizes and modifications, mvd_l1_zero, cabac_init, the
collocated picture, pred_weight_table and MaxNumMergeCand; CUs with
cu_skip_flag, pred_mode_flag, part_mode (AMP), merge_idx or
inter_pred_idc / ref_idx / mvd / mvp flag per PU, rqt_root_cbf.

hpvcd_parse_hevc reads it the way the reference's video decoder does
(src/video.rs:1045 decode_access_unit): POC from the previous TemporalId-0
picture (src/rps.rs:208), RPS marking (src/dpb.rs:217), lists from
StCurrBefore / StCurrAfter with list_entry (src/dpb.rs:298, :465), weights
(src/inter.rs:159), and per PU the merge / AMVP / temporal derivations of
src/motion.rs (pinned separately in test_motion_pins.py).

The generator decides its reference lists directly as batch pictures (the
GOP structure, gen_lists), not through RPSs, so a parse that equals the
generated records byte for byte pins POC derivation, RPS marking and list
construction against that independent definition, and the inter syntax
against the generator's decisions.
"""
import pytest

import hpvcd_amd as H
from hpvcd_amd import N

BASE = dict(tile_width=256, tile_height=192, out_width=256, out_height=192, grid_cols=1, inter=1, num_refs=3,
            qp=30, grid_rows=8, num_outputs=8, p_bi=0.5)

CASES = {
    "p_low_delay": dict(p_bi=0.0),
    "b_low_delay_list1_modified": dict(),
    "gop8_tmvp_weighted": dict(grid_rows=12, num_outputs=12, gop=8, tmvp=1, weighted=1),
    "rps_in_sps_gop4": dict(gop=4, rps_in_sps=1, tmvp=1),
    "cabac_init": dict(cabac_init=1, tmvp=1, gop=4),
    "one_merge_candidate": dict(merge_cand=1, p_merge=0.8),
    "two_candidates_merge_level3": dict(merge_cand=2, par_merge_level=3, p_merge=0.7),
    "merge_level6_tmvp": dict(par_merge_level=6, p_merge=0.7, tmvp=1),
    "mvd_l1_zero": dict(mvd_l1_zero=1, p_bi=0.9),
    "no_amp_ctb32": dict(amp=0, log2_ctb=5, log2_max_tb=5),
    "ctb16": dict(log2_ctb=4, log2_min_cb=3, log2_max_tb=4, max_transform_depth_intra=2),
    "min_cu16_nxn": dict(log2_ctb=6, log2_min_cb=4),
    "inter_split_depth0": dict(max_transform_depth_intra=0),
    "slices_tiles": dict(num_slices=3, tile_cols=2, tile_rows=2, tmvp=1),
    "wpp": dict(wpp=1, tmvp=1),
    "poc_wrap_low_delay": dict(grid_rows=40, num_outputs=40, num_refs=2),
    "poc_wrap_gop8": dict(grid_rows=40, num_outputs=40, gop=8, tmvp=1),
    "gop16_four_refs": dict(grid_rows=33, num_outputs=33, gop=16, num_refs=4, tmvp=1, rps_in_sps=1),
    "ten_bit_weighted": dict(bit_depth=10, out_kind=N.OUT_RGB16, weighted=1),
    "yuv444_weighted": dict(chroma_idc=3, weighted=1, tmvp=1),
    "mono_weighted": dict(chroma_idc=0, out_kind=N.OUT_LUMA8, weighted=1),
    "tqb_qp_delta": dict(p_tqb=0.2, cu_qp_delta=1, diff_cu_qp_delta_depth=1, qp_delta_max=6),
    "mostly_skip": dict(p_skip=0.8, p_merge=0.8, tmvp=1),
    "mostly_intra": dict(p_intra_cu=0.6),
    "one_ref": dict(num_refs=1, tmvp=1),
    "eight_refs_gop8": dict(num_refs=8, gop=8, grid_rows=17, num_outputs=17, tmvp=1),
    "p_gop4_weighted": dict(p_bi=0.0, gop=4, tmvp=1, weighted=1),
    "ts_rdpcm": dict(transform_skip=1, implicit_rdpcm=1, p_ts=0.4),
    "sign_hiding": dict(sign_hiding=1, tmvp=1),
    "odd_size": dict(tile_width=200, tile_height=136, out_width=200, out_height=136, tmvp=1),
}

ARRAYS = ("tbs", "coeffs", "sao", "cell_info", "scaling", "pus", "refs", "pictures", "outputs")


@pytest.mark.parametrize("name", sorted(CASES))
def test_parse_of_coded_sequence_equals_records(name):
    params = dict(BASE, **CASES[name])
    bs = H.gen_hevc(**params)
    g = H.RecordSet.generate(**params)
    p = H.RecordSet.parse_hevc(bs, out_kind=H.gen_config(**params).out_kind)
    assert p.num_pictures == g.num_pictures == params["grid_rows"]
    for a in ARRAYS:
        assert g.raw(a) == p.raw(a), a
    assert len(p.raw("pus")) > 0


def test_sequence_structure_of_a_gop():
    """Pictures come in decode order and are output in display order: a GOP
    of 8 decodes 0, 8, 4, 2, 1, 3, 6, 5, 7 (each picture's output index is its
    POC), and every reference entry names an earlier picture."""
    params = dict(BASE, grid_rows=9, num_outputs=9, gop=8, tmvp=1)
    p = H.RecordSet.parse_hevc(H.gen_hevc(**params))
    assert [p.picture(i).out_index for i in range(9)] == [0, 8, 4, 2, 1, 3, 6, 5, 7]
    for i in range(1, 9):
        pic = p.picture(i)
        for r in range(pic.ref_count):
            assert p.ref(pic.ref_begin + r).picture < i


def test_truncated_and_damaged_sequences_return_statuses():
    import random
    params = dict(BASE, grid_rows=4, num_outputs=4, tmvp=1, weighted=1, tile_width=128, tile_height=128,
                  out_width=128, out_height=128)
    bs = H.gen_hevc(**params)
    rnd = random.Random(3)
    statuses = set()
    for trial in range(80):
        b = bytearray(bs)
        if trial % 3 == 0:
            b = b[:rnd.randrange(4, len(b))]
        else:
            for _ in range(1 + trial % 3):
                i = rnd.randrange(40, len(b))
                b[i] ^= 1 << rnd.randrange(8)
        try:
            H.RecordSet.parse_hevc(bytes(b))
            statuses.add(0)
        except H.HpvcdError as e:
            statuses.add(e.status)
    assert statuses - {0} <= {N.E_BITSTREAM, N.E_CABAC_DESYNC, N.E_PARAMSET, N.E_UNSUPPORTED, N.E_BAD_DIMENSIONS,
                             N.E_UNSUPPORTED_BITDEPTH, N.E_UNSUPPORTED_CHROMA, N.E_LIMIT}
