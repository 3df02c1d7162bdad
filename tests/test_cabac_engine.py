"""CABAC tables and engine of the record producer (CPU).

* The context layout, initValues (initType 0, and the P/B rows), rangeTabLps
  and transIdxLps/Mps the parser and the writer use equal the reference's
  (src/cabac/contexts.rs:139-400, src/cabac/engine.rs:35-70), read from the
  committed fixture tests/golden/ref_tables.json.
* 9.3.2.2 initialisation: initValue 154 is the equiprobable state (pStateIdx
  0, valMps 1) at every QP, and the state moves monotonically with QP for
  a nonzero slope.
* tests/cabac_kat.cpp, built from the library's host sources: the
  reference's own engine tests (PCM re-alignment, bypass alignment), the
  encoder/decoder inverse property, Exp-Golomb and emulation-prevention
  known codes, parameter-set round trips and the 4:2:2 refusal.
"""
import ctypes as C
import json
import os
import subprocess

import pytest

import hpvcd_amd as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_tables.json")))


def _layout():
    L = H.N.lib()
    out = {}
    for i in range(L.hpvcd_cabac_num_elements()):
        name, off, cnt = C.c_char_p(), C.c_int32(), C.c_int32()
        assert L.hpvcd_cabac_element(i, C.byref(name), C.byref(off), C.byref(cnt)) == 0
        out[name.value.decode()] = (off.value, cnt.value)
    return out


def _tables():
    L = H.N.lib()
    n = L.hpvcd_cabac_tables(None, None, None, None)
    init = (C.c_uint8 * (3 * n))()
    lps = (C.c_uint8 * 256)()
    tl = (C.c_uint8 * 64)()
    tm = (C.c_uint8 * 64)()
    L.hpvcd_cabac_tables(init, lps, tl, tm)
    rows = [list(init[t * n:(t + 1) * n]) for t in range(3)]
    return n, rows, [list(lps[4 * s:4 * s + 4]) for s in range(64)], list(tl), list(tm)


def test_engine_tables_match_reference():
    _, _, lps, tl, tm = _tables()
    assert lps == FIX["cabac_range_lps"]
    assert tl == FIX["cabac_trans_lps"]
    assert tm == FIX["cabac_trans_mps"]


def test_islice_init_values_match_reference():
    lay = _layout()
    _, rows, _, _, _ = _tables()
    ref = FIX["cabac_init_islice"]
    checked = 0
    for name, (off, cnt) in lay.items():
        assert name in ref, name
        assert rows[0][off:off + cnt] == ref[name], name
        checked += 1
    # every element the I-slice parse reads is covered
    for name in ("sao_merge_flag", "sao_type_idx", "split_cu_flag", "cu_transquant_bypass_flag", "part_mode",
                 "prev_intra_luma_pred_flag", "intra_chroma_pred_mode", "split_transform_flag", "cbf_luma",
                 "cbf_chroma", "cu_qp_delta_abs", "transform_skip_flag", "last_sig_coeff_x_prefix",
                 "last_sig_coeff_y_prefix", "coded_sub_block_flag", "sig_coeff_flag", "coeff_abs_level_greater1",
                 "coeff_abs_level_greater2", "explicit_rdpcm_flag", "explicit_rdpcm_dir",
                 "log2_res_scale_abs_plus1", "res_scale_sign_flag", "cu_residual_act_flag"):
        assert name in lay, name
    assert checked == len(lay)


@pytest.mark.parametrize("it", [1, 2])
def test_pb_init_rows_match_reference(it):
    lay = _layout()
    _, rows, _, _, _ = _tables()
    t = FIX["cabac_init_typed"]
    row = rows[it]

    def got(name):
        off, cnt = lay[name]
        return row[off:off + cnt]

    assert got("split_cu_flag") == t["split_cu"][it]
    assert got("split_transform_flag") == t["split_tf"][it]
    assert got("cbf_luma") == t["cbf_luma"][it]
    c = t["cbf_chroma"][it]
    assert got("cbf_chroma") == c + [c[3]]
    assert got("last_sig_coeff_x_prefix") == t["last_sig_prefix"][it * 18:(it + 1) * 18]
    assert got("last_sig_coeff_y_prefix") == t["last_sig_prefix"][it * 18:(it + 1) * 18]
    s = t["sig_coeff_flag"][it]
    assert got("sig_coeff_flag") == s + [s[41], s[41]]
    assert got("coded_sub_block_flag") == t["coded_sub_block_flag"][it]
    assert got("coeff_abs_level_greater1") == t["coeff_g1"][it * 24:(it + 1) * 24]
    assert got("coeff_abs_level_greater2") == t["coeff_g2"][it * 6:(it + 1) * 6]
    assert got("sao_type_idx") == [t["sao_type_idx"][it]]
    off = it if it != 2 else 5
    assert got("part_mode") == t["part_mode_pm"][off:off + 4]
    assert got("prev_intra_luma_pred_flag") == [t["prev_intra_luma_pred_flag"][it]]
    assert got("intra_chroma_pred_mode") == [t["intra_chroma_pred_mode"][it]]


def test_init_state_special_cases():
    L = H.N.lib()
    n, rows, _, _, _ = _tables()
    out = (C.c_uint8 * n)()
    for qp in (0, 22, 37, 51, -12, 60):
        L.hpvcd_cabac_init_states(0, qp, out)
        for i in range(n):
            if rows[0][i] == 154:  # slope 9 -> m = 0; offset 10 -> n = 64: equiprobable
                assert out[i] == (1 << 6), (qp, i)
    # a positive slope raises preCtxState with QP: the MPS=1 probability grows
    # (initValue 224: slope 14, m = 25)
    lay = _layout()
    off = lay["split_transform_flag"][0]
    prev = None
    for qp in range(0, 52):
        L.hpvcd_cabac_init_states(2, qp, out)
        v = out[off]
        pre = (v & 63) + 64 if v >> 6 else 63 - (v & 63)
        if prev is not None:
            assert pre >= prev
        prev = pre


def test_engine_known_answers(tmp_path):
    exe = str(tmp_path / "cabac_kat")
    src = [os.path.join(ROOT, "tests", "cabac_kat.cpp")] + [
        os.path.join(ROOT, "hpvcd_amd", "csrc", f) for f in ("cabac.cpp", "hevc_ps.cpp", "walker.cpp", "motion.cpp", "parse.cpp")]
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", exe] + src + ["-lpthread"], check=True,
                   timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    lines = r.stdout.strip().splitlines()
    assert r.returncode == 0, r.stdout
    names = {l.split()[0] for l in lines if l.endswith(" ok")}
    assert names == {"bypass_runs", "pcm_realign", "bypass_alignment", "engine_inverse", "expgolomb", "emulation_prevention",
                     "sps_roundtrip", "pps_roundtrip", "tile_uniform_spacing", "chroma_422_refused"}, r.stdout
