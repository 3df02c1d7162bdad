"""Host side: the C-ABI library, the syntax generator and record validation.

No GPU needed: nothing here launches a kernel.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import hpvcd_amd as H
from hpvcd_amd import _native as N
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for hdr in ("hpvcd.h",):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(hpvcd_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = N.lib()
    declared = _declared_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
        assert name in N.SIGNATURES, f"{name} has no ctypes signature"
    assert L.hpvcd_version().startswith(b"hpvcd")


def test_struct_layouts_match_c():
    assert C.sizeof(N.TB) == 24
    assert C.sizeof(N.Sao) == 36
    assert C.sizeof(N.Output) == 12
    assert C.sizeof(N.Picture) == 140
    assert C.sizeof(N.Batch) == 144
    assert C.sizeof(N.PU) == 20
    assert C.sizeof(N.Ref) == 20
    assert C.sizeof(N.GenConfig) == 176
    assert C.sizeof(N.Settings) == 56 and C.sizeof(N.Image) == 40


def test_generator_is_deterministic_and_seeded():
    a = H.RecordSet.generate("single512")
    b = H.RecordSet.generate("single512")
    c = H.RecordSet.generate("single512", base_seed=99)
    assert np.array_equal(a.tbs(), b.tbs()) and np.array_equal(a.coeffs(), b.coeffs())
    assert not np.array_equal(a.coeffs()[:1000], c.coeffs()[:1000])
    # per-tile seeds differ across the grid
    g = H.RecordSet.generate("grid4032", grid_cols=2, grid_rows=1, out_width=1024, out_height=512)
    p0, p1 = g.picture(0), g.picture(1)
    assert p0.tb_count != p1.tb_count or p0.coeff_count != p1.coeff_count


def _luma_tbs(rs, p):
    t = rs.tbs()
    pic = rs.picture(p)
    t = t[pic.tb_begin:pic.tb_begin + pic.tb_count]
    return t, pic


def test_tbs_tile_the_picture_in_z_order():
    rs = H.RecordSet.generate("single512", base_seed=3)
    t, pic = _luma_tbs(rs, 0)
    cover = np.zeros((pic.height // 4, pic.width // 4), np.int32)
    for r in t[t["comp"] == 0]:
        n4 = (1 << r["log2n"]) // 4
        cover[r["y"] // 4:r["y"] // 4 + n4, r["x"] // 4:r["x"] // 4 + n4] += 1
    assert np.all(cover == 1)
    for comp in (1, 2):
        cov = np.zeros((pic.height // 8, pic.width // 8), np.int32)
        for r in t[t["comp"] == comp]:
            n4 = (1 << r["log2n"]) // 4
            cov[r["y"] // 4:r["y"] // 4 + n4, r["x"] // 4:r["x"] // 4 + n4] += 1
        assert np.all(cov == 1)


def test_deblock_edges_are_the_transform_edges():
    """Intra pictures: bS = 2 exactly on TU edges (independent derivation)."""
    rs = H.RecordSet.generate("single512", base_seed=5)
    t, pic = _luma_tbs(rs, 0)
    gw, gh = pic.width // 4, pic.height // 4
    want_v = np.zeros((gh, gw), np.int32)
    want_h = np.zeros((gh, gw), np.int32)
    for r in t[t["comp"] == 0]:
        n4 = (1 << r["log2n"]) // 4
        gx, gy = r["x"] // 4, r["y"] // 4
        if gx > 0:
            want_v[gy:gy + n4, gx] = 2
        if gy > 0:
            want_h[gy, gx:gx + n4] = 2
    cells = rs.cells()[pic.cell_begin:pic.cell_begin + gw * gh].reshape(gh, gw)
    assert np.array_equal(cells & 3, want_v)
    assert np.array_equal((cells >> 2) & 3, want_h)
    qp = ((cells >> 8) & 0xff).astype(np.int8)
    assert np.all(qp == 32)


def test_chroma_modes_follow_table_8_2():
    rs = H.RecordSet.generate("single512", base_seed=7)
    t, pic = _luma_tbs(rs, 0)
    luma_mode = np.zeros((pic.height // 4, pic.width // 4), np.int32)
    for r in t[t["comp"] == 0]:
        n4 = (1 << r["log2n"]) // 4
        luma_mode[r["y"] // 4:r["y"] // 4 + n4, r["x"] // 4:r["x"] // 4 + n4] = r["mode"]
    for r in t[t["comp"] == 1]:
        lm = luma_mode[(r["y"] * 2) // 4, (r["x"] * 2) // 4]
        assert r["mode"] in (0, 26, 10, 1, 34) or r["mode"] == lm
        if r["mode"] == 34:
            assert lm in (0, 26, 10, 1, 34)


def test_qp_prime_of_chroma_tbs():
    rs = H.RecordSet.generate("single512", base_seed=9, qp=37)
    t, _ = _luma_tbs(rs, 0)
    assert np.all(t[t["comp"] == 0]["qp"] == 37)
    assert np.all(t[t["comp"] > 0]["qp"] == O.lib().oracle_qpc(37, 1, 8))


def test_coefficient_statistics_follow_the_distribution():
    rs = H.RecordSet.generate("single512", base_seed=11)
    t, _ = _luma_tbs(rs, 0)
    coded = t[(t["flags"] & N.TB_CBF) != 0]
    area = int(sum((1 << (2 * int(l))) for l in coded["log2n"]))
    nnz = int(coded["coeff_cnt"].sum())
    assert 0.27 < nnz / area < 0.33
    luma = t[t["comp"] == 0]
    frac_cbf = np.mean((luma["flags"] & N.TB_CBF) != 0)  # cbf_luma p = 0.6
    assert 0.55 < frac_cbf < 0.65
    lv = np.abs(rs.coeffs().astype(np.int32) >> 10)
    assert 0.45 < np.mean(lv == 1) < 0.55


def test_rext_records_carry_ts_and_ccp():
    rs = H.RecordSet.generate("rext4096", grid_cols=1, grid_rows=1, out_width=512, out_height=512)
    t = rs.tbs()
    ts = t[(t["flags"] & N.TB_TS) != 0]
    assert len(ts) > 0 and np.all(ts["log2n"] == 2)
    assert set(np.unique(t["res_scale"])) <= {0, 1, -1, 2, -2, 4, -4}
    assert np.any(t["res_scale"] != 0)


def test_act_records_are_colocated_triples_with_act_qps():
    rs = H.RecordSet.generate("rext4096", grid_cols=1, grid_rows=1, out_width=512, out_height=512, act=1,
                              p_act=0.5, base_seed=41)
    t = rs.tbs()
    idx = np.flatnonzero((t["flags"] & N.TB_ACT) != 0)
    assert len(idx) >= 30 and len(idx) % 3 == 0
    for i in idx[t["comp"][idx] == 0]:
        tri = t[i:i + 3]
        assert list(tri["comp"]) == [0, 1, 2] and np.all(tri["flags"] & N.TB_ACT)
        assert len(set(zip(tri["x"], tri["y"], tri["log2n"]))) == 1
        # QP 22, 10-bit: Y qp' = 17 + 12, Cb = qpc(17) + 12, Cr = qpc(19) + 12 (src/decode.rs:3788)
        assert list(tri["qp"]) == [29, 29, 31]
        assert tri["mode"][1] == tri["mode"][2] == tri["mode"][0]


def test_pcm_records_carry_raw_samples():
    rs = H.RecordSet.generate("single512", pcm=1, p_pcm=0.2, pcm_bit_depth=5, base_seed=43)
    t, co = rs.tbs(), rs.coeffs()
    pcm = t[(t["flags"] & N.TB_PCM) != 0]
    assert len(pcm) > 30 and np.all(pcm["qp"] == 3)
    luma = pcm[pcm["comp"] == 0]
    assert set(np.unique(luma["log2n"])) <= {3, 4, 5} and len(pcm) == 3 * len(luma)
    for tb in pcm[:40]:
        n = 1 << int(tb["log2n"])
        w = co[int(tb["coeff_off"]):int(tb["coeff_off"]) + int(tb["coeff_cnt"])]
        assert tb["coeff_cnt"] == n * n and np.array_equal(w & 0x3ff, np.arange(n * n))
        assert np.all((w.view(np.int32) >> 10) >= 0) and np.all((w >> 10) < 32)


def _status(rs):
    return N.lib().hpvcd_validate_batch(C.byref(rs.batch))


def test_validation_accepts_generated_and_rejects_malformed():
    rs = H.RecordSet.generate("single512", base_seed=13)
    assert _status(rs) == 0
    tbs = (N.TB * rs.batch.num_tbs).from_address(C.addressof(rs.batch.tbs.contents))
    coeffs = (C.c_uint32 * rs.batch.num_coeffs).from_address(C.addressof(rs.batch.coeffs.contents))
    pics = rs.batch.pictures

    def expect(code, msg):
        st = _status(rs)
        assert st == code, (st, N.lib().hpvcd_last_error())
        assert msg in N.lib().hpvcd_last_error().decode()

    # a coefficient position outside its n x n block
    i = next(k for k in range(rs.batch.num_tbs) if tbs[k].flags & N.TB_CBF and tbs[k].log2n == 2)
    old = coeffs[pics[0].coeff_begin + tbs[i].coeff_off]
    coeffs[pics[0].coeff_begin + tbs[i].coeff_off] = (old & ~0x3ff) | 16
    expect(N.lib().hpvcd_validate_batch(C.byref(rs.batch)) and 15, "coefficient position")
    coeffs[pics[0].coeff_begin + tbs[i].coeff_off] = old
    # a TB crossing a CTB boundary
    x = tbs[5].x
    tbs[5].x = 62
    expect(15, "CTB boundary")
    tbs[5].x = x
    # bit depth / dimensions / output placement
    pics[0].bit_depth_luma = 16
    expect(8, "bit depth")
    pics[0].bit_depth_luma = 8
    pics[0].width = 500
    expect(9, "multiple of 8")
    pics[0].width = 512
    pics[0].dst_x = 100
    expect(15, "outside its output")
    pics[0].dst_x = 0
    pics[0].out_index = 3
    expect(15, "out_index")
    pics[0].out_index = 0
    pics[0].tb_count += 1
    expect(15, "tb range")
    pics[0].tb_count -= 1
    # ACT outside 4:4:4; PCM shift and sample range
    f, q = tbs[i].flags, tbs[i].qp
    tbs[i].flags = f | N.TB_ACT
    expect(15, "ACT")
    tbs[i].flags = N.TB_PCM
    tbs[i].qp = 8
    expect(15, "PCM sample shift")
    tbs[i].qp = 7
    w = pics[0].coeff_begin + tbs[i].coeff_off
    old = coeffs[w]
    coeffs[w] = (old & 0x3ff) | (5 << 10)
    expect(15, "PCM sample outside")
    coeffs[w] = old
    tbs[i].flags, tbs[i].qp = f, q
    assert _status(rs) == 0


def test_oracle_independent_of_thread_count_and_job_split():
    rs = H.RecordSet.generate("grid4032", grid_cols=3, grid_rows=2, out_width=1500, out_height=1000, base_seed=17)
    outs = []
    for threads, rows in ((1, True), (8, True), (8, False), (3, False)):
        b = rs.output_buffers()
        O.decode_batch(rs.batch, b, threads=threads, rows_parallel=rows)
        outs.append(b[0])
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


def test_loop_filters_are_exercised():
    """The parity cases are not vacuous: deblocking and SAO change samples."""
    rs = H.RecordSet.generate("single512", p_sig=0.02, p_cbf=0.35, qp=37, base_seed=11)
    rec = O.picture_planes(rs.batch, 0, O.STAGE_RECON)
    deb = O.picture_planes(rs.batch, 0, O.STAGE_DEBLOCK)
    sao = O.picture_planes(rs.batch, 0, O.STAGE_SAO)
    for c in range(3):
        assert np.count_nonzero(rec[c] != deb[c]) > 1000
        assert np.count_nonzero(deb[c] != sao[c]) > 1000


def test_lossless_cus_are_prediction_plus_levels():
    """Transquant bypass: residual == levels.  The first TB of each plane has
    no available neighbour, so every intra mode predicts the neutral 1 << (bd - 1)
    and its reconstruction is exactly 128 + its levels; the loop filters leave
    lossless CUs untouched."""
    for seed in (19, 23, 29):
        rs = H.RecordSet.generate("single512", p_tqb=1.0, p_sig=0.05, base_seed=seed)
        rec = O.picture_planes(rs.batch, 0, O.STAGE_RECON)
        out = O.picture_planes(rs.batch, 0, O.STAGE_SAO)
        tbs, coeffs = rs.tbs(), rs.coeffs()
        for comp in range(3):
            t = tbs[np.nonzero(tbs["comp"] == comp)[0][0]]
            n = 1 << int(t["log2n"])
            want = np.full((n, n), 128, np.int64)
            if t["flags"] & N.TB_CBF:
                for w in coeffs[t["coeff_off"]:t["coeff_off"] + t["coeff_cnt"]]:
                    pos, level = int(w) & 0x3FF, int(np.int32(w)) >> 10
                    want[pos // n, pos % n] += level
            got = rec[comp][t["y"]:t["y"] + n, t["x"]:t["x"] + n]
            assert np.array_equal(got, np.clip(want, 0, 255)), (seed, comp)
            assert np.array_equal(rec[comp], out[comp])


# --------------------------------------------------------------------------------------
# generator options that reach the remaining record paths
# --------------------------------------------------------------------------------------
def test_cu_qp_delta_varies_qp_and_chroma_offsets_reach_the_tbs():
    rs = H.RecordSet.generate("single512", cu_qp_delta=1, diff_cu_qp_delta_depth=2, qp_delta_max=6,
                              pps_cb_qp_offset=3, slice_cb_qp_offset=-2, pps_cr_qp_offset=-4, base_seed=71)
    qp = ((rs.cells() >> 8) & 0xFF).astype(np.int8)
    assert len(np.unique(qp)) > 8 and qp.min() >= 0 and qp.max() <= 51
    t = rs.tbs()
    # a coded luma TB's Qp' equals the QpY of its CU's cells (8-bit: no QpBdOffset)
    luma = t[(t["comp"] == 0) & ((t["flags"] & N.TB_CBF) != 0)]
    gw = 512 // 4
    assert np.array_equal(luma["qp"], qp[(luma["y"].astype(int) // 4) * gw + luma["x"].astype(int) // 4])
    P = rs.picture(0)
    assert P.pps_cb_qp_offset == 3 and P.pps_cr_qp_offset == -4


def test_slices_and_tiles_in_the_cells():
    rs = H.RecordSet.generate("single512", num_slices=3, tile_cols=2, tile_rows=2, random_slice_params=1,
                              lf_across_tiles=0, base_seed=73)
    P = rs.picture(0)
    c = rs.cells().reshape(128, 128)
    tile = (c >> 24) & 0xFF
    sl = (c >> 16) & 0xFF
    assert P.num_slices == 3 and P.flags & N.PF_TILES and not P.flags & N.PF_LF_ACROSS_TILES
    # tiles: 2 x 2 uniform over the 8 x 8 CTBs (CTB columns 0-3 / 4-7)
    assert np.array_equal(np.unique(tile), [0, 1, 2, 3])
    assert (tile[:, :64] % 2 == 0).all() and (tile[:, 64:] % 2 == 1).all()
    assert (tile[:64] < 2).all() and (tile[64:] >= 2).all()
    # slices: runs of the tile scan, so slice 0 lies inside tile 0
    assert set(np.unique(sl)) == {0, 1, 2}
    assert (sl[tile == 0] <= 1).all() and (sl[tile == 3] == 2).all()
    assert N.lib().hpvcd_validate_batch(C.byref(rs.batch)) == 0
    # TBs stay in CTB raster order
    t = rs.tbs()
    luma = t[t["comp"] == 0]
    ctb = (luma["y"].astype(int) >> 6) * 8 + (luma["x"].astype(int) >> 6)
    assert (np.diff(ctb) >= 0).all()


def test_scaling_lists_in_the_records():
    for kind in (1, 2):
        rs = H.RecordSet.generate("single512", scaling_list=kind, base_seed=75)
        assert rs.batch.num_scaling == 1 and rs.picture(0).scaling_index == 0
        sc = np.ctypeslib.as_array(rs.batch.scaling, shape=(N.SCALING_SET_BYTES,)).reshape(4, 6, -1)
        assert sc[0, :, 65].all() == (kind == 1)            # default 4x4 lists are flat
        assert not sc[1:, :, 65].any()
        assert np.array_equal(sc[3, 1], sc[2, 1]) and np.array_equal(sc[3, 5], sc[2, 5])
        if kind == 1:
            import json
            T = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_tables.json")))
            assert list(sc[1, 0, :64]) == T["scaling_default_intra_8x8"]
            assert list(sc[2, 4, :64]) == T["scaling_default_inter_8x8"]


def test_conformance_window_places_grid_tiles():
    rs = H.RecordSet.generate("grid4032", grid_cols=3, grid_rows=2, tile_width=256, tile_height=128,
                              out_width=3 * 232, out_height=2 * 112, conf_win=(8, 16, 4, 12), base_seed=77)
    for i in range(6):
        P = rs.picture(i)
        assert (P.crop_left, P.crop_top, P.win_w, P.win_h) == (8, 4, 232, 112)
        assert (P.dst_x, P.dst_y) == ((i % 3) * 232, (i // 3) * 112)
    assert N.lib().hpvcd_validate_batch(C.byref(rs.batch)) == 0


def test_decoder_settings_defaults_and_backends():
    """HeicSettings::new() + ParseLimits::new() (src/settings.rs:59,
    src/limits.rs:85-102); the CPU_REF backend is the reference decoder itself
    and is refused, as is an unknown backend, before any GPU is touched."""
    s = N.Settings()
    N.lib().hpvcd_settings_default(C.byref(s))
    assert (s.threads, s.max_dimension, s.max_pixels) == (0, 16384, 64 << 20)
    assert (s.max_box_size, s.max_item_size, s.max_items, s.max_tiles) == (256 << 20, 256 << 20, 4096, 16384)
    assert s.decode_alpha == 1 and s.decode_gain_map == 1 and s.backend == N.BACKEND_GPU_BATCHED
    h = C.c_void_p()
    s.backend = N.BACKEND_CPU_REF
    assert N.lib().hpvcd_decoder_new(C.byref(s), C.byref(h)) == 11 and not h.value   # E_UNSUPPORTED
    s.backend = 7
    assert N.lib().hpvcd_decoder_new(C.byref(s), C.byref(h)) == 15 and not h.value   # E_INVALID


def test_hierarchical_gop_structure():
    """Random access with hierarchical B GOPs of 8 over 32 frames: pictures
    in decode order (each GOP's end first, then midpoints), outputs in
    display order, references decoded earlier, inside the GOP (or the earlier
    GOP ends for a GOP's end), never the top temporal layer, at most 4; the
    reference chain is 8 pictures deep instead of 32."""
    rs = H.RecordSet.generate("video1080", tile_width=256, tile_height=128, out_width=256, out_height=128)
    n = rs.num_pictures
    poc = [rs.picture(i).out_index for i in range(n)]
    assert sorted(poc) == list(range(32))
    assert poc[:9] == [0, 8, 4, 2, 1, 3, 6, 5, 7]
    assert poc[25:] == [31, 27, 25, 26, 29, 28, 30]
    referenced = set()
    for i in range(1, n):
        p = rs.picture(i)
        refs = {rs.batch.refs[p.ref_begin + k].picture for k in range(p.ref_count)}
        assert refs and all(r < i for r in refs) and len(refs) <= 4
        rp = [poc[r] for r in refs]
        a = (poc[i] - 1) // 8 * 8
        b = min(a + 8, 31)
        if poc[i] == b:
            assert all(q % 8 == 0 and q < poc[i] for q in rp)
        else:
            assert all(a <= q <= b for q in rp) and any(q > poc[i] for q in rp)
        referenced.update(rp)
    # the top temporal layer: odd POCs of the full GOPs, the even ones of the 7-frame last GOP
    assert not referenced & (set(range(1, 24, 2)) | {26, 28, 30})
    assert rs.dependency_depth() == 8
    low_delay = H.RecordSet.generate("video1080", tile_width=256, tile_height=128, out_width=256, out_height=128,
                                     gop=0)
    assert [low_delay.picture(i).out_index for i in range(n)] == list(range(32))
    assert low_delay.dependency_depth() == 32
    with pytest.raises(H.HpvcdError):
        H.RecordSet.generate("single512", gop=8)


def test_tuning_defaults_checks_and_cache_file(tmp_path):
    """Launch-shape knobs: defaults, the value lists hpvcd_tuning_check
    accepts, and the cache file (one line per GPU key, the key's line
    replaced on save, malformed or out-of-list lines skipped on load)."""
    d = H.default_tuning()
    assert d == {"residual_block": [256, 256, 256, 256], "deblock_block": 256, "rgb_rows": 4, "intra_per_cu": 0,
                 "sao_parts": 4}
    for bad in ({"residual_block": [256, 96, 256, 256]}, {"deblock_block": 32}, {"rgb_rows": 3},
                {"intra_per_cu": 9}, {"sao_parts": 0}):
        t = H.tuning_struct(bad)
        assert N.lib().hpvcd_tuning_check(C.byref(t)) == 15, bad
    f = str(tmp_path / "tune.txt")
    with pytest.raises(H.HpvcdError):
        H.load_tuning(f, "gfx950:256cu")
    a = dict(d, residual_block=[64, 128, 256, 64], rgb_rows=2, intra_per_cu=2, sao_parts=1)
    b = dict(d, deblock_block=128)
    H.save_tuning(f, "gfx950:256cu", a)
    H.save_tuning(f, "gfx942:304cu", b)
    assert H.load_tuning(f, "gfx950:256cu") == a and H.load_tuning(f, "gfx942:304cu") == b
    c = dict(a, rgb_rows=1)
    H.save_tuning(f, "gfx950:256cu", c)
    lines = open(f).read().splitlines()
    assert len(lines) == 2 and H.load_tuning(f, "gfx950:256cu") == c
    with open(f, "a") as fh:
        fh.write("gfx950:256cu 64 64 64\n")                   # too few fields
        fh.write("gfx950:256cu 64 64 64 64 100 4 0 4\n")      # deblock 100 is not a choice
        fh.write("# a comment line\n")
    assert H.load_tuning(f, "gfx950:256cu") == c
    with pytest.raises(H.HpvcdError):
        H.load_tuning(f, "gfx950:1cu")
    with pytest.raises(H.HpvcdError):
        H.save_tuning(f, "two words", a)


def test_server_config_and_refusals():
    """The serving loop's defaults, and the arguments it refuses before any
    GPU is touched: the CPU_REF backend, zero caps, null arguments."""
    c = N.ServerConfig()
    N.lib().hpvcd_server_config_default(C.byref(c))
    assert (c.max_batch_pictures, c.parse_workers) == (256, 4)
    assert (c.max_batch_bytes, c.hbm_reserve) == (16 << 30, 8 << 30)
    s = N.Settings()
    N.lib().hpvcd_settings_default(C.byref(s))
    h = C.c_void_p()
    s.backend = N.BACKEND_CPU_REF
    assert N.lib().hpvcd_server_new(C.byref(s), C.byref(c), C.byref(h)) == 11 and not h.value   # E_UNSUPPORTED
    s.backend = N.BACKEND_GPU_BATCHED
    c.parse_workers = 0
    assert N.lib().hpvcd_server_new(C.byref(s), C.byref(c), C.byref(h)) == 15 and not h.value   # E_INVALID
    assert N.lib().hpvcd_server_new(C.byref(s), C.byref(c), None) == 15
    t = C.c_uint64()
    assert N.lib().hpvcd_server_submit(None, b"x", 1, 0, C.byref(t)) == 15


def test_batch_preparation_is_the_same_on_every_thread_count():
    """Job preparation (availability stamps, wavefront units, dependency
    levels, reference-map layout, residual and ticket lists) runs per picture
    on host threads and merges in picture order: its digest does not depend
    on the thread count."""
    import ctypes as C

    sets = [H.RecordSet.generate("grid4032", grid_cols=3, grid_rows=2, out_width=1536, out_height=1024),
            H.RecordSet.generate("rext4096", grid_cols=2, grid_rows=1, out_width=1024, out_height=512, act=1,
                                 p_act=0.5, p_tqb=0.2),
            H.RecordSet.generate("video1080", tile_width=256, tile_height=128, grid_rows=5, num_outputs=5,
                                 out_width=256, out_height=128)]
    sets.append(H.RecordSet.concat(sets))
    for rs in sets:
        digests = set()
        for th in (1, 2, 5, 16):
            d, ms = C.c_uint64(), C.c_double()
            assert N.lib().hpvcd_prep_digest(C.byref(rs.batch), th, C.byref(d), C.byref(ms)) == 0
            digests.add(d.value)
        assert len(digests) == 1
