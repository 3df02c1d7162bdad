"""Small hand-built record sets whose reconstruction has a closed form.

Used by the CPU oracle checks (the closed form pins the oracle) and by the
GPU parity tests (the same records through the HIP back end).
"""
import ctypes as C

import numpy as np

import hpvcd_amd as H
from hpvcd_amd import _native as N


def picture(w, h, chroma_idc, bd, log2_ctb=4, flags=N.PF_DEBLOCK_DISABLED, deblock=False):
    p = N.Picture()
    p.width, p.height, p.chroma_idc = w, h, chroma_idc
    p.bit_depth_luma = p.bit_depth_chroma = bd
    p.log2_ctb = log2_ctb
    p.flags = flags | N.PF_LF_ACROSS_SLICES
    p.num_slices = 1
    p.slices[0].disabled = 0 if deblock else 1
    p.slices[0].lf_across = 1
    p.slice_qp = 30
    p.win_w, p.win_h, p.copy_w, p.copy_h = w, h, w, h
    p.scaling_index = -1
    return p


def output(w, h, chroma_idc, kind=N.OUT_RGB8):
    """An output image record; the channel count follows the kind (a 4:0:0
    picture written as RGB gives gray triples)."""
    return N.Output(width=w, height=h, kind=kind, channels=1 if kind in (N.OUT_LUMA8, N.OUT_LUMA16) else 3,
                    bytes_per_sample=2 if kind in (N.OUT_RGB16, N.OUT_LUMA16) else 1)


def tb(x, y, log2n, comp, mode, qp, flags, coeff_off=0, coeff_cnt=0, res_scale=0):
    return (x, y, log2n, comp, mode, qp, flags, 1 << log2n, res_scale, coeff_off, coeff_cnt, 0)


def cells(w, h, qp, extra=0):
    n = ((w + 3) // 4) * ((h + 3) // 4)
    return np.full(n, N.CELL_INTRA | extra | ((qp & 0xff) << 8), dtype=np.uint32)


def dense_levels(block):
    """Coefficient words listing every position of an n x n block."""
    flat = np.asarray(block, dtype=np.int64).reshape(-1)
    pos = np.arange(len(flat), dtype=np.int64)
    return ((pos & 0x3ff) | (flat << 10)).astype(np.uint32)


def act_forward(r0, r1, r2):
    """Forward YCgCo-R lifting (the inverse of src/act.rs:43), elementwise."""
    co = r2 - r1
    t = r1 + (co >> 1)
    cg = r0 - t
    y = t + (cg >> 1)
    return y, cg, co


def act_lossless_case(seed=0, n=16, scale_cb=0, scale_cr=0):
    """A 4:4:4 8-bit picture of lossless ACT TUs of n x n (at least 8 x 8
    samples, the TUs in z-order).

    The first TU has no neighbour, so its DC predictions are the neutral 128
    and its reconstruction is 128 + the target residual triple, which the
    levels encode through cross-component prediction and the forward colour
    transform.  Returns (records, target residuals (Y, Cb, Cr) of the first
    TU)."""
    rng = np.random.default_rng(seed)
    size = max(n, 8)
    log2n = n.bit_length() - 1
    tbs, words, first = [], [], None
    off = 0
    for k in range((size // n) ** 2):
        x, y0 = (k & 1) * n, (k >> 1) * n
        r = [rng.integers(-100, 101, size=(n, n)) for _ in range(3)]
        y, cg, co = act_forward(*r)
        levels = [y, cg - ((scale_cb * y) >> 3), co - ((scale_cr * y) >> 3)]
        fl = N.TB_ACT | N.TB_TQB | N.TB_CBF
        for comp in range(3):
            words.append(dense_levels(levels[comp]))
            tbs.append(tb(x, y0, log2n, comp, 1, 30, fl, off, n * n, (0, scale_cb, scale_cr)[comp]))
            off += n * n
        first = first or r
    p = picture(size, size, 3, 8, log2_ctb=max(4, log2n))
    p.tb_count = len(tbs)
    p.coeff_count = off
    rs = H.RecordSet.from_arrays([p], np.array(tbs, dtype=H._TB_DTYPE), np.concatenate(words),
                                 cells(size, size, 30, N.CELL_TQB), [output(size, size, 3)])
    return rs, first


def pcm_case(seed=0, chroma_idc=1, bd=8, pcm_bits=5, log2_cu=4, size=32):
    """A size x size picture tiled by I_PCM CUs of 1 << log2_cu.  Returns
    (records, expected Y, Cb, Cr planes)."""
    rng = np.random.default_rng(seed)
    shift = bd - pcm_bits
    cu = 1 << log2_cu
    sw = 2 if chroma_idc in (1, 2) else 1
    sh = 2 if chroma_idc == 1 else 1
    planes = [np.zeros((size, size), np.uint16)]
    if chroma_idc:
        planes += [np.zeros((size // sh, size // sw), np.uint16) for _ in range(2)]
    tbs, words = [], []
    off = 0
    ctb = max(16, cu)
    order = [(bx + x, by + y) for by in range(0, size, ctb) for bx in range(0, size, ctb)
             for y in range(0, ctb, cu) for x in range(0, ctb, cu)]
    for cx, cy in order:
            for comp in range(3 if chroma_idc else 1):
                l2 = log2_cu if comp == 0 or chroma_idc == 3 else log2_cu - 1
                m = 1 << l2
                x, y = (cx, cy) if comp == 0 else (cx // sw, cy // sh)
                s = rng.integers(0, 1 << pcm_bits, size=(m, m))
                planes[comp][y:y + m, x:x + m] = s << shift
                words.append(dense_levels(s))
                tbs.append(tb(x, y, l2, comp, 1, shift, N.TB_PCM, off, m * m))
                off += m * m
    p = picture(size, size, chroma_idc, bd, log2_ctb=ctb.bit_length() - 1)
    p.tb_count = len(tbs)
    p.coeff_count = off
    rs = H.RecordSet.from_arrays([p], np.array(tbs, dtype=H._TB_DTYPE), np.concatenate(words),
                                 cells(size, size, 30, N.CELL_PCM),
                                 [output(size, size, chroma_idc, N.OUT_RGB16 if bd > 8 else N.OUT_RGB8)])
    return rs, planes


def inter_case(ref_luma, ref_cb=None, ref_cr=None, mvs=((0, 0),), bi=False, weights=None, pu=16, bd=8):
    """Two pictures of bit depth bd: picture 0 is I_PCM carrying the given
    planes exactly (PcmBitDepth = bd); picture 1 is tiled by pu x pu inter PUs (the i-th
    takes mvs[i % len(mvs)], list 0, and list 1 too when bi) whose CUs code
    no residual, with the loop filters off, so its reconstruction is the
    motion-compensated prediction.  weights: None, or ((w, o) per Y/Cb/Cr for
    list 0, the same for list 1) with log2 denominators 6.
    Returns the record set (outputs: one image per picture)."""
    ref_luma = np.asarray(ref_luma)
    h, w = ref_luma.shape
    chroma = ref_cb is not None
    cidc = 1 if chroma else 0
    tbs, words, off = [], [], 0
    order = [(bx + x, by + y) for by in range(0, h, 16) for bx in range(0, w, 16)
             for y in range(0, 16, 16) for x in range(0, 16, 16)]
    for cx, cy in order:  # picture 0: 16x16 PCM CUs
        for comp, plane in enumerate([ref_luma] + ([ref_cb, ref_cr] if chroma else [])):
            m = 16 if comp == 0 else 8
            x, y = (cx, cy) if comp == 0 else (cx // 2, cy // 2)
            words.append(dense_levels(np.asarray(plane)[y:y + m, x:x + m]))
            tbs.append(tb(x, y, 4 if comp == 0 else 3, comp, 1, 0, N.TB_PCM, off, m * m))
            off += m * m
    p0 = picture(w, h, cidc, bd)
    p0.tb_count, p0.coeff_count = len(tbs), off
    t1 = []
    for cx, cy in order:  # picture 1: inter TBs without residual
        t1.append(tb(cx, cy, 4, 0, 1, 30, N.TB_INTER))
        if chroma:
            for comp in (1, 2):
                t1.append(tb(cx // 2, cy // 2, 3, comp, 1, 30, N.TB_INTER))
    pus = []
    k = 0
    for y in range(0, h, pu):
        for x in range(0, w, pu):
            mv = mvs[k % len(mvs)]
            k += 1
            u = N.PU(x=x, y=y, w=pu, h=pu, flags=N.PU_L0 | (N.PU_L1 if bi else 0))
            u.ref[0], u.ref[1] = 0, 1 if bi else 0
            for l in range(2):
                u.mv[l][0], u.mv[l][1] = mv
            pus.append(u)
    refs = []
    for l in range(2 if bi else 1):
        e = N.Ref(picture=0, weighted=1 if weights else 0)
        e.log2_denom[0] = e.log2_denom[1] = 6
        for c in range(3):
            e.weight[c], e.offset[c] = weights[l][c] if weights else (64, 0)
        refs.append(e)
    p1 = picture(w, h, cidc, bd)
    p1.tb_begin, p1.tb_count = len(tbs), len(t1)
    p1.coeff_begin, p1.coeff_count = off, 0
    p1.ctb_begin = ((w + 15) // 16) * ((h + 15) // 16)
    p1.cell_begin = ((w + 3) // 4) * ((h + 3) // 4)
    p1.pu_begin, p1.pu_count, p1.ref_begin, p1.ref_count = 0, len(pus), 0, len(refs)
    p1.out_index = 1
    c0 = cells(w, h, 30, N.CELL_PCM)
    c1 = np.full_like(c0, (30 & 0xff) << 8)
    kind = N.OUT_RGB16 if bd > 8 else N.OUT_RGB8
    if not cidc:
        kind = N.OUT_LUMA16 if bd > 8 else N.OUT_LUMA8
    rs = H.RecordSet.from_arrays([p0, p1], np.array(tbs + t1, dtype=H._TB_DTYPE), np.concatenate(words),
                                 np.concatenate([c0, c1]), [output(w, h, cidc, kind), output(w, h, cidc, kind)],
                                 pus=pus, refs=refs)
    return rs


def cell_word(bsv=0, bsh=0, extra=N.CELL_INTRA, qp=30, slice_idx=0, tile=0):
    """One per-4x4 cell info word (HPVCD_CELL_PACK)."""
    return (bsv & 3) | ((bsh & 3) << 2) | extra | ((qp & 0xff) << 8) | ((slice_idx & 0xff) << 16) | \
        ((tile & 0xff) << 24)


def pcm_picture(planes, chroma_idc, bd, cell_words, log2_ctb=4, cu=8, flags=0, slices=((0, 0, 0, 1),),
                pps_cb_qp_offset=0, pps_cr_qp_offset=0, sao=None, deblock=True):
    """One picture whose CUs are all I_PCM CUs of cu x cu carrying `planes`
    exactly (PcmBitDepth = bit depth, loop-filtered), so the loop filters see
    chosen samples.  cell_words: the per-4x4 cell words (bS, QP, slice, tile)
    in raster order; slices: (disabled, beta_offset_div2, tc_offset_div2,
    lf_across) per slice; sao: one hpvcd_sao per CTB or None (all off).
    Returns (picture, tbs, coefficient words, cells, sao list)."""
    y = np.asarray(planes[0])
    h, w = y.shape
    sw = 2 if chroma_idc in (1, 2) else 1
    sh = 2 if chroma_idc == 1 else 1
    ctb = 1 << log2_ctb
    tbs, words = [], []
    off = 0
    for by in range(0, h, ctb):
        for bx in range(0, w, ctb):
            for k in range((ctb // cu) ** 2):
                # z-order of the CUs inside the CTB
                zx = sum(((k >> (2 * b)) & 1) << b for b in range(4))
                zy = sum(((k >> (2 * b + 1)) & 1) << b for b in range(4))
                cx, cy = bx + zx * cu, by + zy * cu
                if cx >= w or cy >= h:
                    continue
                for comp in range(3 if chroma_idc else 1):
                    m = cu if comp == 0 or chroma_idc == 3 else cu // 2
                    x0, y0 = (cx, cy) if comp == 0 else (cx // sw, cy // sh)
                    blk = np.asarray(planes[comp])[y0:y0 + m, x0:x0 + m]
                    words.append(dense_levels(blk))
                    tbs.append(tb(x0, y0, m.bit_length() - 1, comp, 1, 0, N.TB_PCM, off, m * m))
                    off += m * m
    p = picture(w, h, chroma_idc, bd, log2_ctb=log2_ctb, flags=flags, deblock=deblock)
    p.num_slices = len(slices)
    for i, (dis, bo, to, acr) in enumerate(slices):
        p.slices[i].disabled = 0 if deblock and not dis else 1
        p.slices[i].beta_offset_div2, p.slices[i].tc_offset_div2, p.slices[i].lf_across = bo, to, acr
    p.pps_cb_qp_offset, p.pps_cr_qp_offset = pps_cb_qp_offset, pps_cr_qp_offset
    p.tb_count, p.coeff_count = len(tbs), off
    nctb = ((w + ctb - 1) // ctb) * ((h + ctb - 1) // ctb)
    sao = list(sao) if sao is not None else [N.Sao() for _ in range(nctb)]
    return p, tbs, np.concatenate(words), np.asarray(cell_words, dtype=np.uint32), sao


def batch_of(pictures, kind=None):
    """A record set over pcm_picture() results (one output image per picture)."""
    pics, tbs, words, cells, sao, outs = [], [], [], [], [], []
    tb0 = co0 = cell0 = ctb0 = 0
    for i, (p, t, wd, c, s) in enumerate(pictures):
        p.tb_begin, p.coeff_begin, p.cell_begin, p.ctb_begin, p.out_index = tb0, co0, cell0, ctb0, i
        pics.append(p)
        tbs += t
        words.append(wd)
        cells.append(c)
        sao += s
        tb0, co0, cell0, ctb0 = tb0 + len(t), co0 + len(wd), cell0 + len(c), ctb0 + len(s)
        k = kind if kind is not None else (N.OUT_RGB16 if p.bit_depth_luma > 8 else N.OUT_RGB8)
        if not p.chroma_idc:
            k = N.OUT_LUMA16 if p.bit_depth_luma > 8 else N.OUT_LUMA8
        outs.append(output(p.width, p.height, p.chroma_idc, k))
    return H.RecordSet.from_arrays(pics, np.array(tbs, dtype=H._TB_DTYPE), np.concatenate(words),
                                   np.concatenate(cells), outs, sao=sao)


def shifted(plane, dx, dy):
    """plane sampled at (x + dx, y + dy) with edge clamping (src/mc.rs:310)."""
    plane = np.asarray(plane)
    h, w = plane.shape
    ys = np.clip(np.arange(h) + dy, 0, h - 1)
    xs = np.clip(np.arange(w) + dx, 0, w - 1)
    return plane[np.ix_(ys, xs)]


# levels of the reference's dequantisation stress tests (src/transform.rs:1374-1382)
SAFE_LEVELS = [0, 1, -1, 2, -2, 3, -3, 7, -7, 15, -15, 31, -31, 63, -63, 5]
STRESS_LEVELS = [0, 1, -1, 255, -255, 4095, -4095, 32767, -32767, 100_000, -100_000, 1_000_000, -1_000_000, 17, -33,
                 511]


def stress_case(bd=8, ep=False, stress=True, ts=False, qp=None, seed=0, chroma_idc=3):
    """A 64x64 intra picture (CTB 64, DC prediction) whose TBs carry every
    position with levels from the reference's SAFE/STRESS tables, rotated per
    TB: a 32x32, four 16x16, sixteen 8x8 and sixty-four 4x4 TBs in z-order,
    each TU luma then (4:4:4) Cb and Cr of the same size.  ts sets transform
    skip on the 4x4 TBs.  Exercises the clamp before the multiply, the
    clamps after dequantisation and both transform stages at the given bit
    depth with and without extended precision."""
    src = STRESS_LEVELS if stress else SAFE_LEVELS
    qp = (22 + 6 * (bd - 8)) if qp is None else qp
    tbs, words = [], []
    off = 0
    rot = seed
    tus = [(0, 0, 5)]
    tus += [(32 + (k & 1) * 16, (k >> 1) * 16, 4) for k in range(4)]
    tus += [(((k & 1) | ((k >> 1) & 2)) * 8, 32 + (((k >> 1) & 1) | ((k >> 2) & 2)) * 8, 3) for k in range(16)]
    tus += [(32 + sum(((k >> (2 * b)) & 1) << b for b in range(3)) * 4,
             32 + sum(((k >> (2 * b + 1)) & 1) << b for b in range(3)) * 4, 2) for k in range(64)]
    for x, y, l2 in tus:
        n = 1 << l2
        for comp in range(3 if chroma_idc == 3 else 1):
            lv = np.array([src[(i + rot) % len(src)] for i in range(n * n)], np.int64).reshape(n, n)
            rot += 5
            words.append(dense_levels(lv))
            fl = N.TB_CBF | (N.TB_TS if ts and l2 == 2 else 0)
            tbs.append(tb(x, y, l2, comp, 1, qp, fl, off, n * n))
            off += n * n
    flags = N.PF_DEBLOCK_DISABLED | (N.PF_EXTENDED_PRECISION if ep else 0)
    p = picture(64, 64, chroma_idc, bd, log2_ctb=6, flags=flags)
    p.tb_count, p.coeff_count = len(tbs), off
    kind = N.OUT_RGB16 if bd > 8 else N.OUT_RGB8
    if not chroma_idc:
        kind = N.OUT_LUMA16 if bd > 8 else N.OUT_LUMA8
    return H.RecordSet.from_arrays([p], np.array(tbs, dtype=H._TB_DTYPE), np.concatenate(words),
                                   cells(64, 64, qp - 6 * (bd - 8)), [output(64, 64, chroma_idc, kind)])
