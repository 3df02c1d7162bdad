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
    return N.Output(width=w, height=h, kind=kind, channels=3 if chroma_idc else 1,
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
