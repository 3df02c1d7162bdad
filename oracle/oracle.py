"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker.  It reads the same
packed records (hpvcd_records.h) as the GPU back end.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import List, Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

STAGE_RECON, STAGE_DEBLOCK, STAGE_SAO = 0, 1, 2

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        I32P = C.POINTER(C.c_int32)
        L.oracle_dequant.argtypes = [I32P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, I32P]
        L.oracle_dequant_ts.argtypes = [I32P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, I32P]
        L.oracle_inv_dct.argtypes = [I32P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, I32P]
        L.oracle_inv_dst4.argtypes = [I32P, C.c_int, C.c_int, C.c_int, I32P]
        L.oracle_inv_dct_dense.argtypes = [I32P, C.c_int, C.c_int, C.c_int, I32P]
        L.oracle_dct_coef.argtypes = [C.c_int, C.c_int, C.c_int]
        L.oracle_dct_coef.restype = C.c_int
        L.oracle_max_range.argtypes = [C.c_int, C.c_int]
        L.oracle_max_range.restype = C.c_int
        L.oracle_intra_predict.argtypes = [C.c_int, I32P, I32P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_int, C.POINTER(C.c_uint16)]
        L.oracle_intra_filter_refs.argtypes = [I32P, I32P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                               I32P, I32P]
        L.oracle_sao_ctb.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.oracle_deblock_luma_segment.argtypes = [I32P, C.c_int, C.c_int, C.c_int]
        L.oracle_luma_decision.argtypes = [I32P, C.c_int, C.c_int, I32P, I32P]
        L.oracle_luma_decision.restype = C.c_int
        L.oracle_qpc.argtypes = [C.c_int, C.c_int, C.c_int]
        L.oracle_qpc.restype = C.c_int
        L.oracle_weighted_uni.argtypes = [C.c_int] * 5
        L.oracle_weighted_uni.restype = C.c_int
        L.oracle_weighted_bi.argtypes = [C.c_int] * 8
        L.oracle_weighted_bi.restype = C.c_int
        L.oracle_act_inverse.argtypes = [I32P, I32P, I32P]
        L.oracle_picture_planes.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_picture_rgb.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_decode_batch.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, C.c_int]
        _lib = L
    return _lib


def _i32(a) -> Tuple[np.ndarray, C.POINTER(C.c_int32)]:
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int32))


# ---- kernel-level helpers -------------------------------------------------------------
def dequant(levels, n, qp, bd, extended=False, scaling: Optional[np.ndarray] = None, ts=False) -> np.ndarray:
    lv, lp = _i32(np.asarray(levels).reshape(-1))
    out = np.zeros(n * n, np.int32)
    sc = None
    if scaling is not None:
        sc = np.ascontiguousarray(scaling, dtype=np.uint8)
    fn = lib().oracle_dequant_ts if ts else lib().oracle_dequant
    fn(lp, n, qp, bd, int(extended), sc.ctypes.data if sc is not None else None,
       out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out.reshape(n, n)


def inv_dct(coeff, n, bd, max_range=15, nx=None, i16=False) -> np.ndarray:
    cv, cp = _i32(np.asarray(coeff).reshape(-1))
    out = np.zeros(n * n, np.int32)
    lib().oracle_inv_dct(cp, n, bd, max_range, n if nx is None else nx, int(i16), out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out.reshape(n, n)


def inv_dct_dense(coeff, n, bd, max_range=15) -> np.ndarray:
    cv, cp = _i32(np.asarray(coeff).reshape(-1))
    out = np.zeros(n * n, np.int32)
    lib().oracle_inv_dct_dense(cp, n, bd, max_range, out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out.reshape(n, n)


def inv_dst4(coeff, bd, max_range=15, i16=False) -> np.ndarray:
    cv, cp = _i32(np.asarray(coeff).reshape(-1))
    out = np.zeros(16, np.int32)
    lib().oracle_inv_dst4(cp, bd, max_range, int(i16), out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out.reshape(4, 4)


def dct_matrix(n) -> np.ndarray:
    return np.array([[lib().oracle_dct_coef(n, k, m) for m in range(n)] for k in range(n)], dtype=np.int64)


def intra_predict(corner, above, left, n, mode, bd=8, filter_is_luma=True, smoothing=True, strong=True,
                  edge_filters=True) -> np.ndarray:
    a, ap = _i32(above)
    l, lp = _i32(left)
    out = np.zeros(n * n, np.uint16)
    lib().oracle_intra_predict(int(corner), ap, lp, n, mode, int(filter_is_luma), int(smoothing), int(strong),
                               int(edge_filters), bd, out.ctypes.data_as(C.POINTER(C.c_uint16)))
    return out.reshape(n, n)


def intra_filter_refs(above, left, n, mode, is_luma=True, enabled=True, strong=True, bd=8):
    a, ap = _i32(above)
    l, lp = _i32(left)
    fa = np.zeros(2 * n + 1, np.int32)
    fl = np.zeros(2 * n + 1, np.int32)
    lib().oracle_intra_filter_refs(ap, lp, n, mode, int(is_luma), int(enabled), int(strong), bd,
                                   fa.ctypes.data_as(C.POINTER(C.c_int32)), fl.ctypes.data_as(C.POINTER(C.c_int32)))
    return fa, fl


def sao_ctb(src, x0, y0, x1, y1, sao_type, offsets, band_pos=0, eo_class=0, bd=8) -> np.ndarray:
    src = np.ascontiguousarray(src, dtype=np.uint16)
    dst = src.copy()
    off = np.ascontiguousarray(offsets, dtype=np.int16)
    h, w = src.shape
    lib().oracle_sao_ctb(dst.ctypes.data, src.ctypes.data, w, h, x0, y0, x1, y1, sao_type, off.ctypes.data,
                         band_pos, eo_class, bd)
    return dst


def picture_rgb(batch, p, y, cb, cr) -> np.ndarray:
    pic = batch.pictures[p]
    o = batch.outputs[pic.out_index]
    ch = 1 if o.kind in (3, 4) else 3
    dt = np.uint16 if o.kind in (1, 4) else np.uint8
    out = np.zeros((pic.win_h, pic.win_w, ch), dt)
    y = np.ascontiguousarray(y, np.uint16)
    cb = np.ascontiguousarray(cb if cb is not None else np.zeros(1), np.uint16)
    cr = np.ascontiguousarray(cr if cr is not None else np.zeros(1), np.uint16)
    rc = lib().oracle_picture_rgb(C.addressof(batch), p, y.ctypes.data, cb.ctypes.data, cr.ctypes.data, out.ctypes.data)
    if rc:
        raise RuntimeError("oracle_picture_rgb failed")
    return out


def deblock_luma_segment(lines, beta, tc, max_val) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(lines, dtype=np.int32).reshape(4, 8)).copy()
    lib().oracle_deblock_luma_segment(a.ctypes.data_as(C.POINTER(C.c_int32)), beta, tc, max_val)
    return a


def luma_decision(lines, beta, tc):
    a = np.ascontiguousarray(np.asarray(lines, dtype=np.int32).reshape(4, 8))
    p1 = C.c_int32()
    q1 = C.c_int32()
    d = lib().oracle_luma_decision(a.ctypes.data_as(C.POINTER(C.c_int32)), beta, tc, C.byref(p1), C.byref(q1))
    return d, p1.value, q1.value


# ---- picture-level helpers ------------------------------------------------------------
def picture_planes(batch, p: int, stage: int = STAGE_SAO):
    """(y, cb, cr) u16 planes of picture p after `stage` (batch: hpvcd_amd Batch struct)."""
    pic = batch.pictures[p]
    y = np.zeros((pic.height, pic.width), np.uint16)
    if pic.chroma_idc:
        sw = 2 if pic.chroma_idc in (1, 2) else 1
        sh = 2 if pic.chroma_idc == 1 else 1
        cb = np.zeros(((pic.height + sh - 1) // sh, (pic.width + sw - 1) // sw), np.uint16)
        cr = np.zeros_like(cb)
    else:
        cb = np.zeros(1, np.uint16)
        cr = np.zeros(1, np.uint16)
    rc = lib().oracle_picture_planes(C.addressof(batch), p, stage, y.ctypes.data, cb.ctypes.data, cr.ctypes.data)
    if rc:
        raise RuntimeError(f"oracle_picture_planes failed ({rc})")
    return y, (cb if pic.chroma_idc else None), (cr if pic.chroma_idc else None)


def decode_batch(batch, buffers: List[np.ndarray], threads: int = 1, rows_parallel: bool = True) -> None:
    """Decode every picture into the caller's zeroed output buffers."""
    ptrs = (C.c_void_p * len(buffers))(*[b.ctypes.data for b in buffers])
    rc = lib().oracle_decode_batch(C.addressof(batch), ptrs, threads, int(rows_parallel))
    if rc:
        raise RuntimeError(f"oracle_decode_batch failed ({rc})")
