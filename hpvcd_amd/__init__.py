"""hpvcd_amd — MI355X reconstruction back end for HEVC/HEIC decoding.

Python host mirror of the C ABI in ``include/hpvcd.h``.  The reference
(awxkee/hpvcd v0.3.1) exposes ``Decoder::{decode, decode_rgb8, decode_yuv}``
(src/decoder.rs:224-234) over ``decode_heic`` (src/lib.rs:1476); its
reconstruction half runs per TU through ``ExecContext`` (src/exec.rs:33).
Here the record producer (the seeded syntax generator, standing in for the
CABAC parse) emits reconstruction records, and ``Job`` reconstructs whole
batches of pictures on one GPU through hand-written gfx950 kernels.

Errors surface as :class:`HpvcdError` carrying the ``hpvcd_status`` code
(1..13 mirror ``DecodeError``, src/error.rs:30-58).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from ._native import HpvcdError, check

__all__ = ["RecordSet", "Gpu", "Job", "PinnedBuffer", "Decoder", "Server", "HpvcdError", "WORKLOADS", "gen_config", "gen_hevc",
           "default_tuning", "save_tuning", "load_tuning",
           "gen_heic", "N"]


# --------------------------------------------------------------------------------------
# workloads: the synthetic picture sets the benchmark and the parity tests use
# --------------------------------------------------------------------------------------
WORKLOADS: Dict[str, dict] = {
    # 512x512 HEVC Main intra, 4:2:0 8-bit, CTB 64, min CU 8, QP 32 -> RGB8
    "single512": dict(tile_width=512, tile_height=512, grid_cols=1, grid_rows=1, out_width=512, out_height=512,
                      chroma_idc=1, bit_depth=8, qp=32, config_id=0, out_kind=N.OUT_RGB8),
    # 4032x3024 HEIC grid of 8x6 tiles of 512x512, 4:2:0 8-bit, QP 27 -> RGB8
    "grid4032": dict(tile_width=512, tile_height=512, grid_cols=8, grid_rows=6, out_width=4032, out_height=3024,
                     chroma_idc=1, bit_depth=8, qp=27, config_id=1, out_kind=N.OUT_RGB8),
    # 4096x2048 RExt 4:4:4 10-bit grid of 8x4 tiles, QP 22, TS on 4x4 p=0.2, implicit RDPCM,
    # cross-component prediction log2_res_scale in [-3, 3] -> RGB16
    "rext4096": dict(tile_width=512, tile_height=512, grid_cols=8, grid_rows=4, out_width=4096, out_height=2048,
                     chroma_idc=3, bit_depth=10, qp=22, config_id=2, out_kind=N.OUT_RGB16,
                     transform_skip=1, log2_max_ts=2, p_ts=0.2, implicit_rdpcm=1, cross_component=1,
                     log2_res_scale_max=3),
    # 1920x1080 4:2:0 8-bit sequence of 32 pictures: 1 I + 31 B in hierarchical GOPs of 8 with up to
    # 4 references, bi-prediction p=0.5, explicit weighted prediction, merge p=0.4 with temporal MVP,
    # MV ~ N(0, 16 px), QP 30 -> one RGB8 image per picture in display order
    "video1080": dict(tile_width=1920, tile_height=1080, grid_cols=1, grid_rows=32, out_width=1920, out_height=1080,
                      num_outputs=32, chroma_idc=1, bit_depth=8, qp=30, config_id=3, out_kind=N.OUT_RGB8,
                      inter=1, num_refs=4, weighted=1, p_bi=0.5, mv_sigma=16.0, gop=8, tmvp=1),
}


def gen_config(**overrides) -> N.GenConfig:
    """Default distributions (split 0.5, NxN 0.5, TU split 0.4, cbf 0.6, sig 0.3,
    geometric levels 0.5, SAO uniform, deblocking on) with overrides applied."""
    cfg = N.GenConfig()
    N.lib().hpvcd_gen_default_config(C.byref(cfg))
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise KeyError(f"unknown generator field {k!r}")
        setattr(cfg, k, v)
    return cfg


class RecordSet:
    """An owned set of reconstruction records (one hpvcd_records*)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        self.batch = N.Batch()
        N.lib().hpvcd_records_view(self._h, C.byref(self.batch))

    @classmethod
    def generate(cls, workload: Optional[str] = None, **overrides) -> "RecordSet":
        params = dict(WORKLOADS[workload]) if workload else {}
        params.update(overrides)
        cfg = gen_config(**params)
        h = C.c_void_p()
        check(N.lib().hpvcd_gen_records(C.byref(cfg), C.byref(h)))
        return cls(h.value)

    @classmethod
    def parse_hevc(cls, data: bytes, out_kind: int = -1) -> "RecordSet":
        """Records of an Annex B stream of intra pictures (hpvcd_parse_hevc)."""
        h = C.c_void_p()
        check(N.lib().hpvcd_parse_hevc(data, len(data), out_kind, C.byref(h)))
        return cls(h.value)

    @classmethod
    def parse_heic(cls, data: bytes, out_kind: int = -1, settings: Optional[N.Settings] = None) -> "RecordSet":
        """Records of a HEIC file's primary image (hpvcd_parse_heic)."""
        h = C.c_void_p()
        check(N.lib().hpvcd_parse_heic(data, len(data), C.byref(settings) if settings is not None else None, out_kind,
                                       C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_arrays(cls, pictures: Sequence[N.Picture], tbs: np.ndarray, coeffs: np.ndarray, cells: np.ndarray,
                    outputs: Sequence[N.Output], sao: Optional[Sequence[N.Sao]] = None,
                    scaling: Optional[np.ndarray] = None, pus: Optional[Sequence[N.PU]] = None,
                    refs: Optional[Sequence[N.Ref]] = None) -> "RecordSet":
        """A batch over caller-built record arrays (what a syntax parser emits):
        ``tbs`` in the hpvcd_tb layout, ``coeffs`` and ``cells`` as uint32 words,
        SAO parameters per CTB (all off when omitted).  The arrays are copied
        and kept alive by the returned set."""
        self = cls.__new__(cls)
        self._h = C.c_void_p()
        tbs = np.ascontiguousarray(np.asarray(tbs, dtype=_TB_DTYPE))
        coeffs = np.ascontiguousarray(np.asarray(coeffs, dtype=np.uint32))
        cells = np.ascontiguousarray(np.asarray(cells, dtype=np.uint32))
        pics = (N.Picture * len(pictures))(*pictures)
        outs = (N.Output * len(outputs))(*outputs)
        nctb = 0
        for p in pictures:
            ctb = 1 << p.log2_ctb
            nctb = max(nctb, p.ctb_begin + ((p.width + ctb - 1) // ctb) * ((p.height + ctb - 1) // ctb))
        sao_arr = (N.Sao * max(1, nctb))(*(sao or []))
        b = N.Batch()
        b.num_pictures, b.pictures = len(pictures), C.cast(pics, C.POINTER(N.Picture))
        b.num_tbs, b.tbs = len(tbs), tbs.ctypes.data_as(C.POINTER(N.TB))
        b.num_coeffs, b.coeffs = len(coeffs), coeffs.ctypes.data_as(C.POINTER(C.c_uint32))
        b.num_ctbs, b.sao = nctb, C.cast(sao_arr, C.POINTER(N.Sao))
        b.num_cells, b.cell_info = len(cells), cells.ctypes.data_as(C.POINTER(C.c_uint32))
        if scaling is not None:
            scaling = np.ascontiguousarray(np.asarray(scaling, dtype=np.uint8))
            b.num_scaling = len(scaling) // N.SCALING_SET_BYTES
            b.scaling = scaling.ctypes.data_as(C.POINTER(C.c_uint8))
        b.num_outputs, b.outputs = len(outputs), C.cast(outs, C.POINTER(N.Output))
        pu_arr = (N.PU * max(1, len(pus or [])))(*(pus or []))
        ref_arr = (N.Ref * max(1, len(refs or [])))(*(refs or []))
        b.num_pus, b.pus = len(pus or []), C.cast(pu_arr, C.POINTER(N.PU))
        b.num_refs, b.refs = len(refs or []), C.cast(ref_arr, C.POINTER(N.Ref))
        self.batch = b
        self._keep = (tbs, coeffs, cells, pics, outs, sao_arr, scaling, pu_arr, ref_arr)
        return self

    @classmethod
    def concat(cls, sets: Sequence["RecordSet"]) -> "RecordSet":
        arr = (C.c_void_p * len(sets))(*[s._h.value for s in sets])
        h = C.c_void_p()
        check(N.lib().hpvcd_records_concat(arr, len(sets), C.byref(h)))
        return cls(h.value)

    def close(self) -> None:
        if self._h:
            N.lib().hpvcd_records_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- views -------------------------------------------------------------------
    @property
    def num_pictures(self) -> int:
        return self.batch.num_pictures

    def picture(self, i: int) -> N.Picture:
        """A copy of picture record i (valid after the set is freed)."""
        return N.Picture.from_buffer_copy(self.batch.pictures[i])

    def ref(self, i: int) -> N.Ref:
        """A copy of reference-list entry i of the batch."""
        return N.Ref.from_buffer_copy(self.batch.refs[i])

    def tbs(self) -> np.ndarray:
        n = self.batch.num_tbs
        if n == 0:
            return np.zeros(0, dtype=_TB_DTYPE)
        buf = (N.TB * n).from_address(C.addressof(self.batch.tbs.contents))
        return np.frombuffer(buf, dtype=_TB_DTYPE).copy()

    def coeffs(self) -> np.ndarray:
        n = self.batch.num_coeffs
        return np.ctypeslib.as_array(self.batch.coeffs, shape=(n,)).copy() if n else np.zeros(0, np.uint32)

    def cells(self) -> np.ndarray:
        n = self.batch.num_cells
        return np.ctypeslib.as_array(self.batch.cell_info, shape=(n,)).copy() if n else np.zeros(0, np.uint32)

    _RAW = {"pictures": ("num_pictures", N.Picture), "tbs": ("num_tbs", N.TB), "coeffs": ("num_coeffs", C.c_uint32),
            "sao": ("num_ctbs", N.Sao), "cell_info": ("num_cells", C.c_uint32), "scaling": ("num_scaling", C.c_uint8),
            "outputs": ("num_outputs", N.Output), "pus": ("num_pus", N.PU), "refs": ("num_refs", N.Ref)}

    def raw(self, name: str) -> bytes:
        """The bytes of one record array (pictures, tbs, coeffs, sao,
        cell_info, scaling, outputs, pus, refs)."""
        count, ctype = self._RAW[name]
        n = getattr(self.batch, count)
        if name == "scaling":
            n *= N.SCALING_SET_BYTES
        if n == 0:
            return b""
        return C.string_at(C.addressof(getattr(self.batch, name).contents), n * C.sizeof(ctype))

    def outputs(self) -> List[N.Output]:
        """Copies of the output records."""
        return [N.Output.from_buffer_copy(self.batch.outputs[i]) for i in range(self.batch.num_outputs)]

    def output_buffers(self) -> List[np.ndarray]:
        bufs = []
        for o in self.outputs():
            if o.kind in N.OUT_YUV:
                dt = np.uint16 if (o.kind - N.OUT_YUV420_8) & 1 else np.uint8
                bufs.append(np.zeros(int(N.lib().hpvcd_output_bytes(C.byref(o))) // np.dtype(dt).itemsize, dt))
                continue
            ch = 1 if o.kind in (N.OUT_LUMA8, N.OUT_LUMA16) else 3
            dt = np.uint16 if o.kind in (N.OUT_RGB16, N.OUT_LUMA16) else np.uint8
            bufs.append(np.zeros((o.height, o.width, ch), dtype=dt))
        return bufs

    def stats(self) -> dict:
        """Coded samples S, nonzero levels, visible output pixels and output bytes."""
        s = 0
        sample_bytes = 0
        for i in range(self.num_pictures):
            p = self.picture(i)
            sw = 2 if p.chroma_idc in (1, 2) else 1
            sh = 2 if p.chroma_idc == 1 else 1
            cs = 0 if p.chroma_idc == 0 else 2 * ((p.width + sw - 1) // sw) * ((p.height + sh - 1) // sh)
            n = p.width * p.height + cs
            s += n
            sample_bytes += n * (1 if p.bit_depth_luma == 8 else 2)
        out_px = sum(o.width * o.height for o in self.outputs())
        out_bytes = sum(N.lib().hpvcd_output_bytes(C.byref(o)) for o in self.outputs())
        # motion compensation: reference samples read per list and predicted samples written
        mc_read = mc_write = 0
        for i in range(self.num_pictures):
            p = self.picture(i)
            if not p.pu_count:
                continue
            pus = (N.PU * p.pu_count).from_address(C.addressof(self.batch.pus.contents) + p.pu_begin * C.sizeof(N.PU))
            sw = 2 if p.chroma_idc in (1, 2) else 1
            sh = 2 if p.chroma_idc == 1 else 1
            b = 1 if p.bit_depth_luma == 8 else 2
            for u in pus:
                n = u.w * u.h + (2 * (u.w // sw) * (u.h // sh) if p.chroma_idc else 0)
                mc_read += n * b * bin(u.flags & 3).count("1")
                mc_write += n * b
        return dict(samples=s, sample_bytes=sample_bytes, nnz=int(self.batch.num_coeffs), out_pixels=out_px,
                    out_bytes=int(out_bytes), pictures=self.num_pictures, tbs=int(self.batch.num_tbs),
                    pus=int(self.batch.num_pus), mc_read_bytes=mc_read, mc_write_bytes=mc_write,
                    dependency_depth=self.dependency_depth())

    def dependency_depth(self) -> int:
        """Length of the longest reference chain in pictures (1 + the deepest
        picture's level): the number of dependent reconstruction waves."""
        lv = []
        for i in range(self.num_pictures):
            p = self.picture(i)
            refs = [self.batch.refs[p.ref_begin + k].picture for k in range(p.ref_count)]
            lv.append(1 + max((lv[r] for r in refs), default=-1))
        return 1 + max(lv, default=-1)


_TB_DTYPE = np.dtype([("x", "<u2"), ("y", "<u2"), ("log2n", "u1"), ("comp", "u1"), ("mode", "u1"), ("qp", "u1"),
                      ("flags", "<u2"), ("nx", "u1"), ("res_scale", "i1"), ("coeff_off", "<u4"),
                      ("coeff_cnt", "<u4"), ("order", "<u4")])
assert _TB_DTYPE.itemsize == C.sizeof(N.TB)


def _take_bytes(p, n) -> bytes:
    try:
        return C.string_at(p, n.value)
    finally:
        N.lib().hpvcd_bytes_free(p)


def gen_hevc(workload: Optional[str] = None, **overrides) -> bytes:
    """The generator's pictures as an Annex B HEVC stream (hpvcd_gen_hevc)."""
    params = dict(WORKLOADS[workload]) if workload else {}
    params.update(overrides)
    cfg = gen_config(**params)
    p = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    check(N.lib().hpvcd_gen_hevc(C.byref(cfg), C.byref(p), C.byref(n)))
    return _take_bytes(p, n)


def gen_heic(workload: Optional[str] = None, **overrides) -> bytes:
    """The generator's pictures as a HEIC file (hpvcd_gen_heic)."""
    params = dict(WORKLOADS[workload]) if workload else {}
    params.update(overrides)
    cfg = gen_config(**params)
    p = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    check(N.lib().hpvcd_gen_heic(C.byref(cfg), C.byref(p), C.byref(n)))
    return _take_bytes(p, n)


class PinnedBuffer:
    """Page-locked host memory (hpvcd_host_alloc) viewed as a numpy array."""

    def __init__(self, shape, dtype):
        dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * dtype.itemsize
        p = C.c_void_p()
        check(N.lib().hpvcd_host_alloc(n, C.byref(p)))
        self._p = p
        self.array = np.ctypeslib.as_array((C.c_uint8 * max(n, 1)).from_address(p.value))[:n].view(dtype).reshape(shape)

    def close(self) -> None:
        if self._p:
            self.array = None
            N.lib().hpvcd_host_free(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Gpu:
    """One GPU and its stream (hpvcd_gpu_open)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(N.lib().hpvcd_gpu_open(device, C.byref(h)))
        self._h = h

    def mem_info(self) -> Tuple[int, int]:
        f, t = C.c_size_t(), C.c_size_t()
        check(N.lib().hpvcd_gpu_mem_info(self._h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def sync(self) -> None:
        check(N.lib().hpvcd_gpu_sync(self._h))

    def close(self) -> None:
        if self._h:
            N.lib().hpvcd_gpu_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def job(self, records: RecordSet) -> "Job":
        return Job(self, records)

    @property
    def tuning(self) -> dict:
        """The launch shapes this GPU's launches use (hpvcd_gpu_get_tuning)."""
        t = N.Tuning()
        check(N.lib().hpvcd_gpu_get_tuning(self._h, C.byref(t)))
        return tuning_dict(t)

    @tuning.setter
    def tuning(self, d: dict) -> None:
        check(N.lib().hpvcd_gpu_set_tuning(self._h, C.byref(tuning_struct(d))))

    @property
    def tuning_key(self) -> str:
        b = C.create_string_buffer(64)
        check(N.lib().hpvcd_gpu_tuning_key(self._h, b, 64))
        return b.value.decode()

    def decode(self, records: RecordSet) -> List[np.ndarray]:
        """Upload, reconstruct and download every output image of the batch."""
        with Job(self, records) as j:
            j.run()
            return j.download()


class Job:
    """A batch resident in HBM (hpvcd_job_create)."""

    def __init__(self, gpu: Gpu, records: RecordSet):
        self.gpu = gpu
        self.records = records
        h = C.c_void_p()
        check(N.lib().hpvcd_job_create(gpu._h, C.byref(records.batch), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h:
            N.lib().hpvcd_job_free(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def device_bytes(self) -> int:
        return int(N.lib().hpvcd_job_device_bytes(self._h))

    def host_ms(self) -> Tuple[float, float]:
        """(record preparation, upload) milliseconds spent in job creation."""
        a, b = C.c_double(), C.c_double()
        check(N.lib().hpvcd_job_host_ms(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def download_into(self, buffers: Sequence[np.ndarray]) -> None:
        """Copy the output images into caller buffers (e.g. PinnedBuffer arrays)."""
        ptrs = (C.c_void_p * len(buffers))(*[b.ctypes.data for b in buffers])
        check(N.lib().hpvcd_job_download(self._h, ptrs))

    def run(self, mode: int = N.MODE_BATCHED, stop: int = N.STAGE_RGB) -> None:
        check(N.lib().hpvcd_job_run(self._h, mode, stop))

    def tune(self, iters: int = 10) -> Tuple[dict, float, float]:
        """The autotuner on this job (hpvcd_tune): (best shapes, their ms per
        run, ms per run of the shapes it started from); the GPU keeps the best."""
        t, b, d = N.Tuning(), C.c_double(), C.c_double()
        check(N.lib().hpvcd_tune(self._h, iters, C.byref(t), C.byref(b), C.byref(d)))
        return tuning_dict(t), b.value, d.value

    def download(self) -> List[np.ndarray]:
        bufs = self.records.output_buffers()
        ptrs = (C.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
        check(N.lib().hpvcd_job_download(self._h, ptrs))
        return bufs

    def download_planes(self, picture: int) -> Tuple[np.ndarray, Optional[np.ndarray], Optional[np.ndarray]]:
        p = self.records.picture(picture)
        y = np.zeros((p.height, p.width), np.uint16)
        if p.chroma_idc:
            sw = 2 if p.chroma_idc in (1, 2) else 1
            sh = 2 if p.chroma_idc == 1 else 1
            cb = np.zeros(((p.height + sh - 1) // sh, (p.width + sw - 1) // sw), np.uint16)
            cr = np.zeros_like(cb)
            check(N.lib().hpvcd_job_download_planes(self._h, picture, y.ctypes.data, cb.ctypes.data, cr.ctypes.data))
            return y, cb, cr
        check(N.lib().hpvcd_job_download_planes(self._h, picture, y.ctypes.data, None, None))
        return y, None, None

    def time(self, mode: int = N.MODE_GRAPH, warmup: int = 2, iters: int = 10, breakdown: bool = True):
        ms = C.c_double()
        km = (C.c_double * len(N.KERNEL_NAMES))()
        check(N.lib().hpvcd_job_time(self._h, mode, warmup, iters, C.byref(ms), km if breakdown else None))
        return ms.value, {N.KERNEL_NAMES[i]: km[i] for i in range(len(N.KERNEL_NAMES))} if breakdown else None


def raw_array(img: N.Image) -> np.ndarray:
    """A copy of a raw-plane image's bytes as a flat array (planar Y, Cb, Cr)."""
    dt = np.uint16 if img.sample_bytes == 2 else np.uint8
    raw = (C.c_uint8 * img.pixels_len).from_address(img.pixels)
    return np.frombuffer(raw, dtype=dt).copy()


def yuv_planes(buf: np.ndarray, output) -> Tuple[np.ndarray, Optional[np.ndarray], Optional[np.ndarray]]:
    """(y, cb, cr) views of a raw-plane output buffer (HPVCD_OUT_YUV* layout)."""
    w, h = output.width, output.height
    chroma = {N.OUT_YUV420_8: 1, N.OUT_YUV420_16: 1, N.OUT_YUV444_8: 3, N.OUT_YUV444_16: 3}.get(output.kind, 0)
    y = buf[:w * h].reshape(h, w)
    if not chroma:
        return y, None, None
    s = 2 if chroma == 1 else 1
    cw, ch = (w + s - 1) // s, (h + s - 1) // s
    cb = buf[w * h:w * h + cw * ch].reshape(ch, cw)
    cr = buf[w * h + cw * ch:w * h + 2 * cw * ch].reshape(ch, cw)
    return y, cb, cr


def image_array(img: N.Image) -> np.ndarray:
    """A copy of a decoded image's pixels as (height, width, channels)."""
    dt = np.uint16 if img.sample_bytes == 2 else np.uint8
    if not img.pixels:
        raise ValueError("image has no pixels")
    raw = (C.c_uint8 * img.pixels_len).from_address(img.pixels)
    return np.frombuffer(raw, dtype=dt).reshape(img.height, img.width, img.channels).copy()


class Decoder:
    """The drop-in decoder (hpvcd_decoder_new / src/decoder.rs Decoder)."""

    def __init__(self, backend: int = N.BACKEND_GPU_BATCHED, device: int = 0, threads: int = 0, **limits):
        s = N.Settings()
        N.lib().hpvcd_settings_default(C.byref(s))
        s.backend, s.device, s.threads = backend, device, threads
        for k, v in limits.items():
            if not hasattr(s, k):
                raise KeyError(f"unknown setting {k!r}")
            setattr(s, k, v)
        h = C.c_void_p()
        check(N.lib().hpvcd_decoder_new(C.byref(s), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h:
            N.lib().hpvcd_decoder_free(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def decode(self, data: bytes) -> np.ndarray:
        """A HEIC file's image (hpvcd_decode_heic / Decoder::decode)."""
        return self._heic(N.lib().hpvcd_decode_heic, data)

    def decode_rgb8(self, data: bytes) -> np.ndarray:
        """A HEIC file's image as 8-bit RGB (hpvcd_decode_heic_rgb8 / Decoder::decode_rgb8)."""
        return self._heic(N.lib().hpvcd_decode_heic_rgb8, data)

    def decode_yuv(self, data: bytes) -> Tuple[np.ndarray, Optional[np.ndarray], Optional[np.ndarray], dict]:
        """A HEIC file's raw planes (hpvcd_decode_heic_yuv / Decoder::decode_yuv):
        (y, cb, cr, info); cb and cr are None for 4:0:0; orientation reported, not applied."""
        yv = N.Yuv()
        st = N.lib().hpvcd_decode_heic_yuv(self._h, data, len(data), C.byref(yv))
        try:
            check(st)
            dt = np.uint16 if yv.sample_bytes == 2 else np.uint8

            def plane(ptr, w, h):
                raw = (C.c_uint8 * (w * h * yv.sample_bytes)).from_address(ptr)
                return np.frombuffer(raw, dtype=dt).reshape(h, w).copy()

            y = plane(yv.y, yv.width, yv.height)
            cb = plane(yv.cb, yv.chroma_width, yv.chroma_height) if yv.cb else None
            cr = plane(yv.cr, yv.chroma_width, yv.chroma_height) if yv.cr else None
            info = {f: getattr(yv, f) for f in ("chroma_idc", "bit_depth", "orientation", "matrix", "full_range")}
            return y, cb, cr, info
        finally:
            N.lib().hpvcd_yuv_free(C.byref(yv))

    def decode_hevc(self, data: bytes, out_kind: int = -1) -> List[np.ndarray]:
        """An HEVC stream's frames in output order (hpvcd_decode_hevc / decode_hevc)."""
        frames = C.POINTER(N.Image)()
        n = C.c_uint32()
        check(N.lib().hpvcd_decode_hevc(self._h, data, len(data), out_kind, C.byref(frames), C.byref(n)))
        try:
            return [image_array(frames[i]) for i in range(n.value)]
        finally:
            N.lib().hpvcd_frames_free(frames, n)

    def _heic(self, fn, data: bytes) -> np.ndarray:
        img = N.Image()
        st = fn(self._h, data, len(data), C.byref(img))
        try:
            check(st)
            return image_array(img)
        finally:
            N.lib().hpvcd_image_free(C.byref(img))

    def decode_records(self, records: RecordSet) -> List[np.ndarray]:
        """Every output image of a record set (hpvcd_decode_records)."""
        n = records.batch.num_outputs
        imgs = (N.Image * max(n, 1))()
        st = N.lib().hpvcd_decode_records(self._h, C.byref(records.batch), imgs)
        try:
            check(st)
            outs = records.outputs()
            return [raw_array(imgs[i]) if outs[i].kind in N.OUT_YUV else image_array(imgs[i]) for i in range(n)]
        finally:
            for i in range(n):
                N.lib().hpvcd_image_free(C.byref(imgs[i]))


class Server:
    """The serving loop (hpvcd_server_*): submit HEIC bytes from any thread,
    wait for each image; requests are parsed on host workers and batched
    onto the GPU continuously."""

    def __init__(self, backend: int = N.BACKEND_GPU_BATCHED, device: int = 0, threads: int = 0,
                 max_batch_pictures: Optional[int] = None, max_batch_bytes: Optional[int] = None,
                 hbm_reserve: Optional[int] = None, parse_workers: Optional[int] = None,
                 batch_delay_us: Optional[int] = None, **limits):
        s = N.Settings()
        N.lib().hpvcd_settings_default(C.byref(s))
        s.backend, s.device, s.threads = backend, device, threads
        for k, v in limits.items():
            if not hasattr(s, k):
                raise KeyError(f"unknown setting {k!r}")
            setattr(s, k, v)
        c = N.ServerConfig()
        N.lib().hpvcd_server_config_default(C.byref(c))
        for k, v in (("max_batch_pictures", max_batch_pictures), ("max_batch_bytes", max_batch_bytes),
                     ("hbm_reserve", hbm_reserve), ("parse_workers", parse_workers),
                     ("batch_delay_us", batch_delay_us)):
            if v is not None:
                setattr(c, k, v)
        h = C.c_void_p()
        check(N.lib().hpvcd_server_new(C.byref(s), C.byref(c), C.byref(h)))
        self._h = h

    def submit(self, data: bytes, rgb8: bool = False) -> int:
        t = C.c_uint64()
        check(N.lib().hpvcd_server_submit(self._h, data, len(data), 1 if rgb8 else 0, C.byref(t)))
        return t.value

    def wait(self, ticket: int, image: bool = True):
        """(image as (height, width, channels), stats dict) of a ticket;
        image=False drops the pixels unread and gives (None, stats)."""
        img = N.Image()
        st = N.RequestStats()
        rc = N.lib().hpvcd_server_wait(self._h, ticket, C.byref(img), C.byref(st))
        try:
            check(rc)
            return (image_array(img) if image else None), {f: getattr(st, f) for f, _ in N.RequestStats._fields_}
        finally:
            N.lib().hpvcd_image_free(C.byref(img))

    def totals(self) -> Tuple[int, int, int]:
        b, r, p = C.c_uint64(), C.c_uint64(), C.c_uint64()
        N.lib().hpvcd_server_totals(self._h, C.byref(b), C.byref(r), C.byref(p))
        return b.value, r.value, p.value

    def close(self) -> None:
        if self._h:
            N.lib().hpvcd_server_free(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def tuning_dict(t: N.Tuning) -> dict:
    return {"residual_block": list(t.residual_block), "deblock_block": t.deblock_block, "rgb_rows": t.rgb_rows,
            "intra_per_cu": t.intra_per_cu, "sao_parts": t.sao_parts}


def tuning_struct(d: dict) -> N.Tuning:
    t = N.Tuning()
    N.lib().hpvcd_tuning_default(C.byref(t))
    for k, v in d.items():
        if k == "residual_block":
            for i, x in enumerate(v):
                t.residual_block[i] = x
        elif k in ("deblock_block", "rgb_rows", "intra_per_cu", "sao_parts"):
            setattr(t, k, v)
        else:
            raise KeyError(f"unknown tuning field {k!r}")
    return t


def default_tuning() -> dict:
    t = N.Tuning()
    N.lib().hpvcd_tuning_default(C.byref(t))
    return tuning_dict(t)


def save_tuning(path: str, key: str, d: dict) -> None:
    """Write (or replace) the key's line of a tuning cache file (hpvcd_tuning_save)."""
    check(N.lib().hpvcd_tuning_save(path.encode(), key.encode(), C.byref(tuning_struct(d))))


def load_tuning(path: str, key: str) -> dict:
    t = N.Tuning()
    check(N.lib().hpvcd_tuning_load(path.encode(), key.encode(), C.byref(t)))
    return tuning_dict(t)
