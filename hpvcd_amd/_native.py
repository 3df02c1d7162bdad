"""ctypes mirror of include/hpvcd_records.h and include/hpvcd.h.

Loads ``hpvcd_amd/libhpvcd.so`` (built in-tree by ``hpvcd_amd.build``).  A
missing library raises ``ImportError`` with the build command; there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhpvcd.so")

# ---- records -----------------------------------------------------------------
PF_STRONG_INTRA_SMOOTHING = 1 << 0
PF_INTRA_SMOOTHING_DISABLED = 1 << 1
PF_INTRA_BOUNDARY_FILTER_OFF = 1 << 2
PF_EXTENDED_PRECISION = 1 << 3
PF_IMPLICIT_RDPCM = 1 << 4
PF_TS_ROTATION = 1 << 5
PF_SAO_LUMA = 1 << 6
PF_SAO_CHROMA = 1 << 7
PF_PCM_LF_DISABLED = 1 << 8
PF_CONSTRAINED_INTRA = 1 << 9
PF_LF_ACROSS_TILES = 1 << 10
PF_LF_ACROSS_SLICES = 1 << 11
PF_FULL_RANGE = 1 << 12
PF_TILES = 1 << 13
PF_DEBLOCK_DISABLED = 1 << 14

TB_CBF = 1 << 0
TB_TS = 1 << 1
TB_TQB = 1 << 2
TB_INTER = 1 << 3
TB_RDPCM_H = 1 << 4
TB_RDPCM_V = 1 << 5
TB_ACT = 1 << 6
TB_PCM = 1 << 7
PU_L0 = 1 << 0
PU_L1 = 1 << 1
REF_MISSING = 0xFFFFFFFF

CELL_TQB = 1 << 4
CELL_PCM = 1 << 5
CELL_INTRA = 1 << 6

OUT_RGB8 = 0
OUT_RGB16 = 1
OUT_RGB8_FROM16 = 2
OUT_LUMA8 = 3
OUT_LUMA16 = 4
OUT_YUV420_8, OUT_YUV420_16, OUT_YUV444_8, OUT_YUV444_16, OUT_YUV400_8, OUT_YUV400_16 = 5, 6, 7, 8, 9, 10
OUT_YUV = (OUT_YUV420_8, OUT_YUV420_16, OUT_YUV444_8, OUT_YUV444_16, OUT_YUV400_8, OUT_YUV400_16)

ORIENT_NORMAL, ORIENT_FLIP_H, ORIENT_ROTATE_180, ORIENT_FLIP_V = 1, 2, 3, 4
ORIENT_TRANSPOSE, ORIENT_ROTATE_90, ORIENT_TRANSVERSE, ORIENT_ROTATE_270 = 5, 6, 7, 8

# hpvcd_status codes (1..13 mirror DecodeError, src/error.rs:30-58)
OK, E_NOT_HEIF, E_TRUNCATED_BOX, E_MISSING_BOX, E_UNSUPPORTED_ITEM, E_BITSTREAM, E_CABAC_DESYNC = 0, 1, 2, 3, 4, 5, 6
E_UNSUPPORTED_CHROMA, E_UNSUPPORTED_BITDEPTH, E_BAD_DIMENSIONS, E_PARAMSET, E_UNSUPPORTED = 7, 8, 9, 10, 11
E_LIMIT, E_ALLOC, E_DEVICE, E_INVALID = 12, 13, 14, 15

SCALING_ENTRY_BYTES = 66
SCALING_SET_BYTES = 4 * 6 * SCALING_ENTRY_BYTES

MODE_BATCHED = 0
MODE_PLAIN = 1
MODE_GRAPH = 2

STAGE_RECON = 0
STAGE_DEBLOCK = 1
STAGE_SAO = 2
STAGE_RGB = 3

KERNEL_NAMES = ("residual", "intra", "deblock", "sao", "rgb", "prep", "mc")


class SliceParams(C.Structure):
    _fields_ = [("disabled", C.c_uint8), ("beta_offset_div2", C.c_int8), ("tc_offset_div2", C.c_int8),
                ("lf_across", C.c_uint8)]


class Picture(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32),
        ("chroma_idc", C.c_uint8), ("bit_depth_luma", C.c_uint8), ("bit_depth_chroma", C.c_uint8),
        ("log2_ctb", C.c_uint8),
        ("flags", C.c_uint32),
        ("pps_cb_qp_offset", C.c_int8), ("pps_cr_qp_offset", C.c_int8),
        ("matrix", C.c_uint8), ("num_slices", C.c_uint8),
        ("log2_sao_offset_scale_luma", C.c_int8), ("log2_sao_offset_scale_chroma", C.c_int8),
        ("out_kind", C.c_uint8), ("_pad0", C.c_uint8),
        ("slices", SliceParams * 8),
        ("slice_qp", C.c_int32),
        ("crop_left", C.c_int32), ("crop_top", C.c_int32), ("win_w", C.c_int32), ("win_h", C.c_int32),
        ("dst_x", C.c_int32), ("dst_y", C.c_int32), ("copy_w", C.c_int32), ("copy_h", C.c_int32),
        ("out_index", C.c_uint32),
        ("scaling_index", C.c_int32),
        ("tb_begin", C.c_uint32), ("tb_count", C.c_uint32),
        ("coeff_begin", C.c_uint32), ("coeff_count", C.c_uint32),
        ("ctb_begin", C.c_uint32),
        ("cell_begin", C.c_uint32),
        ("pu_begin", C.c_uint32), ("pu_count", C.c_uint32),
        ("ref_begin", C.c_uint32), ("ref_count", C.c_uint32),
    ]


class PU(C.Structure):
    _fields_ = [
        ("x", C.c_uint16), ("y", C.c_uint16), ("w", C.c_uint8), ("h", C.c_uint8),
        ("flags", C.c_uint8), ("_pad", C.c_uint8),
        ("ref", C.c_uint16 * 2), ("mv", (C.c_int16 * 2) * 2),
    ]


class Ref(C.Structure):
    _fields_ = [
        ("picture", C.c_uint32), ("weighted", C.c_uint8), ("log2_denom", C.c_uint8 * 2), ("_pad", C.c_uint8),
        ("weight", C.c_int16 * 3), ("offset", C.c_int16 * 3),
    ]


class TB(C.Structure):
    _fields_ = [
        ("x", C.c_uint16), ("y", C.c_uint16),
        ("log2n", C.c_uint8), ("comp", C.c_uint8), ("mode", C.c_uint8), ("qp", C.c_uint8),
        ("flags", C.c_uint16),
        ("nx", C.c_uint8), ("res_scale", C.c_int8),
        ("coeff_off", C.c_uint32), ("coeff_cnt", C.c_uint32), ("order", C.c_uint32),
    ]


class Sao(C.Structure):
    _fields_ = [("type_idx", C.c_uint8 * 3), ("band_pos", C.c_uint8 * 3), ("eo_class", C.c_uint8 * 3),
                ("_pad", C.c_uint8 * 3), ("offsets", (C.c_int16 * 4) * 3)]


class Output(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("kind", C.c_uint8), ("channels", C.c_uint8),
                ("bytes_per_sample", C.c_uint8), ("orientation", C.c_uint8)]


class Batch(C.Structure):
    _fields_ = [
        ("num_pictures", C.c_uint32), ("pictures", C.POINTER(Picture)),
        ("num_tbs", C.c_uint32), ("tbs", C.POINTER(TB)),
        ("num_coeffs", C.c_uint32), ("coeffs", C.POINTER(C.c_uint32)),
        ("num_ctbs", C.c_uint32), ("sao", C.POINTER(Sao)),
        ("num_cells", C.c_uint32), ("cell_info", C.POINTER(C.c_uint32)),
        ("num_scaling", C.c_uint32), ("scaling", C.POINTER(C.c_uint8)),
        ("num_outputs", C.c_uint32), ("outputs", C.POINTER(Output)),
        ("num_pus", C.c_uint32), ("pus", C.POINTER(PU)),
        ("num_refs", C.c_uint32), ("refs", C.POINTER(Ref)),
    ]


class ServerConfig(C.Structure):
    _fields_ = [("max_batch_pictures", C.c_uint32), ("parse_workers", C.c_uint32), ("max_batch_bytes", C.c_uint64),
                ("hbm_reserve", C.c_uint64), ("batch_delay_us", C.c_uint64)]


class Yuv(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("chroma_width", C.c_uint32),
                ("chroma_height", C.c_uint32), ("chroma_idc", C.c_uint8), ("bit_depth", C.c_uint8),
                ("sample_bytes", C.c_uint8), ("orientation", C.c_uint8), ("matrix", C.c_uint8),
                ("full_range", C.c_uint8), ("_pad", C.c_uint8 * 2), ("y", C.c_void_p), ("cb", C.c_void_p),
                ("cr", C.c_void_p), ("bytes", C.c_size_t), ("_reserved", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("residual_block", C.c_uint16 * 4), ("deblock_block", C.c_uint16), ("rgb_rows", C.c_uint16),
                ("intra_per_cu", C.c_uint8), ("sao_parts", C.c_uint8), ("_pad", C.c_uint8 * 2)]


class RequestStats(C.Structure):
    _fields_ = [("latency_ms", C.c_double), ("parse_ms", C.c_double), ("queue_ms", C.c_double),
                ("device_ms", C.c_double), ("batch_requests", C.c_uint32), ("batch_pictures", C.c_uint32)]


class GenConfig(C.Structure):
    _fields_ = [
        ("tile_width", C.c_int32), ("tile_height", C.c_int32),
        ("grid_cols", C.c_int32), ("grid_rows", C.c_int32),
        ("out_width", C.c_int32), ("out_height", C.c_int32),
        ("chroma_idc", C.c_uint8), ("bit_depth", C.c_uint8), ("log2_ctb", C.c_uint8), ("log2_min_cb", C.c_uint8),
        ("log2_min_tb", C.c_uint8), ("log2_max_tb", C.c_uint8), ("max_transform_depth_intra", C.c_uint8),
        ("qp", C.c_uint8),
        ("log2_max_ts", C.c_uint8),
        ("transform_skip", C.c_uint8), ("implicit_rdpcm", C.c_uint8), ("cross_component", C.c_uint8),
        ("extended_precision", C.c_uint8),
        ("strong_intra_smoothing", C.c_uint8), ("sao", C.c_uint8), ("deblock", C.c_uint8),
        ("ts_rotation", C.c_uint8),
        ("matrix", C.c_uint8), ("full_range", C.c_uint8), ("out_kind", C.c_uint8), ("num_outputs", C.c_uint8),
        ("log2_res_scale_max", C.c_uint8),
        ("act", C.c_uint8), ("pcm", C.c_uint8), ("pcm_bit_depth", C.c_uint8), ("pcm_lf_disabled", C.c_uint8),
        ("inter", C.c_uint8), ("num_refs", C.c_uint8), ("weighted", C.c_uint8), ("amp", C.c_uint8),
        ("gop", C.c_uint8), ("_pad", C.c_uint8),
        ("p_split_cu", C.c_float), ("p_nxn", C.c_float), ("p_split_tu", C.c_float), ("p_cbf", C.c_float),
        ("p_sig", C.c_float), ("p_level_geom", C.c_float), ("p_ts", C.c_float), ("p_tqb", C.c_float),
        ("p_act", C.c_float), ("p_pcm", C.c_float),
        ("p_intra_cu", C.c_float), ("p_skip", C.c_float), ("p_bi", C.c_float), ("mv_sigma", C.c_float),
        ("config_id", C.c_uint32),
        ("base_seed", C.c_uint64),
        ("scaling_list", C.c_uint8), ("cu_qp_delta", C.c_uint8), ("diff_cu_qp_delta_depth", C.c_uint8),
        ("qp_delta_max", C.c_uint8),
        ("pps_cb_qp_offset", C.c_int8), ("pps_cr_qp_offset", C.c_int8),
        ("slice_cb_qp_offset", C.c_int8), ("slice_cr_qp_offset", C.c_int8),
        ("beta_offset_div2", C.c_int8), ("tc_offset_div2", C.c_int8),
        ("num_slices", C.c_uint8), ("tile_cols", C.c_uint8), ("tile_rows", C.c_uint8),
        ("lf_across_tiles", C.c_uint8), ("lf_across_slices", C.c_uint8), ("random_slice_params", C.c_uint8),
        ("orientation", C.c_uint8), ("sign_hiding", C.c_uint8),
        ("conf_win", C.c_uint16 * 4),
        ("p_sao_merge", C.c_float),
        ("persistent_rice", C.c_uint8), ("ts_context", C.c_uint8),
        ("bypass_alignment", C.c_uint8), ("wpp", C.c_uint8),
        ("tmvp", C.c_uint8), ("merge_cand", C.c_uint8), ("par_merge_level", C.c_uint8),
        ("mvd_l1_zero", C.c_uint8), ("rps_in_sps", C.c_uint8), ("cabac_init", C.c_uint8),
        ("_pad2", C.c_uint8 * 2),
        ("p_merge", C.c_float),
    ]


class Settings(C.Structure):
    _fields_ = [
        ("threads", C.c_int32), ("max_dimension", C.c_uint32),
        ("max_pixels", C.c_uint64), ("max_box_size", C.c_uint64), ("max_item_size", C.c_uint64),
        ("max_items", C.c_uint32), ("max_tiles", C.c_uint32),
        ("decode_alpha", C.c_uint8), ("decode_gain_map", C.c_uint8), ("_pad", C.c_uint8 * 2),
        ("backend", C.c_int32), ("device", C.c_int32),
    ]


class Image(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32),
        ("channels", C.c_uint8), ("bit_depth", C.c_uint8), ("sample_bytes", C.c_uint8), ("orientation", C.c_uint8),
        ("pixels", C.c_void_p), ("pixels_len", C.c_size_t),
        ("matrix", C.c_uint8), ("full_range", C.c_uint8), ("_pad", C.c_uint8 * 6),
    ]


class MvCell(C.Structure):
    _fields_ = [
        ("inter", C.c_uint8), ("pred", C.c_uint8), ("lt", C.c_uint8), ("avail", C.c_uint8),
        ("ref_idx", C.c_int8 * 2), ("_pad", C.c_uint8 * 2),
        ("mv", (C.c_int16 * 2) * 2), ("ref_poc", C.c_int32 * 2),
    ]


class MotionCase(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32),
        ("log2_ctb", C.c_int32), ("par_mrg_level", C.c_int32), ("max_merge", C.c_int32),
        ("is_b", C.c_int32), ("tmvp", C.c_int32), ("cur_poc", C.c_int32), ("col_poc", C.c_int32),
        ("col_from_l0", C.c_int32),
        ("num_ref", C.c_int32 * 2),
        ("ref_poc", (C.c_int32 * 16) * 2),
        ("ref_lt", (C.c_uint8 * 16) * 2),
        ("pu_x", C.c_int32), ("pu_y", C.c_int32), ("pu_w", C.c_int32), ("pu_h", C.c_int32),
        ("cu_x", C.c_int32), ("cu_y", C.c_int32), ("cu_size", C.c_int32),
        ("part_idx", C.c_int32),
        ("cur", C.POINTER(MvCell)), ("col", C.POINTER(MvCell)),
    ]


class MotionResult(C.Structure):
    _fields_ = [
        ("num_merge", C.c_int32),
        ("merge", MvCell * 5),
        ("amvp", (((C.c_int16 * 2) * 2) * 16) * 2),
    ]


BACKEND_CPU_REF, BACKEND_GPU_PLAIN, BACKEND_GPU_BATCHED = 0, 1, 2

# every symbol include/hpvcd.h declares, with its ctypes signature
_P = C.c_void_p
SIGNATURES = {
    "hpvcd_last_error": (C.c_char_p, []),
    "hpvcd_version": (C.c_char_p, []),
    "hpvcd_gen_default_config": (None, [C.POINTER(GenConfig)]),
    "hpvcd_gen_records": (C.c_int, [C.POINTER(GenConfig), C.POINTER(_P)]),
    "hpvcd_records_concat": (C.c_int, [C.POINTER(_P), C.c_uint32, C.POINTER(_P)]),
    "hpvcd_records_view": (None, [_P, C.POINTER(Batch)]),
    "hpvcd_records_free": (None, [_P]),
    "hpvcd_gen_hevc": (C.c_int, [C.POINTER(GenConfig), C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t)]),
    "hpvcd_bytes_free": (None, [C.POINTER(C.c_uint8)]),
    "hpvcd_parse_hevc": (C.c_int, [C.c_char_p, C.c_size_t, C.c_int32, C.POINTER(_P)]),
    "hpvcd_parse_heic": (C.c_int, [C.c_char_p, C.c_size_t, C.c_void_p, C.c_int32, C.POINTER(_P)]),
    "hpvcd_gen_heic": (C.c_int, [C.POINTER(GenConfig), C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t)]),
    "hpvcd_cabac_num_elements": (C.c_int32, []),
    "hpvcd_motion_probe": (C.c_int, [C.POINTER(MotionCase), C.POINTER(MotionResult)]),
    "hpvcd_cabac_element": (C.c_int32, [C.c_int32, C.POINTER(C.c_char_p), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "hpvcd_cabac_tables": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hpvcd_cabac_init_states": (C.c_int32, [C.c_int32, C.c_int32, C.c_void_p]),
    "hpvcd_output_bytes": (C.c_size_t, [C.POINTER(Output)]),
    "hpvcd_validate_batch": (C.c_int, [C.POINTER(Batch)]),
    "hpvcd_gpu_open": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "hpvcd_gpu_close": (None, [_P]),
    "hpvcd_gpu_mem_info": (C.c_int, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "hpvcd_job_create": (C.c_int, [_P, C.POINTER(Batch), C.POINTER(_P)]),
    "hpvcd_job_free": (None, [_P]),
    "hpvcd_job_device_bytes": (C.c_size_t, [_P]),
    "hpvcd_job_host_ms": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hpvcd_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(_P)]),
    "hpvcd_host_free": (None, [_P]),
    "hpvcd_job_run": (C.c_int, [_P, C.c_int, C.c_int]),
    "hpvcd_gpu_sync": (C.c_int, [_P]),
    "hpvcd_job_download": (C.c_int, [_P, C.POINTER(_P)]),
    "hpvcd_job_download_planes": (C.c_int, [_P, C.c_uint32, _P, _P, _P]),
    "hpvcd_job_time": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hpvcd_gpu_decode_batch": (C.c_int, [_P, C.POINTER(Batch), C.POINTER(_P)]),
    "hpvcd_gpu_decode_batch_noalloc": (C.c_int, [_P, C.POINTER(Batch), C.POINTER(_P)]),
    "hpvcd_noalloc_scratch_bytes": (C.c_size_t, []),
    "hpvcd_prep_digest": (C.c_int, [C.POINTER(Batch), C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "hpvcd_settings_default": (None, [C.POINTER(Settings)]),
    "hpvcd_decoder_new": (C.c_int, [C.POINTER(Settings), C.POINTER(_P)]),
    "hpvcd_decoder_free": (None, [_P]),
    "hpvcd_decode_records": (C.c_int, [_P, C.POINTER(Batch), C.POINTER(Image)]),
    "hpvcd_image_free": (None, [C.POINTER(Image)]),
    "hpvcd_decode_heic": (C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(Image)]),
    "hpvcd_server_config_default": (None, [C.POINTER(ServerConfig)]),
    "hpvcd_server_new": (C.c_int, [C.POINTER(Settings), C.POINTER(ServerConfig), C.POINTER(_P)]),
    "hpvcd_server_submit": (C.c_int, [_P, C.c_char_p, C.c_size_t, C.c_int32, C.POINTER(C.c_uint64)]),
    "hpvcd_server_wait": (C.c_int, [_P, C.c_uint64, C.POINTER(Image), C.POINTER(RequestStats)]),
    "hpvcd_server_totals": (None, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "hpvcd_server_free": (None, [_P]),
    "hpvcd_decode_heic_rgb8": (C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(Image)]),
    "hpvcd_decode_heic_yuv": (C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(Yuv)]),
    "hpvcd_yuv_free": (None, [C.POINTER(Yuv)]),
    "hpvcd_decode_hevc": (C.c_int, [_P, C.c_char_p, C.c_size_t, C.c_int32, C.POINTER(C.POINTER(Image)),
                                    C.POINTER(C.c_uint32)]),
    "hpvcd_frames_free": (None, [C.POINTER(Image), C.c_uint32]),
    "hpvcd_tuning_default": (None, [C.POINTER(Tuning)]),
    "hpvcd_tuning_check": (C.c_int, [C.POINTER(Tuning)]),
    "hpvcd_gpu_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "hpvcd_gpu_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "hpvcd_gpu_tuning_key": (C.c_int, [_P, C.c_char_p, C.c_size_t]),
    "hpvcd_tune": (C.c_int, [_P, C.c_int, C.POINTER(Tuning), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hpvcd_tuning_save": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(Tuning)]),
    "hpvcd_tuning_load": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(Tuning)]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded native library (raises ImportError when it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m hpvcd_amd.build` "
                          "(or __graft_entry__.build()); there is no non-native fallback")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class HpvcdError(RuntimeError):
    """A non-OK hpvcd_status; ``status`` holds the code."""

    def __init__(self, status: int, message: str):
        super().__init__(f"hpvcd status {status}: {message}")
        self.status = status


def check(status: int) -> None:
    if status != 0:
        raise HpvcdError(status, lib().hpvcd_last_error().decode(errors="replace"))
