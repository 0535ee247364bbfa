"""ctypes mirror of include/rivet_hip.h (librivet_hip.so) and include/rivet_gen.h
(librivet_gen.so).

The library is loaded from this directory (built in-tree by build.py);
there is no fallback: a missing library is an ImportError naming the
build step.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "librivet_hip.so"
GEN_LIB_PATH = Path(__file__).resolve().parent / "librivet_gen.so"

RV_OK = 0
RV_E_INVALID = -1
RV_E_NODEVICE = -2
RV_E_HIP = -3
RV_E_NOMEM = -4
RV_E_RANGE = -5
RV_E_STATE = -6
RV_E_DEVICE = -7
RV_MAX_REFS = 4
RV_TU_DST = 1
RV_TU_CODED = 2
RV_TU_LEVELS = 4
RV_TU_INTRA = 8
RV_BLK_LEVELS = 1
RV_INTRA_HOST = 1
RV_INTRA_DEVICE = 2
RV_CU_LF_EXEMPT = 1
RV_CU_BYPASS = 2
RV_CU_DBK_EXEMPT = 4
RV_CODEC_HEVC = 0
RV_CODEC_H264 = 1
RV_H264_DEFAULT = 0
RV_H264_EXPLICIT = 1
RV_H264_IMPLICIT = 2
RV_H264_I4X4 = 0
RV_H264_I8X8 = 1
RV_H264_I16X16 = 2
RV_LAUNCH_PLAIN = 0
RV_LAUNCH_BATCHED = 1
RV_LAUNCH_GRAPH = 2
RV_HASH_MD5 = 0
RV_HASH_CRC = 1
RV_HASH_CHECKSUM = 2

i8, u8, i16, u16, i32, u32, i64 = C.c_int8, C.c_uint8, C.c_int16, C.c_uint16, C.c_int32, C.c_uint32, C.c_int64
R = RV_MAX_REFS


class Mv(C.Structure):
    _fields_ = [("x", i16), ("y", i16)]


class Sao(C.Structure):
    _fields_ = [("type_idx", u8), ("band_or_class", u8), ("offsets", i16 * 4)]


class HevcCu(C.Structure):
    _fields_ = [("x", u16), ("y", u16), ("log2_size", u8), ("qp_y", i8), ("intra", u8), ("part_mode", u8),
                ("pred_off", u32), ("luma_mode", u8 * 4), ("chroma_mode", u8), ("exempt", u8), ("pad", u8 * 2)]


class HevcPu(C.Structure):
    _fields_ = [("x", u16), ("y", u16), ("w", u8), ("h", u8), ("dir", u8), ("pad0", u8), ("ref_idx", i8 * 2),
                ("pad1", u16), ("mv", Mv * 2)]


class Nz(C.Structure):
    _fields_ = [("pos", u16), ("val", i16)]


class HevcTu(C.Structure):
    _fields_ = [("x", u16), ("y", u16), ("comp", u8), ("log2", u8), ("flags", u8), ("max_x", u8), ("max_y", u8),
                ("qp", u8), ("n_nz", u16), ("coef_off", u32)]


class HevcScaling(C.Structure):
    _fields_ = [("list4", (u8 * 16) * 6), ("list8", ((u8 * 64) * 6) * 3), ("dc", (u8 * 6) * 2)]


class HevcSlice(C.Structure):
    _fields_ = [("deblocking_disabled", i32), ("beta_offset", i32), ("tc_offset", i32),
                ("loop_filter_across_slices", i32)]


class HevcPic(C.Structure):
    _fields_ = [
        ("width", i32), ("height", i32), ("bit_depth", i32), ("log2_ctb", i32), ("poc", i32),
        ("num_ref", i32 * 2), ("ref_slot", (i32 * R) * 2), ("ref_poc", (i32 * R) * 2),
        ("ref_long_term", (i32 * R) * 2),
        ("weighted", i32), ("log2_wd", i32 * 3), ("wp_w", ((i32 * 3) * R) * 2), ("wp_o", ((i32 * 3) * R) * 2),
        ("deblock", i32), ("beta_offset", i32), ("tc_offset", i32), ("cb_qp_offset", i32), ("cr_qp_offset", i32),
        ("sao_luma", i32), ("sao_chroma", i32),
        ("n_cus", i32), ("n_pus", i32), ("n_tus", i32),
        ("cus", C.POINTER(HevcCu)), ("pus", C.POINTER(HevcPu)), ("tus", C.POINTER(HevcTu)),
        ("nz", C.POINTER(Nz)), ("n_nz", u32), ("n_intra_pred", u32),
        ("intra_pred", C.POINTER(u16)), ("sao", C.POINTER(Sao)), ("scaling", C.POINTER(HevcScaling)),
        ("strong_intra_smoothing", i32), ("constrained_intra_pred", i32),
        ("n_slices", i32), ("slices", C.POINTER(HevcSlice)), ("ctb_slice", C.POINTER(u16)),
        ("ctb_slice_addr", C.POINTER(u32)), ("ctb_tile", C.POINTER(u16)), ("ctb_ts", C.POINTER(u32)),
        ("loop_filter_across_tiles", i32),
    ]


class H264Mb(C.Structure):
    _fields_ = [("intra", u8), ("transform_8x8", u8), ("qp", i8), ("qpc", i8 * 2), ("pad", u8), ("nz_mask", u16),
                ("pred_off", u32), ("intra_kind", u8), ("chroma_mode", u8), ("pad2", u8 * 2), ("modes", u8 * 16)]


class H264Part(C.Structure):
    _fields_ = [("x", u16), ("y", u16), ("w", u8), ("h", u8), ("dir", u8), ("pad0", u8), ("ref_idx", i8 * 2),
                ("pad1", u16), ("mv", Mv * 2)]


class H264Blk(C.Structure):
    _fields_ = [("x", u16), ("y", u16), ("comp", u8), ("size", u8), ("has_dc", u8), ("flags", u8), ("dc", i32),
                ("coef_off", u32)]


class H264Scaling(C.Structure):
    _fields_ = [("list4", (u8 * 16) * 6), ("list8", (u8 * 64) * 2)]


class H264Pic(C.Structure):
    _fields_ = [
        ("width", i32), ("height", i32), ("bit_depth", i32), ("poc", i32),
        ("num_ref", i32 * 2), ("ref_slot", (i32 * R) * 2), ("ref_poc", (i32 * R) * 2),
        ("ref_long_term", (i32 * R) * 2),
        ("weighting", i32), ("log_wd", i32 * 3), ("wp_w", ((i32 * 3) * R) * 2), ("wp_o", ((i32 * 3) * R) * 2),
        ("deblock", i32), ("offset_a", i32), ("offset_b", i32),
        ("n_mbs", i32), ("n_parts", i32), ("n_blks", i32),
        ("mbs", C.POINTER(H264Mb)), ("parts", C.POINTER(H264Part)), ("blks", C.POINTER(H264Blk)),
        ("coefs", C.POINTER(i32)), ("n_coefs", u32), ("n_intra_pred", u32), ("intra_pred", C.POINTER(u16)),
        ("scaling", C.POINTER(H264Scaling)), ("constrained_intra_pred", i32),
    ]


class GenHevcCfg(C.Structure):
    _fields_ = [(n, i32) for n in ("width", "height", "bit_depth", "b_picture", "bi_percent", "explicit_wp",
                                   "intra_percent", "nz_permille", "lap_b", "qp_min", "qp_max", "mv_range",
                                   "deblock", "sao", "poc")] + [("ref_poc", i32 * 2), ("ref_slot", i32 * 2),
                                                                ("levels", i32), ("scaling", i32),
                                                                ("intra_device", i32),
                                                                ("strong_intra_smoothing", i32),
                                                                ("constrained_intra_pred", i32)] + [
        (n, i32) for n in ("tile_cols", "tile_rows", "n_slices", "lf_across_slices", "lf_across_tiles",
                           "slice_vary", "exempt_percent")]


class GenH264Cfg(C.Structure):
    _fields_ = [(n, i32) for n in ("width", "height", "b_picture", "bi_percent", "weighting", "intra_percent",
                                   "nz_permille", "lap_b", "qp_min", "qp_max", "mv_range", "deblock", "poc")] + [
        ("ref_poc", i32 * 2), ("ref_slot", i32 * 2), ("levels", i32), ("scaling", i32), ("intra_device", i32),
        ("constrained_intra_pred", i32)]


ABI_STRUCTS = [HevcCu, HevcPu, HevcTu, Sao, HevcPic, H264Mb, H264Part, H264Blk, H264Pic, Nz, HevcScaling,
               H264Scaling, HevcSlice]

# Every extern "C" entry of include/rivet_hip.h (librivet_hip.so):
# name -> (restype, argtypes).
P = C.c_void_p
PROTOTYPES = {
    "rv_init": (i32, [i32]),
    "rv_tier_name": (C.c_char_p, []),
    "rv_last_error": (C.c_char_p, []),
    "rv_device_count": (i32, []),
    "rv_session_create": (i32, [i32, i32, i32, i32, i32, C.POINTER(P)]),
    "rv_session_destroy": (i32, [P]),
    "rv_session_footprint": (i64, [i32, i32, i32, i32, i32]),
    "rv_dpb_upload": (i32, [P, i32, P, P, P]),
    "rv_plane_download": (i32, [P, i32, P, P, P]),
    "rv_picture_hash": (i32, [P, i32, i32, P]),
    "rv_frame_md5": (i32, [i32, P, P, P, P, P]),
    "rv_picture_hash_noalloc": (i32, [P, i32, i32, P]),
    "rv_hevc_prepare": (i32, [P, C.POINTER(HevcPic), i32, C.POINTER(P)]),
    "rv_h264_prepare": (i32, [P, C.POINTER(H264Pic), i32, C.POINTER(P)]),
    "rv_job_free": (i32, [P]),
    "rv_batch_create": (i32, [C.POINTER(P), i32, C.POINTER(P)]),
    "rv_batch_free": (i32, [P]),
    "rv_batch_launch": (i32, [P, i32]),
    "rv_sync": (i32, []),
    "rv_batch_profile": (i32, [P, i32, C.POINTER(C.c_float)]),
    "rv_batch_tune": (i32, [P, i32]),
    "rv_batch_variants": (i32, [P, C.POINTER(i32)]),
    "rv_batch_set_variant": (i32, [P, i32, i32]),
    "rv_variant_count": (i32, [i32, i32]),
    "rv_variant_name": (C.c_char_p, [i32, i32, i32]),
    "rv_mark": (i32, [C.POINTER(i32)]),
    "rv_mark_query": (i32, [i32]),
    "rv_mark_wait": (i32, [i32]),
    "rv_mark_elapsed": (i32, [i32, i32, C.POINTER(C.c_float)]),
    "rv_set_stage_mask": (i32, [i32]),
    "rv_job_bytes": (i32, [P, C.POINTER(i64)]),
    "rv_abi_size": (i32, [i32]),
    "rv_probe_copy": (i32, [i64, i32, C.POINTER(C.c_double)]),
}

# The per-entry API (include/rivet_hip.h): buffers as void*, lengths,
# strides and sizes as size_t, neighbour offsets as ptrdiff_t.
_sz, _ss, _u = C.c_size_t, C.c_ssize_t, C.c_uint
ENTRY_PROTOTYPES = {
    "rv_hevc_idct": (i32, [i32, P, _sz, i32, _sz, _sz]),
    "rv_hevc_idst4": (i32, [P, _sz, i32, _sz, _sz]),
    "rv_hevc_qpel_v2": (i32, [P, _sz, P, _sz, _sz, _sz, _sz, _sz]),
    "rv_hevc_epel_v2": (i32, [P, _sz, P, _sz, _sz, _sz, _sz, _sz]),
    "rv_h264_u8_qpel": (i32, [i32, P, _sz, P, _sz, _sz, _sz, _sz, i32]),
    "rv_h264_u8_chroma": (i32, [P, _sz, P, _sz, _sz, _sz, _sz, i32, i32]),
    "rv_h264_u8_copy": (i32, [P, _sz, _sz, P, _sz, _sz, _sz]),
    "rv_h264_u8_avg": (i32, [P, _sz, _sz, P, _sz, P, _sz, _sz, _sz]),
    "rv_h264_u8_weighted_uni": (i32, [P, _sz, _sz, P, _sz, _sz, _sz, i32, i32, i32, i32]),
    "rv_h264_u8_weighted_bi": (i32, [P, _sz, _sz, P, _sz, P, _sz, _sz, _sz] + [i32] * 6),
    "rv_h264_u8_idct4_add": (i32, [P, _sz, _sz, P, i32]),
    "rv_h264_u8_idct8_add": (i32, [P, _sz, _sz, P, i32]),
    "rv_h264_u8_idct4_dc_add": (i32, [P, _sz, _sz, i32, i32]),
    "rv_h264_u8_idct8_dc_add": (i32, [P, _sz, _sz, i32, i32]),
    "rv_h264_u8_residual4": (i32, [P, _sz, _sz, P, i32, i32]),
    "rv_h264_u8_residual8": (i32, [P, _sz, _sz, P, i32]),
}
for _e in ("luma_v", "luma_h", "chroma_v", "chroma_h"):
    ENTRY_PROTOTYPES[f"rv_h264_u8_deblock_{_e}"] = (i32, [P, _sz, _sz, _sz, i32, i32, P, i32])
    ENTRY_PROTOTYPES[f"rv_h264_u8_deblock_{_e}_intra"] = (i32, [P, _sz, _sz, _sz, i32, i32, i32])
for _s in ("u8", "u16"):
    _t = {
        "add_residual": [P, _sz, _sz, P, _sz, _sz, i32],
        "qpel_copy": [P, _sz, P, _sz, _sz, _sz, _sz, i32],
        "epel_copy": [P, _sz, P, _sz, _sz, _sz, _sz, i32],
        "qpel_h": [P, _sz, P, _sz, _sz, _sz, _sz, _sz, i32],
        "qpel_v": [P, _sz, P, _sz, _sz, _sz, _sz, _sz, i32],
        "epel_h": [P, _sz, P, _sz, _sz, _sz, _sz, _sz, i32],
        "epel_v": [P, _sz, P, _sz, _sz, _sz, _sz, _sz, i32],
        "uni": [P, _sz, _sz, P, _sz, _sz, _sz, i32, i32],
        "bi": [P, _sz, _sz, P, _sz, P, _sz, _sz, _sz, i32, i32],
        "weighted_uni": [P, _sz, _sz, P, _sz, _sz, _sz, i32, i32, i32, i32],
        "weighted_bi": [P, _sz, _sz, P, _sz, P, _sz, _sz, _sz] + [i32] * 6,
        "qpel_uni": [P, _sz, _sz, P, _sz, _sz, _sz, _sz, _sz, _sz, _u],
        "epel_uni": [P, _sz, _sz, P, _sz, _sz, _sz, _sz, _sz, _sz, _u],
        "qpel_bi": [P, _sz, _sz, P, _sz, _sz, _sz, _sz, _sz, _sz, P, _sz, _u],
        "epel_bi": [P, _sz, _sz, P, _sz, _sz, _sz, _sz, _sz, _sz, P, _sz, _u],
        "sao_band": [P, _sz, _sz, P, _sz, _sz, _sz, _sz, P, i32, i32],
        "sao_edge": [P, _sz, P, _sz, _sz, _sz, _sz, _sz, _ss, _ss, P, i32],
        "deblock_luma_v": [P, _sz, _sz, _sz, P, P, P, P, i32],
        "deblock_luma_h": [P, _sz, _sz, _sz, P, P, P, P, i32],
        "deblock_chroma_v": [P, _sz, _sz, _sz, P, P, P, i32],
        "deblock_chroma_h": [P, _sz, _sz, _sz, P, P, P, i32],
    }
    for _n, _a in _t.items():
        ENTRY_PROTOTYPES[f"rv_hevc_{_s}_{_n}"] = (i32, _a)
PROTOTYPES.update(ENTRY_PROTOTYPES)

# include/rivet_gen.h (librivet_gen.so): the synthetic work-list generator,
# a library of its own, so the device tier carries no test-workload code.
GEN_PROTOTYPES = {
    "rv_gen_hevc": (i32, [C.POINTER(GenHevcCfg), C.c_uint64, C.POINTER(P)]),
    "rv_gen_h264": (i32, [C.POINTER(GenH264Cfg), C.c_uint64, C.POINTER(P)]),
    "rv_gen_hevc_pic": (C.POINTER(HevcPic), [P]),
    "rv_gen_h264_pic": (C.POINTER(H264Pic), [P]),
    "rv_gen_free": (None, [P]),
    "rv_gen_planes": (i32, [i32, i32, i32, C.c_uint64, P, P, P]),
    "rv_gen_subseed": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64]),
}


def _bind(lib: C.CDLL, protos: dict) -> C.CDLL:
    for name, (res, args) in protos.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def load_gen() -> C.CDLL:
    if not GEN_LIB_PATH.exists():
        raise ImportError(f"{GEN_LIB_PATH} is missing: run `python rivet-h26x-codecs_amd/build.py` first")
    return _bind(C.CDLL(str(GEN_LIB_PATH)), GEN_PROTOTYPES)


def load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run `python rivet-h26x-codecs_amd/build.py` "
                          "(or __graft_entry__.build()) first")
    lib = _bind(C.CDLL(str(LIB_PATH)), PROTOTYPES)
    for i, st in enumerate(ABI_STRUCTS):
        n = lib.rv_abi_size(i)
        if n != C.sizeof(st):
            raise ImportError(f"ABI mismatch: {st.__name__} is {C.sizeof(st)} bytes here, {n} in the library")
    return lib
