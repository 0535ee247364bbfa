"""TEST INFRASTRUCTURE ONLY — ctypes binding of the CPU oracle.

Loads oracle/liboracle.so (built by oracle/Makefile). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker / CPU baseline; the product path never
does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def _load():
    if not LIB.exists():
        raise ImportError(f"{LIB} missing: run `make -C oracle`")
    return C.CDLL(str(LIB))


lib = _load()
P = C.c_void_p
i16p = C.POINTER(C.c_int16)
i32p = C.POINTER(C.c_int32)
u8p = C.POINTER(C.c_uint8)
_sig = {
    "orc_hevc_idct": (None, [i16p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_hevc_idst4": (None, [i16p, C.c_int]),
    "orc_hevc_scale_coefficients": (None, [i16p, C.c_int, C.c_int, C.c_int, u8p, C.c_int, C.c_int, C.c_int]),
    "orc_hevc_intra_tables": (None, [i32p, i32p]),
    "orc_hevc_intra_substitute": (None, [P, P, C.c_int, C.c_int]),
    "orc_hevc_intra_pred": (None, [P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_hevc_intra_avail": (C.c_int, [P, C.c_int, P]),
    "orc_hevc_idct_matrix": (None, [i16p, C.c_int, C.c_int, i16p]),
    "orc_hevc_add_residual": (None, [C.c_int, P, C.c_int, i16p, C.c_int, C.c_int]),
    "orc_hevc_copy": (None, [C.c_int, i16p, P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_hevc_qpel_h": (None, [C.c_int, i16p, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_hevc_qpel_v": (None, [C.c_int, i16p, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_hevc_qpel_v2": (None, [i16p, i16p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_hevc_epel_h": (None, [C.c_int, i16p, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_hevc_epel_v": (None, [C.c_int, i16p, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_hevc_epel_v2": (None, [i16p, i16p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_hevc_uni": (None, [C.c_int, P, C.c_int, i16p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_hevc_bi": (None, [C.c_int, P, C.c_int, i16p, i16p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_hevc_weighted_uni": (None, [C.c_int, P, C.c_int, i16p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int]),
    "orc_hevc_weighted_bi": (None, [C.c_int, P, C.c_int, i16p, i16p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_int, C.c_int]),
    "orc_hevc_mc_fused": (None, [C.c_int, C.c_int, P, C.c_int, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, i16p,
                                 C.c_int]),
    "orc_hevc_sao_band": (None, [C.c_int, P, C.c_int, P, C.c_int, C.c_int, C.c_int, i16p, C.c_int, C.c_int]),
    "orc_hevc_sao_edge": (None, [C.c_int, P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, i16p,
                                 C.c_int]),
    "orc_hevc_deblock_luma": (None, [C.c_int, C.c_int, P, C.c_int, C.c_int, i32p, i32p, u8p, u8p, C.c_int]),
    "orc_hevc_deblock_chroma": (None, [C.c_int, C.c_int, P, C.c_int, C.c_int, i32p, u8p, u8p, C.c_int]),
    "orc_hevc_reconstruct": (C.c_int, [P, C.POINTER(P), C.c_int, C.POINTER(P), C.c_int]),
    "orc_hevc_reconstruct_mt": (C.c_int, [P, C.POINTER(P), C.c_int, C.POINTER(P), C.c_int, C.c_int]),
    "orc_hevc_bs_map": (C.c_int, [P, P, P]),
    "orc_h264_bs_map": (C.c_int, [P, P]),
    "orc_h264_qpel": (None, [C.c_int, u8p, u8p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_h264_chroma": (None, [u8p, u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_h264_copy": (None, [u8p, C.c_int, u8p, C.c_int, C.c_int]),
    "orc_h264_avg": (None, [u8p, C.c_int, u8p, u8p, C.c_int, C.c_int]),
    "orc_h264_idct4_add": (None, [u8p, C.c_int, i16p, C.c_int]),
    "orc_h264_idct8_add": (None, [u8p, C.c_int, i16p, C.c_int]),
    "orc_h264_dc_add": (None, [u8p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_h264_weighted_uni": (None, [u8p, C.c_int, u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_h264_weighted_bi": (None, [u8p, C.c_int, u8p, u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_int]),
    "orc_h264_deblock": (None, [C.c_int, C.c_int, u8p, C.c_int, C.c_int, C.c_int, C.c_int, i16p, C.c_int]),
    "orc_h264_residual4": (None, [u8p, C.c_int, i32p, C.c_int, C.c_int, C.c_int]),
    "orc_h264_residual8": (None, [u8p, C.c_int, i32p, C.c_int]),
    "orc_h264_level_scale4": (C.c_int, [P, C.c_int, C.c_int, C.c_int]),
    "orc_h264_level_scale8": (C.c_int, [P, C.c_int, C.c_int, C.c_int]),
    "orc_h264_dequant_level": (C.c_int32, [C.c_int32, C.c_int32, C.c_int]),
    "orc_h264_luma_dc": (None, [i32p, C.c_int, C.c_int]),
    "orc_h264_chroma_dc": (None, [i32p, C.c_int, C.c_int]),
    "orc_h264_implicit_pair": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, i32p]),
    "orc_h264_reconstruct": (C.c_int, [P, C.POINTER(P), C.c_int, C.POINTER(P), C.c_int]),
    "orc_h264_intra_pred": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, P, C.c_int]),
    "orc_h264_filter8": (None, [i32p, C.c_int, i32p]),
    "orc_h264_block_avail": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_crc16": (C.c_uint16, [P, C.c_size_t]),
    "orc_checksum": (C.c_uint32, [P, C.c_size_t, C.c_int, C.c_int]),
    "orc_md5": (None, [P, C.c_size_t, u8p]),
}
for _n, (_r, _a) in _sig.items():
    _f = getattr(lib, _n)
    _f.restype = _r
    _f.argtypes = _a


def ptr(a: np.ndarray, t=None):
    a = np.ascontiguousarray(a)
    return a.ctypes.data_as(t) if t is not None else C.c_void_p(a.ctypes.data)


def component_message(plane: np.ndarray) -> bytes:
    """One component's hash message (src/hevc/hash.rs:131-142): its samples
    in raster order as bytes, low byte first for uint16 planes."""
    a = np.ascontiguousarray(plane)
    return a.astype("<u2").tobytes() if a.dtype == np.uint16 else a.astype(np.uint8).tobytes()


def crc16(msg: bytes) -> int:
    return int(lib.orc_crc16(msg, len(msg)))


def checksum(msg: bytes, width: int, wide: bool) -> int:
    return int(lib.orc_checksum(msg, len(msg), width, int(wide)))


def md5(msg: bytes) -> bytes:
    out = (C.c_uint8 * 16)()
    lib.orc_md5(msg, len(msg), out)
    return bytes(out)


def picture_hash(planes, kind: int) -> list:
    """The three component hashes of hash::verify (src/hevc/hash.rs:109-158)
    for kind 0 (MD5), 1 (CRC) or 2 (checksum)."""
    out = []
    for p in planes:
        m = component_message(p)
        if kind == 0:
            out.append(md5(m))
        elif kind == 1:
            out.append(crc16(m))
        else:
            out.append(checksum(m, p.shape[1], p.dtype == np.uint16))
    return out


def hevc_bs_map(pic):
    """(ver, hor) boundary-strength maps of an HEVC picture description, one
    entry per 4x4 block (orc_hevc_bs_map)."""
    c = pic.contents
    ver = np.zeros((c.height // 4, c.width // 4), np.uint8)
    hor = np.zeros_like(ver)
    rc = lib.orc_hevc_bs_map(C.cast(pic, C.c_void_p), ver.ctypes.data, hor.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle failed: {rc}")
    return ver, hor


def h264_bs_map(pic):
    """Per macroblock (mby, mbx): bS[2][4][4] = [vertical | horizontal][edge
    0..3][segment 0..3] (orc_h264_bs_map)."""
    c = pic.contents
    raw = np.zeros(c.n_mbs * 8, np.uint32)
    rc = lib.orc_h264_bs_map(C.cast(pic, C.c_void_p), raw.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle failed: {rc}")
    b = raw.view(np.uint8).reshape(c.height // 16, c.width // 16, 2, 4, 4)
    return b


def reconstruct(codec: int, pic, slots: list, stop_after: int = 0, threads: int = 1):
    """Run the oracle on a picture description (ctypes pointer).

    slots[k] = (y, cb, cr) numpy planes of DPB slot k (or None).
    Returns the three output planes. threads > 1 (HEVC) splits the
    picture's work over that many threads.
    """
    c = pic.contents
    w, h = c.width, c.height
    dt = np.uint8 if c.bit_depth <= 8 else np.uint16
    flat = (C.c_void_p * (3 * len(slots)))()
    keep = []
    for k, s in enumerate(slots):
        for j in range(3):
            if s is None:
                flat[3 * k + j] = None
            else:
                a = np.ascontiguousarray(s[j], dtype=dt)
                keep.append(a)
                flat[3 * k + j] = a.ctypes.data
    outs = [np.zeros((h, w), dt), np.zeros((h // 2, w // 2), dt), np.zeros((h // 2, w // 2), dt)]
    optr = (C.c_void_p * 3)(*(o.ctypes.data for o in outs))
    if codec == 0 and threads > 1:
        rc = lib.orc_hevc_reconstruct_mt(C.cast(pic, C.c_void_p), flat, len(slots), optr, stop_after, threads)
    else:
        f = lib.orc_hevc_reconstruct if codec == 0 else lib.orc_h264_reconstruct
        rc = f(C.cast(pic, C.c_void_p), flat, len(slots), optr, stop_after)
    if rc != 0:
        raise RuntimeError(f"oracle failed: {rc}")
    return outs
