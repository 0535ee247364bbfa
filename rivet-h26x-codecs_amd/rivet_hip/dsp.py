"""The kernel tables of the HIP tier, entry for entry.

`HevcDsp(sample)` mirrors `HevcDsp<S>` (src/dsp/hevc.rs:165-229) and
`H264Dsp()` mirrors `H264Dsp<u8>` (src/dsp/h264.rs:123-170): the same entry
names, argument order and meaning. `idct[log2 - 2]` and `qpel[yf * 4 + xf]`
are indexed as in the reference.

A Rust slice is a C-contiguous numpy array of the entry's element type.
Its length is the array's size, and `&mut` slices are written in place.
Where the reference panics (an index outside a slice, a value its
arithmetic cannot take), the call raises RivetError with code RV_E_RANGE
and changes nothing. Every call runs one kernel on the device and returns
when its outputs are back in the arrays (the per-entry C API,
include/rivet_hip.h).

The intra entries (intra_planar / intra_dc / intra_angular) and the
MBAFF-only deblock_luma8_* entries are not part of this tier.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import check, lib, tier_name

NO_DC = -(2**31)  # src/dsp/h264.rs:118
PRED_STRIDE = 16  # src/dsp/h264.rs:22-24


def _buf(a: np.ndarray, dtype, what: str, writable: bool = False):
    if not isinstance(a, np.ndarray) or a.dtype != np.dtype(dtype) or not a.flags.c_contiguous:
        raise TypeError(f"{what}: expected a C-contiguous {np.dtype(dtype).name} array")
    if writable and not a.flags.writeable:
        raise TypeError(f"{what}: array is read-only")
    return C.c_void_p(a.ctypes.data), a.size


def _arr(values, ctype, n: int, what: str):
    v = list(values)
    if len(v) != n:
        raise TypeError(f"{what}: expected {n} values")
    return (ctype * n)(*v)


class HevcDsp:
    """HevcDsp<S> for S = uint8 (8-bit) or uint16 (9..16-bit)."""

    fused_mc = True

    def __init__(self, sample=np.uint8):
        self.S = np.dtype(sample)
        if self.S not in (np.dtype(np.uint8), np.dtype(np.uint16)):
            raise TypeError("sample must be uint8 or uint16")
        self.sfx = "u8" if self.S == np.uint8 else "u16"
        self.cpu = tier_name()
        self.idct = [self._idct_n(log2) for log2 in (2, 3, 4, 5)]

    def _f(self, name):
        return getattr(lib, f"rv_hevc_{self.sfx}_{name}")

    # -- transforms ---------------------------------------------------------
    def _idct_n(self, log2):
        def idct(coeffs, bd_shift, max_x, max_y):
            p, n = _buf(coeffs, np.int16, "coeffs", True)
            check(lib.rv_hevc_idct(log2, p, n, bd_shift, max_x, max_y))
        return idct

    def idst4(self, coeffs, bd_shift, max_x=0, max_y=0):
        p, n = _buf(coeffs, np.int16, "coeffs", True)
        check(lib.rv_hevc_idst4(p, n, bd_shift, max_x, max_y))

    def add_residual(self, dst, stride, res, n, max):
        d, dn = _buf(dst, self.S, "dst", True)
        r, rn = _buf(res, np.int16, "res")
        check(self._f("add_residual")(d, dn, stride, r, rn, n, max))

    # -- interpolation ------------------------------------------------------
    def _copy(self, name, dst, src, src_stride, w, h, shift):
        d, dn = _buf(dst, np.int16, "dst", True)
        s, sn = _buf(src, self.S, "src")
        check(self._f(name)(d, dn, s, sn, src_stride, w, h, shift))

    def _interp(self, name, dst, src, src_stride, w, h, frac, shift):
        d, dn = _buf(dst, np.int16, "dst", True)
        s, sn = _buf(src, self.S, "src")
        check(self._f(name)(d, dn, s, sn, src_stride, w, h, frac, shift))

    def qpel_copy(self, dst, src, src_stride, w, h, shift):
        self._copy("qpel_copy", dst, src, src_stride, w, h, shift)

    def epel_copy(self, dst, src, src_stride, w, h, shift):
        self._copy("epel_copy", dst, src, src_stride, w, h, shift)

    def qpel_h(self, dst, src, src_stride, w, h, frac, shift):
        self._interp("qpel_h", dst, src, src_stride, w, h, frac, shift)

    def qpel_v(self, dst, src, src_stride, w, h, frac, shift):
        self._interp("qpel_v", dst, src, src_stride, w, h, frac, shift)

    def epel_h(self, dst, src, src_stride, w, h, frac, shift):
        self._interp("epel_h", dst, src, src_stride, w, h, frac, shift)

    def epel_v(self, dst, src, src_stride, w, h, frac, shift):
        self._interp("epel_v", dst, src, src_stride, w, h, frac, shift)

    @staticmethod
    def _v2(fn, dst, src, src_stride, w, h, frac):
        d, dn = _buf(dst, np.int16, "dst", True)
        s, sn = _buf(src, np.int16, "src")
        check(fn(d, dn, s, sn, src_stride, w, h, frac))

    def qpel_v2(self, dst, src, src_stride, w, h, frac):
        self._v2(lib.rv_hevc_qpel_v2, dst, src, src_stride, w, h, frac)

    def epel_v2(self, dst, src, src_stride, w, h, frac):
        self._v2(lib.rv_hevc_epel_v2, dst, src, src_stride, w, h, frac)

    # -- combination --------------------------------------------------------
    def uni(self, dst, stride, src, w, h, shift, max):
        d, dn = _buf(dst, self.S, "dst", True)
        s, sn = _buf(src, np.int16, "src")
        check(self._f("uni")(d, dn, stride, s, sn, w, h, shift, max))

    def bi(self, dst, stride, a, b, w, h, shift, max):
        d, dn = _buf(dst, self.S, "dst", True)
        pa, an = _buf(a, np.int16, "a")
        pb, bn = _buf(b, np.int16, "b")
        check(self._f("bi")(d, dn, stride, pa, an, pb, bn, w, h, shift, max))

    def weighted_uni(self, dst, stride, src, w, h, log2_wd, wt, o, max):
        d, dn = _buf(dst, self.S, "dst", True)
        s, sn = _buf(src, np.int16, "src")
        check(self._f("weighted_uni")(d, dn, stride, s, sn, w, h, log2_wd, wt, o, max))

    def weighted_bi(self, dst, stride, a, b, w, h, log2_wd, w0, w1, o0, o1, max):
        d, dn = _buf(dst, self.S, "dst", True)
        pa, an = _buf(a, np.int16, "a")
        pb, bn = _buf(b, np.int16, "b")
        check(self._f("weighted_bi")(d, dn, stride, pa, an, pb, bn, w, h, log2_wd, w0, w1, o0, o1, max))

    # -- fused interpolation + combination ------------------------------------
    def _fused_uni(self, name, dst, dst_stride, src, src_stride, w, h, fx, fy, tmp, bit_depth):
        d, dn = _buf(dst, self.S, "dst", True)
        s, sn = _buf(src, self.S, "src")
        check(self._f(name)(d, dn, dst_stride, s, sn, src_stride, w, h, fx, fy, bit_depth))

    def _fused_bi(self, name, dst, dst_stride, src, src_stride, w, h, fx, fy, tmp, other, bit_depth):
        d, dn = _buf(dst, self.S, "dst", True)
        s, sn = _buf(src, self.S, "src")
        o, on = _buf(other, np.int16, "other")
        check(self._f(name)(d, dn, dst_stride, s, sn, src_stride, w, h, fx, fy, o, on, bit_depth))

    # `tmp` is accepted for the reference's signature; the device needs no
    # host scratch (pass None).
    def qpel_uni(self, dst, dst_stride, src, src_stride, w, h, fx, fy, tmp, bit_depth):
        self._fused_uni("qpel_uni", dst, dst_stride, src, src_stride, w, h, fx, fy, tmp, bit_depth)

    def epel_uni(self, dst, dst_stride, src, src_stride, w, h, fx, fy, tmp, bit_depth):
        self._fused_uni("epel_uni", dst, dst_stride, src, src_stride, w, h, fx, fy, tmp, bit_depth)

    def qpel_bi(self, dst, dst_stride, src, src_stride, w, h, fx, fy, tmp, other, bit_depth):
        self._fused_bi("qpel_bi", dst, dst_stride, src, src_stride, w, h, fx, fy, tmp, other, bit_depth)

    def epel_bi(self, dst, dst_stride, src, src_stride, w, h, fx, fy, tmp, other, bit_depth):
        self._fused_bi("epel_bi", dst, dst_stride, src, src_stride, w, h, fx, fy, tmp, other, bit_depth)

    # -- SAO ----------------------------------------------------------------
    def sao_band(self, dst, dst_stride, src, src_stride, w, h, table, shift, max):
        d, dn = _buf(dst, self.S, "dst", True)
        s, sn = _buf(src, self.S, "src")
        check(self._f("sao_band")(d, dn, dst_stride, s, sn, src_stride, w, h,
                                  _arr(table, C.c_int16, 32, "table"), shift, max))

    def sao_edge(self, dst, src, origin, stride, w, h, na, nb, off, max):
        d, dn = _buf(dst, self.S, "dst", True)
        s, sn = _buf(src, self.S, "src")
        check(self._f("sao_edge")(d, dn, s, sn, origin, stride, w, h, na, nb, _arr(off, C.c_int16, 5, "off"), max))

    # -- deblocking ---------------------------------------------------------
    def _luma(self, name, data, off, stride, beta, tc, no_p, no_q, max):
        d, dn = _buf(data, self.S, "data", True)
        check(self._f(name)(d, dn, off, stride, _arr(beta, C.c_int32, 2, "beta"), _arr(tc, C.c_int32, 2, "tc"),
                            _arr(map(int, no_p), C.c_uint8, 2, "no_p"), _arr(map(int, no_q), C.c_uint8, 2, "no_q"),
                            max))

    def _chroma(self, name, data, off, stride, tc, no_p, no_q, max):
        d, dn = _buf(data, self.S, "data", True)
        check(self._f(name)(d, dn, off, stride, _arr(tc, C.c_int32, 4, "tc"),
                            _arr(map(int, no_p), C.c_uint8, 4, "no_p"), _arr(map(int, no_q), C.c_uint8, 4, "no_q"),
                            max))

    def deblock_luma_v(self, data, off, stride, beta, tc, no_p, no_q, max):
        self._luma("deblock_luma_v", data, off, stride, beta, tc, no_p, no_q, max)

    def deblock_luma_h(self, data, off, stride, beta, tc, no_p, no_q, max):
        self._luma("deblock_luma_h", data, off, stride, beta, tc, no_p, no_q, max)

    def deblock_chroma_v(self, data, off, stride, tc, no_p, no_q, max):
        self._chroma("deblock_chroma_v", data, off, stride, tc, no_p, no_q, max)

    def deblock_chroma_h(self, data, off, stride, tc, no_p, no_q, max):
        self._chroma("deblock_chroma_h", data, off, stride, tc, no_p, no_q, max)


class H264Dsp:
    """H264Dsp<u8>."""

    def __init__(self):
        self.cpu = tier_name()
        self.qpel = [self._qpel_pos(p) for p in range(16)]

    def _qpel_pos(self, pos):
        def qpel(dst, src, src_stride, w, h, max):
            d, dn = _buf(dst, np.uint8, "dst", True)
            s, sn = _buf(src, np.uint8, "src")
            check(lib.rv_h264_u8_qpel(pos, d, dn, s, sn, src_stride, w, h, max))
        return qpel

    def chroma(self, dst, src, stride, w, h, xf, yf):
        d, dn = _buf(dst, np.uint8, "dst", True)
        s, sn = _buf(src, np.uint8, "src")
        check(lib.rv_h264_u8_chroma(d, dn, s, sn, stride, w, h, xf, yf))

    def copy(self, dst, stride, src, w, h):
        d, dn = _buf(dst, np.uint8, "dst", True)
        s, sn = _buf(src, np.uint8, "src")
        check(lib.rv_h264_u8_copy(d, dn, stride, s, sn, w, h))

    def avg(self, dst, stride, a, b, w, h):
        d, dn = _buf(dst, np.uint8, "dst", True)
        pa, an = _buf(a, np.uint8, "a")
        pb, bn = _buf(b, np.uint8, "b")
        check(lib.rv_h264_u8_avg(d, dn, stride, pa, an, pb, bn, w, h))

    def weighted_uni(self, dst, stride, src, w, h, log_wd, wt, o, max):
        d, dn = _buf(dst, np.uint8, "dst", True)
        s, sn = _buf(src, np.uint8, "src")
        check(lib.rv_h264_u8_weighted_uni(d, dn, stride, s, sn, w, h, log_wd, wt, o, max))

    def weighted_bi(self, dst, stride, a, b, w, h, log_wd, w0, w1, o0, o1, max):
        d, dn = _buf(dst, np.uint8, "dst", True)
        pa, an = _buf(a, np.uint8, "a")
        pb, bn = _buf(b, np.uint8, "b")
        check(lib.rv_h264_u8_weighted_bi(d, dn, stride, pa, an, pb, bn, w, h, log_wd, w0, w1, o0, o1, max))

    def _dbk(self, name, data, off, stride, alpha, beta, tc0, max):
        d, dn = _buf(data, np.uint8, "data", True)
        check(getattr(lib, name)(d, dn, off, stride, alpha, beta, _arr(tc0, C.c_int16, 4, "tc0"), max))

    def _dbk_intra(self, name, data, off, stride, alpha, beta, max):
        d, dn = _buf(data, np.uint8, "data", True)
        check(getattr(lib, name)(d, dn, off, stride, alpha, beta, max))

    def deblock_luma_v(self, data, off, stride, alpha, beta, tc0, max):
        self._dbk("rv_h264_u8_deblock_luma_v", data, off, stride, alpha, beta, tc0, max)

    def deblock_luma_h(self, data, off, stride, alpha, beta, tc0, max):
        self._dbk("rv_h264_u8_deblock_luma_h", data, off, stride, alpha, beta, tc0, max)

    def deblock_chroma_v(self, data, off, stride, alpha, beta, tc0, max):
        self._dbk("rv_h264_u8_deblock_chroma_v", data, off, stride, alpha, beta, tc0, max)

    def deblock_chroma_h(self, data, off, stride, alpha, beta, tc0, max):
        self._dbk("rv_h264_u8_deblock_chroma_h", data, off, stride, alpha, beta, tc0, max)

    def deblock_luma_v_intra(self, data, off, stride, alpha, beta, max):
        self._dbk_intra("rv_h264_u8_deblock_luma_v_intra", data, off, stride, alpha, beta, max)

    def deblock_luma_h_intra(self, data, off, stride, alpha, beta, max):
        self._dbk_intra("rv_h264_u8_deblock_luma_h_intra", data, off, stride, alpha, beta, max)

    def deblock_chroma_v_intra(self, data, off, stride, alpha, beta, max):
        self._dbk_intra("rv_h264_u8_deblock_chroma_v_intra", data, off, stride, alpha, beta, max)

    def deblock_chroma_h_intra(self, data, off, stride, alpha, beta, max):
        self._dbk_intra("rv_h264_u8_deblock_chroma_h_intra", data, off, stride, alpha, beta, max)

    def _fixed(self, a, dtype, n, what):
        if not isinstance(a, np.ndarray) or a.dtype != np.dtype(dtype) or a.size != n or not a.flags.c_contiguous:
            raise TypeError(f"{what}: expected a C-contiguous {np.dtype(dtype).name}[{n}]")
        return C.c_void_p(a.ctypes.data)

    def idct4_add(self, dst, stride, coeffs, max):
        d, dn = _buf(dst, np.uint8, "dst", True)
        check(lib.rv_h264_u8_idct4_add(d, dn, stride, self._fixed(coeffs, np.int16, 16, "coeffs"), max))

    def idct8_add(self, dst, stride, coeffs, max):
        d, dn = _buf(dst, np.uint8, "dst", True)
        check(lib.rv_h264_u8_idct8_add(d, dn, stride, self._fixed(coeffs, np.int16, 64, "coeffs"), max))

    def idct4_dc_add(self, dst, stride, dc, max):
        d, dn = _buf(dst, np.uint8, "dst", True)
        check(lib.rv_h264_u8_idct4_dc_add(d, dn, stride, dc, max))

    def idct8_dc_add(self, dst, stride, dc, max):
        d, dn = _buf(dst, np.uint8, "dst", True)
        check(lib.rv_h264_u8_idct8_dc_add(d, dn, stride, dc, max))

    def residual4(self, dst, stride, coefs, dc, max):
        d, dn = _buf(dst, np.uint8, "dst", True)
        check(lib.rv_h264_u8_residual4(d, dn, stride, self._fixed(coefs, np.int32, 16, "coefs"), dc, max))

    def residual8(self, dst, stride, coefs, max):
        d, dn = _buf(dst, np.uint8, "dst", True)
        check(lib.rv_h264_u8_residual8(d, dn, stride, self._fixed(coefs, np.int32, 64, "coefs"), max))
