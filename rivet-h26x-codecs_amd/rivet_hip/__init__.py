"""rivet_hip — the MI355X (gfx950) picture-reconstruction tier.

Host-side wrapper over the C ABI (include/rivet_hip.h): sessions with a
device-resident DPB, prepared jobs (one picture's work lists), batches,
and the seeded generator of synthetic work lists (librivet_gen.so). All
device work is done by librivet_hip.so; nothing here computes picture
samples.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _abi
from ._abi import (RV_CODEC_H264, RV_CODEC_HEVC, RV_HASH_CHECKSUM, RV_HASH_CRC, RV_HASH_MD5,  # noqa: F401
                   RV_LAUNCH_BATCHED, RV_LAUNCH_GRAPH, RV_LAUNCH_PLAIN)

lib = _abi.load()
genlib = _abi.load_gen()  # the synthetic work-list generator (host only)


class RivetError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rivet_hip error {code}: {msg}")
        self.code = code


def check(rc: int) -> None:
    if rc != 0:
        raise RivetError(rc, (lib.rv_last_error() or b"").decode())


def device_count() -> int:
    return int(lib.rv_device_count())


_initialised = False


def init(device: int = 0) -> None:
    global _initialised
    check(lib.rv_init(device))
    _initialised = True


def tier_name() -> str:
    return lib.rv_tier_name().decode()


def sample_dtype(bit_depth: int):
    return np.uint8 if bit_depth <= 8 else np.uint16


def plane_shapes(width: int, height: int):
    return [(height, width), (height // 2, width // 2), (height // 2, width // 2)]


def gen_planes(width: int, height: int, bit_depth: int, seed: int):
    """Uniform random 4:2:0 planes (the generator's reference pictures)."""
    dt = sample_dtype(bit_depth)
    ys = [np.empty(s, dtype=dt) for s in plane_shapes(width, height)]
    check(genlib.rv_gen_planes(width, height, bit_depth, seed, *(p.ctypes.data for p in ys)))
    return ys


def subseed(seed: int, a: int = 0, b: int = 0, c: int = 0) -> int:
    return int(genlib.rv_gen_subseed(seed, a, b, c))


class GenPicture:
    """A generated picture description (owns its arrays)."""

    def __init__(self, codec: int, cfg: dict, seed: int):
        self.codec = codec
        self.cfg = dict(cfg)
        self.seed = seed
        h = C.c_void_p()
        if codec == RV_CODEC_HEVC:
            c = _abi.GenHevcCfg()
            for k, v in cfg.items():
                if not hasattr(c, k):
                    raise KeyError(f"no generator setting {k!r}")
                if k in ("ref_poc", "ref_slot"):
                    getattr(c, k)[:] = v
                else:
                    setattr(c, k, v)
            check(genlib.rv_gen_hevc(C.byref(c), seed, C.byref(h)))
            self._h = h
            self.pic = genlib.rv_gen_hevc_pic(h)
        else:
            c = _abi.GenH264Cfg()
            for k, v in cfg.items():
                if not hasattr(c, k):
                    raise KeyError(f"no generator setting {k!r}")
                if k in ("ref_poc", "ref_slot"):
                    getattr(c, k)[:] = v
                else:
                    setattr(c, k, v)
            check(genlib.rv_gen_h264(C.byref(c), seed, C.byref(h)))
            self._h = h
            self.pic = genlib.rv_gen_h264_pic(h)

    @property
    def width(self) -> int:
        return self.pic.contents.width

    @property
    def height(self) -> int:
        return self.pic.contents.height

    @property
    def bit_depth(self) -> int:
        return self.pic.contents.bit_depth

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            genlib.rv_gen_free(h)
            self._h = None


class Session:
    """One stream: DPB planes resident on the device."""

    def __init__(self, codec: int, width: int, height: int, bit_depth: int, dpb_slots: int):
        self.codec, self.width, self.height, self.bit_depth, self.slots = codec, width, height, bit_depth, dpb_slots
        h = C.c_void_p()
        check(lib.rv_session_create(codec, width, height, bit_depth, dpb_slots, C.byref(h)))
        self._h = h

    @staticmethod
    def footprint(codec: int, width: int, height: int, bit_depth: int, dpb_slots: int) -> int:
        return int(lib.rv_session_footprint(codec, width, height, bit_depth, dpb_slots))

    def upload(self, slot: int, planes: Sequence[np.ndarray]) -> None:
        dt = sample_dtype(self.bit_depth)
        ps = [np.ascontiguousarray(p, dtype=dt) for p in planes]
        for p, s in zip(ps, plane_shapes(self.width, self.height)):
            if p.shape != s:
                raise ValueError(f"plane shape {p.shape} != {s}")
        check(lib.rv_dpb_upload(self._h, slot, *(p.ctypes.data for p in ps)))

    def download(self, slot: int):
        dt = sample_dtype(self.bit_depth)
        ps = [np.empty(s, dtype=dt) for s in plane_shapes(self.width, self.height)]
        check(lib.rv_plane_download(self._h, slot, *(p.ctypes.data for p in ps)))
        return ps

    def picture_hash(self, slot: int, kind: int) -> list:
        """Decoded picture hash of DPB slot `slot`, computed on the device
        (rv_picture_hash; hash::verify, src/hevc/hash.rs:109-158): per
        component (Y, Cb, Cr), a 16-byte MD5 digest (RV_HASH_MD5), the CRC-16
        (RV_HASH_CRC) or the checksum (RV_HASH_CHECKSUM)."""
        buf = (C.c_uint8 * 48)()
        check(lib.rv_picture_hash(self._h, slot, kind, buf))
        raw = bytes(buf)
        if kind == RV_HASH_MD5:
            return [raw[16 * c:16 * c + 16] for c in range(3)]
        if kind == RV_HASH_CRC:
            return [int(v) for v in np.frombuffer(raw[:6], dtype="<u2")]
        return [int(v) for v in np.frombuffer(raw[:12], dtype="<u4")]

    def prepare(self, pic: GenPicture | C._Pointer, out_slot: int) -> "Job":
        return Job(self, pic, out_slot)

    def close(self) -> None:
        if self._h and self._h.value:
            check(lib.rv_session_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def frame_md5(pictures) -> list:
    """MD5 digests (16 bytes each) of pictures packed planar, computed on
    the device in one launch (rv_frame_md5): `pictures` is a sequence of
    (session, slot, width, height), width x height the output (cropped)
    luma size."""
    n = len(pictures)
    if n == 0:
        return []
    sess = (C.c_void_p * n)(*[p[0]._h.value for p in pictures])
    slots = (C.c_int * n)(*[p[1] for p in pictures])
    ws = (C.c_int * n)(*[p[2] for p in pictures])
    hs = (C.c_int * n)(*[p[3] for p in pictures])
    out = (C.c_uint8 * (16 * n))()
    check(lib.rv_frame_md5(n, sess, slots, ws, hs, out))
    raw = bytes(out)
    return [raw[16 * i:16 * i + 16] for i in range(n)]


def frame_line(index: int, poc: int, decode_index: int, width: int, height: int, digest: bytes) -> str:
    """One output frame as examples/h26xdec.rs:124-132 prints it:
    frame index, POC, decode index, WxH and the MD5 of the packed picture
    in lowercase hex."""
    return f"{index},{poc},{decode_index},{width}x{height},{digest.hex()}"


class Job:
    """A prepared picture (work lists on the device)."""

    def __init__(self, session: Session, pic, out_slot: int):
        self.session = session
        self.out_slot = out_slot
        p = pic.pic if isinstance(pic, GenPicture) else pic
        h = C.c_void_p()
        if session.codec == RV_CODEC_HEVC:
            check(lib.rv_hevc_prepare(session._h, p, out_slot, C.byref(h)))
        else:
            check(lib.rv_h264_prepare(session._h, p, out_slot, C.byref(h)))
        self._h = h

    def bytes(self):
        out = (C.c_int64 * 4)()
        check(lib.rv_job_bytes(self._h, out))
        return {"mc": out[0], "tx": out[1], "dbk": out[2], "sao": out[3]}

    def close(self) -> None:
        if self._h and self._h.value:
            check(lib.rv_job_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Jobs launched together (tables uploaded once)."""

    def __init__(self, jobs: Sequence[Job]):
        self.jobs = list(jobs)
        arr = (C.c_void_p * len(self.jobs))(*(j._h.value for j in self.jobs))
        h = C.c_void_p()
        check(lib.rv_batch_create(arr, len(self.jobs), C.byref(h)))
        self._h = h

    def launch(self, mode: int = RV_LAUNCH_BATCHED) -> None:
        check(lib.rv_batch_launch(self._h, mode))

    def profile(self, reps: int = 10):
        """Average milliseconds of each of the five stages (HIP events)."""
        out = (C.c_float * 5)()
        check(lib.rv_batch_profile(self._h, reps, out))
        return list(out)

    def tune(self, reps: int = 3) -> list:
        """Time every kernel variant the batch allows; returns the chosen
        variant per stage (also recorded for the batch's depth class)."""
        check(lib.rv_batch_tune(self._h, reps))
        return self.variants()

    def variants(self) -> list:
        out = (C.c_int32 * 5)()
        check(lib.rv_batch_variants(self._h, out))
        return list(out)

    def set_variant(self, stage: int, variant: int) -> None:
        check(lib.rv_batch_set_variant(self._h, stage, variant))

    def close(self) -> None:
        if self._h and self._h.value:
            check(lib.rv_batch_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def variant_names(codec: int = RV_CODEC_HEVC) -> list:
    """Names of the kernel variants of each stage."""
    return [[lib.rv_variant_name(codec, s, v).decode() for v in range(lib.rv_variant_count(codec, s))]
            for s in range(5)]


def sync() -> None:
    check(lib.rv_sync())


def mark() -> int:
    i = C.c_int32()
    check(lib.rv_mark(C.byref(i)))
    return i.value


def mark_wait(i: int) -> None:
    check(lib.rv_mark_wait(i))


def mark_done(i: int) -> bool:
    rc = lib.rv_mark_query(i)
    if rc < 0:
        check(rc)
    return rc == 0


def mark_elapsed_ms(a: int, b: int) -> float:
    v = C.c_float()
    check(lib.rv_mark_elapsed(a, b, C.byref(v)))
    return float(v.value)


def probe_copy_gbs(nbytes: int = 1 << 30, reps: int = 10) -> float:
    """Achievable copy bandwidth (read + write GB/s) of a plain vector copy."""
    v = C.c_double()
    check(lib.rv_probe_copy(nbytes, reps, C.byref(v)))
    return float(v.value)


def set_stage_mask(mask: int) -> None:
    check(lib.rv_set_stage_mask(mask))
