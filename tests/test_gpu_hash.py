"""Decoded picture hashes computed on the device (rv_picture_hash) against
the oracle's restatement of hash::verify (src/hevc/hash.rs:109-158): on
uploaded planes at several sizes and depths, and on the output slot of
reconstructed HEVC and H.264 pictures, hashed right after the launch so the
call's own ordering after the slot's writer is what the test sees."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import configs as cf

pytestmark = pytest.mark.gpu

KINDS = [rh.RV_HASH_MD5, rh.RV_HASH_CRC, rh.RV_HASH_CHECKSUM]
RV_E_INVALID, RV_E_RANGE = -1, -5


@pytest.mark.parametrize("codec,w,h,bd", [(rh.RV_CODEC_HEVC, 416, 240, 8), (rh.RV_CODEC_HEVC, 64, 64, 10),
                                          (rh.RV_CODEC_HEVC, 856, 480, 8), (rh.RV_CODEC_H264, 1920, 1088, 8),
                                          (rh.RV_CODEC_HEVC, 1920, 1080, 10)])
def test_hash_of_uploaded_planes(device, codec, w, h, bd):
    print(f"hash of uploaded planes: codec {codec} {w}x{h} {bd}-bit", flush=True)
    planes = rh.gen_planes(w, h, bd, rh.subseed(77, w, h, bd))
    s = rh.Session(codec, w, h, bd, 2)
    s.upload(1, planes)
    for k in KINDS:
        assert s.picture_hash(1, k) == oracle.picture_hash(planes, k), f"kind {k}"
    s.close()


def _reconstructed(codec, cfg, seed):
    bd = cfg.get("bit_depth", 8)
    refs = [rh.gen_planes(cfg["width"], cfg["height"], bd, cf.ref_seed(seed, k)) for k in range(2)]
    g = rh.GenPicture(codec, cfg, seed)
    s = rh.Session(codec, cfg["width"], cfg["height"], bd, 3)
    for k in range(2):
        s.upload(k, refs[k])
    j = s.prepare(g, 2)
    b = rh.Batch([j])
    b.launch(rh.RV_LAUNCH_BATCHED)  # no host sync before the hash
    want = oracle.reconstruct(codec, g.pic, refs + [None], 0)
    return s, j, b, want


@pytest.mark.parametrize("codec,cfg,seed", [(rh.RV_CODEC_HEVC, cf.HEVC_C1, 4711),
                                            (rh.RV_CODEC_HEVC, dict(cf.HEVC_C2, width=416, height=240), 4712),
                                            (rh.RV_CODEC_H264, dict(cf.H264_C4, width=352, height=288), 4713)])
def test_hash_of_reconstructed_picture(device, codec, cfg, seed):
    print(f"hash of a reconstructed picture: codec {codec} seed {seed}", flush=True)
    s, j, b, want = _reconstructed(codec, cfg, seed)
    for k in KINDS:
        assert s.picture_hash(2, k) == oracle.picture_hash(want, k), f"kind {k}"
    b.close()
    j.close()
    s.close()


def test_hash_argument_errors(device):
    s = rh.Session(rh.RV_CODEC_HEVC, 64, 64, 8, 2)
    buf = (C.c_uint8 * 48)()
    assert rh.lib.rv_picture_hash(s._h, 0, 3, buf) == RV_E_INVALID
    assert rh.lib.rv_picture_hash(s._h, 2, rh.RV_HASH_CRC, buf) == RV_E_RANGE
    assert rh.lib.rv_picture_hash(s._h, 0, rh.RV_HASH_CRC, None) == RV_E_INVALID
    s.close()


# ------------------------------------------------ per-frame MD5 (h26xdec lines)

def packed(planes, w, h):
    """Picture::into_packed of the output size: Y, Cb, Cr rows cropped,
    samples little-endian (written here with numpy, not the oracle)."""
    cw, ch = (w + 1) // 2, (h + 1) // 2
    parts = [planes[0][:h, :w], planes[1][:ch, :cw], planes[2][:ch, :cw]]
    return b"".join(np.ascontiguousarray(p.astype("<u2" if p.dtype != np.uint8 else np.uint8)).tobytes()
                    for p in parts)


def test_frame_md5_of_many_pictures_in_one_call(device):
    import hashlib
    shapes = [(rh.RV_CODEC_HEVC, 856, 480, 8, 854, 480), (rh.RV_CODEC_HEVC, 416, 240, 10, 416, 240),
              (rh.RV_CODEC_H264, 1920, 1088, 8, 1920, 1080), (rh.RV_CODEC_HEVC, 64, 64, 8, 63, 37),
              (rh.RV_CODEC_HEVC, 1280, 720, 10, 1280, 720)]
    sessions, pics, want = [], [], []
    for k in range(24):
        codec, w, h, bd, ow, oh = shapes[k % len(shapes)]
        planes = rh.gen_planes(w, h, bd, rh.subseed(91, k, w, bd))
        s = rh.Session(codec, w, h, bd, 2)
        s.upload(k & 1, planes)
        sessions.append(s)
        pics.append((s, k & 1, ow, oh))
        want.append(hashlib.md5(packed(planes, ow, oh)).digest())
    print(f"frame md5: {len(pics)} pictures in one call", flush=True)
    got = rh.frame_md5(pics)
    assert got == want
    line = rh.frame_line(3, 6, 2, pics[3][2], pics[3][3], got[3])
    assert line == f"3,6,2,63x37,{want[3].hex()}"
    for s in sessions:
        s.close()


def test_frame_md5_refuses_bad_slots_and_sizes(device):
    s = rh.Session(rh.RV_CODEC_HEVC, 64, 64, 8, 2)
    for pic in [(s, 2, 64, 64), (s, 0, 65, 64), (s, 0, 64, 0)]:
        with pytest.raises(rh.RivetError) as e:
            rh.frame_md5([pic])
        assert e.value.code == RV_E_RANGE
    s.close()
