"""Decoded picture hashes computed on the device (rv_picture_hash) against
the oracle's restatement of hash::verify (src/hevc/hash.rs:109-158): on
uploaded planes at several sizes and depths, and on the output slot of
reconstructed HEVC and H.264 pictures, hashed right after the launch so the
call's own ordering after the slot's writer is what the test sees."""
import ctypes as C

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
