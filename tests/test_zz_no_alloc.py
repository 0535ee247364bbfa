"""rv_picture_hash_noalloc (the picture hash that allocates nothing) against
rv_picture_hash, through the C ABI: every hash kind on a reconstructed
1920x1080 10-bit picture and then, reusing the same device array, on a
smaller 8-bit one; a picture taller than the array's limit is refused
before any device work."""
import ctypes as C

import pytest

import rivet_hip as rh
from rivet_hip import configs as cf

pytestmark = pytest.mark.gpu

KINDS = [rh.RV_HASH_MD5, rh.RV_HASH_CRC, rh.RV_HASH_CHECKSUM]
RV_E_RANGE = -5
MAX_ROWS = 4352


def _hashes(fn, s, slot):
    out = []
    for k in KINDS:
        buf = (C.c_uint8 * 48)()
        assert fn(s._h, slot, k, buf) == 0, rh.lib.rv_last_error().decode()
        out.append(bytes(buf))
    return out


def _reconstructed(cfg, seed):
    bd = cfg.get("bit_depth", 8)
    refs = [rh.gen_planes(cfg["width"], cfg["height"], bd, cf.ref_seed(seed, k)) for k in range(2)]
    s = rh.Session(rh.RV_CODEC_HEVC, cfg["width"], cfg["height"], bd, 3)
    for k in range(2):
        s.upload(k, refs[k])
    j = s.prepare(rh.GenPicture(rh.RV_CODEC_HEVC, cfg, seed), 2)
    b = rh.Batch([j])
    b.launch(rh.RV_LAUNCH_BATCHED)  # hashed without a host sync first
    return s, j, b


def test_noalloc_hash_equals_pooled_hash_on_two_calls(device):
    for label, cfg, seed in [("1920x1080 10-bit", cf.HEVC_C2, 9101), ("416x240 8-bit", cf.HEVC_C1, 9102)]:
        print(f"allocation-free hash: {label}", flush=True)
        s, j, b = _reconstructed(cfg, seed)
        got = _hashes(rh.lib.rv_picture_hash_noalloc, s, 2)
        want = _hashes(rh.lib.rv_picture_hash, s, 2)
        assert got == want, label
        # The reference slots too (uploaded planes, no launch in between).
        assert _hashes(rh.lib.rv_picture_hash_noalloc, s, 0) == _hashes(rh.lib.rv_picture_hash, s, 0), label
        b.close()
        j.close()
        s.close()


def test_noalloc_hash_refuses_pictures_above_its_limit(device):
    s = rh.Session(rh.RV_CODEC_HEVC, 64, MAX_ROWS + 8, 8, 1)
    buf = (C.c_uint8 * 48)()
    assert rh.lib.rv_picture_hash_noalloc(s._h, 0, rh.RV_HASH_CRC, buf) == RV_E_RANGE
    s.close()
    s = rh.Session(rh.RV_CODEC_HEVC, 64, MAX_ROWS, 8, 1)
    s.upload(0, rh.gen_planes(64, MAX_ROWS, 8, 9103))
    assert _hashes(rh.lib.rv_picture_hash_noalloc, s, 0) == _hashes(rh.lib.rv_picture_hash, s, 0)
    s.close()
