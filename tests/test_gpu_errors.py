"""Error paths that the host forces, with every kernel finishing normally.

The H.264 deblocking wavefront hands strips between bands through polled
exchange words; a poll that reaches its bound goes on (the workgroup must
not hang) but is counted on the device, and the next synchronising call
(rv_sync, rv_mark_wait, rv_plane_download) reports it once as
RV_E_DEVICE, as the reference turns a lost dependency into an error
(Progress.error, src/threading.rs:30-31). A job prepared while
RIVET_HIP_TEST_EXCHANGE_MISS=1 is set awaits tags that are never published
and gives up after 16 polls."""
import ctypes as C
import os

import numpy as np
import pytest

import rivet_hip as rh
from rivet_hip import _abi
from rivet_hip import configs as cf

pytestmark = pytest.mark.gpu


def _picture(force_miss):
    # 352x288: 18 macroblock rows, two deblocking bands (one exchange).
    cfg = dict(cf.H264_C4, width=352, height=288)
    seed = 4242
    refs = [rh.gen_planes(352, 288, 8, cf.ref_seed(seed, k)) for k in range(2)]
    s = rh.Session(rh.RV_CODEC_H264, 352, 288, 8, 3)
    for k in range(2):
        s.upload(k, refs[k])
    g = rh.GenPicture(rh.RV_CODEC_H264, cfg, seed)
    if force_miss:
        os.environ["RIVET_HIP_TEST_EXCHANGE_MISS"] = "1"
    try:
        j = s.prepare(g, 2)
    finally:
        os.environ.pop("RIVET_HIP_TEST_EXCHANGE_MISS", None)
    return s, j, rh.Batch([j])


def _download_rc(s):
    planes = [np.empty(sh, np.uint8) for sh in rh.plane_shapes(352, 288)]
    return rh.lib.rv_plane_download(s._h, 2, *(p.ctypes.data for p in planes))


def test_exchange_give_up_is_reported_once(device):
    print("forced exchange misses: H.264 352x288", flush=True)
    assert rh.lib.rv_sync() == 0
    s, j, b = _picture(force_miss=True)
    b.launch(rh.RV_LAUNCH_BATCHED)
    assert rh.lib.rv_sync() == _abi.RV_E_DEVICE
    assert b"exchange" in rh.lib.rv_last_error()
    assert rh.lib.rv_sync() == 0  # reported once
    b.launch(rh.RV_LAUNCH_BATCHED)
    assert _download_rc(s) == _abi.RV_E_DEVICE
    b.launch(rh.RV_LAUNCH_BATCHED)
    m = C.c_int32()
    assert rh.lib.rv_mark(C.byref(m)) == 0
    assert rh.lib.rv_mark_wait(m.value) == _abi.RV_E_DEVICE
    b.close()
    j.close()
    # The same picture prepared normally: no error, and the output is good.
    s2, j2, b2 = _picture(force_miss=False)
    b2.launch(rh.RV_LAUNCH_BATCHED)
    assert rh.lib.rv_sync() == 0
    assert _download_rc(s2) == 0
    for x in (b2, j2, s2, s):
        x.close()
