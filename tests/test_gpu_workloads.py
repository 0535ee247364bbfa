"""The batched workloads at their own shapes, checked byte for byte against
the oracle on the GPU:
- configuration 3: eight independent 3840x2160 8-bit streams (P or B) in
  one batch launched from its captured hipGraph, replayed a second time
  after the output slots were overwritten on the host side;
- configuration 5: the serve loop over sessions at every resolution of its
  mix (854x480 coded 856, 1280x720, 1920x1080, 3840x2160), 8-bit and
  Main10 at each, short IPBB GOPs, outputs recorded as they complete: each
  picture equals the oracle run on the session's own recorded references,
  the plain path gives the same bytes, and each session alone gives the
  same bytes as in the mix."""
import concurrent.futures as cfu
import hashlib

import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import configs as cf
from rivet_hip import serve

pytestmark = pytest.mark.gpu


def _mismatch(got, want, label):
    for c in range(3):
        if not np.array_equal(got[c], want[c]):
            bad = np.argwhere(got[c] != want[c])
            y, x = bad[0]
            return (f"{label} plane {c}: {len(bad)} samples differ, first at (x={x}, y={y}): "
                    f"got {got[c][y, x]} want {want[c][y, x]}")
    return None


def test_config3_eight_4k_streams_in_one_graph_launch(device):
    print("config 3: 8 x 3840x2160 HEVC 8-bit, one batch, graph launch", flush=True)
    streams = []
    for k in range(8):
        cfg, seed = cf.c3_stream(k)
        refs = [rh.gen_planes(3840, 2160, 8, cf.ref_seed(seed, r)) for r in range(2)]
        g = rh.GenPicture(rh.RV_CODEC_HEVC, cfg, seed)
        s = rh.Session(rh.RV_CODEC_HEVC, 3840, 2160, 8, 3)
        for r in range(2):
            s.upload(r, refs[r])
        streams.append((g, refs, s, s.prepare(g, 2)))
    with cfu.ThreadPoolExecutor(8) as ex:
        wants = list(ex.map(lambda t: oracle.reconstruct(rh.RV_CODEC_HEVC, t[0].pic, t[1] + [None], 0), streams))
    assert {t[0].pic.contents.num_ref[1] for t in streams} == {0, 1}  # P and B streams both present
    b = rh.Batch([t[3] for t in streams])
    zero = [np.zeros(sh, np.uint8) for sh in rh.plane_shapes(3840, 2160)]
    for replay in range(2):
        b.launch(rh.RV_LAUNCH_GRAPH)
        rh.sync()
        for k, (t, want) in enumerate(zip(streams, wants)):
            err = _mismatch(t[2].download(2), want, f"graph launch {replay}, stream {k}")
            assert err is None, err
            t[2].upload(2, zero)  # the next replay must write every sample again
    b.close()
    for g, refs, s, j in streams:
        j.close()
        s.close()


SIZES = [(856, 480), (1280, 720), (1920, 1080), (3840, 2160)]


def _digest(planes):
    h = hashlib.sha1()
    for a in planes:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _mix():
    # Sessions i and i + 4 share a resolution, one 8-bit and one Main10.
    return serve.build_sessions(8, sizes=SIZES, depths=[8, 8, 8, 8, 10, 10, 10, 10], gop=6, threads=8)


def test_config5_mix_matches_the_oracle_and_alone_equals_mix(device):
    print("config 5: serve loop, 8 sessions over 4 sizes x 2 depths, GOP 6", flush=True)
    mix = _mix()
    assert {(s.width, s.height, s.bit_depth) for s in mix} == {(w, h, d) for w, h in SIZES for d in (8, 10)}
    stats = serve.run(mix, "batched", inflight_max=2, record=True)
    assert stats["pictures"] == 8 * 5
    assert stats["mean_batch"] > 1.0

    def check(s):
        outs = {0: serve.i_planes(s)}
        for p in s.pics[1:]:
            slots = [None] * serve.DPB_SLOTS
            for r, sl in zip(p.refs, p.ref_slots):
                slots[sl] = outs[r]
            want = oracle.reconstruct(0, p.gen.pic, slots, 0)
            err = _mismatch(p.out, want, f"session {s.idx} {s.width}x{s.height} {s.bit_depth}-bit poc {p.poc}")
            if err:
                return err
            outs[p.poc] = p.out
        return None

    with cfu.ThreadPoolExecutor(8) as ex:
        errs = [e for e in ex.map(check, mix) if e]
    assert not errs, errs[:3]
    digests = {s.idx: [_digest(p.out) for p in s.pics[1:]] for s in mix}
    # The plain path, same trace.
    plain = _mix()
    st = serve.run(plain, "plain", inflight_max=2, record=True)
    assert st["mean_batch"] == 1.0
    for s in plain:
        assert [_digest(p.out) for p in s.pics[1:]] == digests[s.idx], f"plain, session {s.idx}"
    # Each session alone.
    for k in range(len(mix)):
        alone = _mix()[k:k + 1]
        serve.run(alone, "batched", inflight_max=2, record=True)
        assert [_digest(p.out) for p in alone[0].pics[1:]] == digests[k], f"alone, session {k}"
