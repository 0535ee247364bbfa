"""Serve-loop scheduling on the CPU: the loop runs against stand-in sessions,
batches and completion markers that check its ordering rules — a picture
is launched only after its references were launched (in an earlier launch
group: the library's one stream runs them first), or completed when the
loop records outputs, and after its session's I picture is resident; each
reference slot still holds the reference when a picture launches (no later
anchor or B picture overwrote it first), no two pictures of a launch write
one slot, and a recorded download returns the picture's own output;
sessions are admitted within the HBM budget and closed once their last
picture completes, and every picture completes."""
import random
import threading
import time

import pytest

from rivet_hip import serve


class World:
    def __init__(self, polls_to_done=2, record=False, prepare_delay_ms=0.0, seed=0, slow_b=False):
        self.record = record
        self.prepare_delay_ms = prepare_delay_ms
        self.slow_b = slow_b  # every B picture's upload takes prepare_delay_ms, anchors none
        self.rng = random.Random(seed)
        self.lock = threading.Lock()
        self.polls_to_done = polls_to_done
        self.marks = {}      # id -> polls left
        self.open = {}       # id(FakeSession) of the open sessions
        self.closed = 0
        self.launched = []
        self.by_gen = {}     # id(gen) -> Pic
        self.errors = []

    def check(self, cond, msg):
        if not cond:
            self.errors.append(msg)


def install(monkeypatch, world, sessions):
    by_planes = {}
    for s in sessions:
        by_planes[id(s.iplanes)] = s.pics[0]
        for p in s.pics:
            if p.gen is not None:
                world.by_gen[id(p.gen)] = p

    class FakeJob:
        def __init__(self, pic):
            self.pic = pic
            self.closed = False

        def close(self):
            world.check(not self.closed, "job closed twice")
            self.closed = True

    class FakeSession:
        @staticmethod
        def footprint(codec, w, h, bd, slots):
            return w * h * 3 // 2 * slots

        def __init__(self, codec, w, h, bd, slots):
            self.resident = False
            self.writer = {}  # slot -> the picture whose output it holds (in stream order)
            with world.lock:
                world.open[id(self)] = 0

        def upload(self, slot, planes):
            world.check(slot == serve.ANCHOR_SLOTS[0], "I picture uploaded to another slot")
            self.writer[slot] = by_planes[id(planes)]
            self.resident = True

        def prepare(self, gen, slot):
            p = world.by_gen[id(gen)]
            world.check(slot == p.slot, "work list prepared for another slot")
            if world.slow_b:
                time.sleep(world.prepare_delay_ms / 1e3 if p.kind == "B" else 0.0)
            elif world.prepare_delay_ms:
                with world.lock:
                    d = world.rng.uniform(0.0, world.prepare_delay_ms)
                time.sleep(d / 1e3)
            return FakeJob(p)

        def download(self, slot):
            return self.writer.get(slot)

        def close(self):
            with world.lock:
                del world.open[id(self)]
                world.closed += 1

    class FakeBatch:
        def __init__(self, jobs):
            self.pics = [j.pic for j in jobs]

        def launch(self, mode):
            # Every picture of a launch reads its references (stage 0) before
            # any writes its slot (the last stage), so reads are checked
            # against the slots as they were before this launch.
            written = set()
            for p in self.pics:
                sess = p.sess.session
                world.check(sess.resident, f"session {p.sess.idx} launched before its I picture")
                for r, sl in zip(p.refs, p.ref_slots):
                    ref = next(q for q in p.sess.pics if q.poc == r)
                    if world.record:
                        world.check(ref.done, f"poc {p.poc} launched before reference {r} completed")
                    else:
                        world.check(ref.done or ref.launched, f"poc {p.poc} launched before reference {r}")
                    w = sess.writer.get(sl)
                    world.check(w is ref, f"session {p.sess.idx} poc {p.poc} read slot {sl} holding poc "
                                          f"{None if w is None else w.poc} instead of reference {r}")
                world.check(not p.launched, f"poc {p.poc} launched twice")
                key = (id(sess), p.slot)
                world.check(key not in written, f"two pictures of one launch write slot {p.slot}")
                written.add(key)
            for p in self.pics:
                p.sess.session.writer[p.slot] = p
            world.launched.extend(self.pics)

        def close(self):
            pass

    counter = iter(range(1, 1 << 30))

    def mark():
        i = next(counter)
        world.marks[i] = world.polls_to_done
        return i

    def mark_done(i):
        world.marks[i] -= 1
        return world.marks[i] <= 0

    def mark_wait(i):
        world.marks[i] = 0

    monkeypatch.setattr(serve, "Session", FakeSession)
    monkeypatch.setattr(serve, "Batch", FakeBatch)
    monkeypatch.setattr(serve, "mark", mark)
    monkeypatch.setattr(serve, "mark_done", mark_done)
    monkeypatch.setattr(serve, "mark_wait", mark_wait)
    monkeypatch.setattr(serve, "sync", lambda: None)


@pytest.fixture(scope="module")
def sessions():
    return serve.build_sessions(6, sizes=[(256, 128), (416, 240)], gop=9, threads=4)


@pytest.mark.parametrize("mode", ["batched", "plain"])
@pytest.mark.parametrize("inflight", [1, 2])
def test_every_picture_completes_in_dependency_order(monkeypatch, sessions, mode, inflight):
    w = World()
    install(monkeypatch, w, sessions)
    st = serve.run(sessions, mode, inflight_max=inflight, upload_threads=3)
    assert not w.errors, w.errors[:5]
    total = sum(len(s.pics) - 1 for s in sessions)
    assert st["pictures"] == total and len(w.launched) == total
    assert w.closed == len(sessions) and not w.open
    assert all(p.done for s in sessions for p in s.pics)
    assert st["latency_ms_p50"] >= 0 and st["latency_ms_max"] >= st["latency_ms_p99"]
    if mode == "plain":
        assert st["mean_batch"] == 1.0


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_recording_waits_for_completed_references(monkeypatch, sessions, mode):
    w = World(record=True)
    install(monkeypatch, w, sessions)
    st = serve.run(sessions, mode, inflight_max=2, record=True, upload_threads=3)
    assert not w.errors, w.errors[:5]
    assert st["pictures"] == sum(len(s.pics) - 1 for s in sessions)
    assert all(p.out is p for s in sessions for p in s.pics if p.kind != "I")


@pytest.mark.parametrize("record", [False, True])
@pytest.mark.parametrize("mode", ["batched", "plain"])
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_slots_are_not_overwritten_before_they_are_read(monkeypatch, sessions, mode, record, seed):
    # Work lists finish uploading out of order: every B picture 30 ms late
    # (seed 0), or each 0-8 ms at random. Later anchors are then ready
    # before the B pictures that read the slot they overwrite, and the
    # loop must hold them back.
    w = World(record=record, prepare_delay_ms=30.0 if seed == 0 else 8.0, seed=seed, slow_b=seed == 0)
    install(monkeypatch, w, sessions)
    st = serve.run(sessions, mode, inflight_max=2, record=record, upload_threads=4)
    assert not w.errors, w.errors[:5]
    assert st["pictures"] == sum(len(s.pics) - 1 for s in sessions)
    if record:
        assert all(p.out is p for s in sessions for p in s.pics if p.kind != "I")


def test_references_launch_ahead_of_completion(monkeypatch, sessions):
    # With two launch groups outstanding, some dependent launches before its
    # reference completes (the stream orders them); never before it launches.
    w = World(polls_to_done=50)
    install(monkeypatch, w, sessions)
    early = []
    orig = serve.Batch.launch

    def launch(self, mode):
        for p in self.pics:
            if any(not next(q for q in p.sess.pics if q.poc == r).done for r in p.refs):
                early.append(p)
        orig(self, mode)

    monkeypatch.setattr(serve.Batch, "launch", launch)
    st = serve.run(sessions, "batched", inflight_max=2, upload_threads=3, preload=True)
    assert not w.errors, w.errors[:5]
    assert st["pictures"] == sum(len(s.pics) - 1 for s in sessions)
    assert early


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_preloaded_run_prepares_everything_first(monkeypatch, sessions, mode):
    w = World()
    install(monkeypatch, w, sessions)
    st = serve.run(sessions, mode, upload_threads=3, preload=True)
    assert not w.errors, w.errors[:5]
    assert st["pictures"] == sum(len(s.pics) - 1 for s in sessions)
    assert st["mode"].startswith(mode) and "prepared" in st["mode"]
    assert w.closed == len(sessions) and not w.open


def test_admission_respects_the_hbm_budget(monkeypatch, sessions):
    w = World()
    install(monkeypatch, w, sessions)
    fps = [serve.footprint(s) for s in sessions]
    budget = max(fps) + min(fps) - 1  # never room for the two largest together
    seen = []
    orig_init = serve.Session.__init__

    def init(self, *a):
        orig_init(self, *a)
        seen.append(len(w.open))

    monkeypatch.setattr(serve.Session, "__init__", init)
    st = serve.run(sessions, "batched", budget=budget, upload_threads=2)
    assert not w.errors, w.errors[:5]
    assert st["pictures"] == sum(len(s.pics) - 1 for s in sessions)
    # With this budget at most two sessions are ever open together.
    assert max(seen) <= 2


def test_a_session_larger_than_the_budget_runs_alone(monkeypatch, sessions):
    w = World()
    install(monkeypatch, w, sessions)
    opened = []
    orig_init = serve.Session.__init__

    def init(self, *a):
        orig_init(self, *a)
        opened.append(len(w.open))

    monkeypatch.setattr(serve.Session, "__init__", init)
    st = serve.run(sessions, "batched", budget=1, upload_threads=2)
    assert not w.errors, w.errors[:5]
    assert st["pictures"] == sum(len(s.pics) - 1 for s in sessions)
    assert max(opened) == 1  # admitted one at a time, each when nothing else is open
