"""Continuous-batching serve loop over decode sessions (config 5).

Sessions arrive by a Poisson process. Each is admitted when its device
footprint (DPB planes + the work lists and scratch of its pictures) fits
the HBM budget, else it waits in FIFO order. An admitted session's DPB
stays resident: its I picture is uploaded once, and every P/B picture is
reconstructed on the device from references already there. A picture is
*ready* when its session is admitted and all its references are complete;
the loop launches every ready picture of every session together (one
batched launch per stage), keeps up to two batches in flight, and records
per-picture latency from ready to complete. The plain path runs the same
trace with one launch sequence per picture.

GOP (decode order): I0 P3 B1 B2 P6 B4 B5 ... P27 B25 B26 P29 B28. P
pictures reference the previous anchor, B pictures the anchors around
them. DPB slots: three anchors in rotation and two B slots.

Work lists are generated before the timed run (they stand for the
parser's output, produced upstream of this path); uploading them to the
device happens inside the loop, at admission.
"""
from __future__ import annotations

import collections
import concurrent.futures as cf_
import heapq
import itertools
import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import (RV_CODEC_HEVC, RV_LAUNCH_BATCHED, RV_LAUNCH_PLAIN, Batch, GenPicture, Job, Session, gen_planes, mark,
               mark_done, mark_wait, subseed, sync)
from . import configs as cf

RESOLUTIONS = [(856, 480), (1280, 720), (1920, 1080), (3840, 2160)]  # 854x480 coded 856 wide
GOP = 30
ANCHOR_SLOTS = (0, 1, 2)
B_SLOTS = (3, 4)
DPB_SLOTS = 5


def gop_order(n: int = GOP):
    """(poc, kind, refs) in decode order; refs are POCs (L0, L1)."""
    anchors = list(range(0, n - 2, 3)) + [n - 1]
    out = [(0, "I", ())]
    for a0, a1 in zip(anchors, anchors[1:]):
        out.append((a1, "P", (a0,)))
        for b in range(a0 + 1, a1):
            out.append((b, "B", (a0, a1)))
    return out


@dataclass(eq=False)  # pictures are compared and hashed by identity
class Pic:
    sess: "Sess"
    poc: int
    kind: str
    refs: tuple
    slot: int
    ref_slots: tuple
    gen: GenPicture | None = None
    job: Job | None = None
    ready_t: float = math.nan
    done_t: float = math.nan
    launched: bool = False
    queued: bool = False  # in the loop's heap of launchable pictures
    done: bool = False
    luma: int = 0
    out: list | None = None
    fut: object = None
    dependents: list = field(default_factory=list)  # pictures referencing this one
    # Earlier pictures (decode order) of the session that write this
    # picture's slot, or read it as a reference: this picture may overwrite
    # the slot only after they have launched (completed, for writers, when
    # outputs are recorded).
    prior_writers: list = field(default_factory=list)
    prior_readers: list = field(default_factory=list)
    slot_successors: list = field(default_factory=list)  # pictures waiting on this one's use of a slot
    ref_pics: list = field(default_factory=list)  # the references, as pictures
    # The loop's counters: pictures this one still waits for (to launch, or
    # to complete), references not yet complete, and who waits for this one.
    pending: int = 0
    refs_pending: int = 0
    started: bool = False  # launched, or (the I picture) resident
    on_start: list = field(default_factory=list)
    on_done: list = field(default_factory=list)


@dataclass
class Sess:
    idx: int
    width: int
    height: int
    bit_depth: int
    seed: int
    arrival: float
    pics: list = field(default_factory=list)
    session: Session | None = None
    admitted_t: float = math.nan
    footprint: int = 0
    remaining: int = 0
    iplanes: list | None = None  # the host-decoded I picture (input data)


def session_cfgs(n: int, seed: int):
    """Resolution uniform over the four sizes; 70 % Main 8-bit, 30 % Main10."""
    rng = np.random.default_rng(subseed(seed, 5, 0, 0) & 0xFFFFFFFF)
    out = []
    for i in range(n):
        w, h = RESOLUTIONS[int(rng.integers(0, 4))]
        bd = 8 if rng.random() < 0.7 else 10
        out.append((w, h, bd, subseed(seed, 5, i, 1)))
    return out


def _pic_cfg(w, h, bd, kind, poc, ref_pocs, ref_slots):
    base = dict(cf.HEVC_C1, width=w, height=h, bit_depth=bd)
    if kind == "P":
        base.update(b_picture=0, poc=poc, ref_poc=[ref_pocs[0], ref_pocs[0]], ref_slot=[ref_slots[0], ref_slots[0]])
    else:
        base.update(b_picture=1, bi_percent=60, poc=poc, ref_poc=list(ref_pocs), ref_slot=list(ref_slots))
    return base


def build_sessions(n: int = 64, seed: int = cf.BASE_SEED, rate: float = 0.0, gop: int = GOP, threads: int = 16,
                   sizes=None, depths=None):
    """Sessions with generated work lists. rate = 0: all arrive at t = 0.
    sizes / depths, when given, replace the drawn resolution / bit depth of
    session i by entry i modulo their length."""
    specs = session_cfgs(n, seed)
    if sizes is not None:
        specs = [(sizes[i % len(sizes)][0], sizes[i % len(sizes)][1], s[2], s[3]) for i, s in enumerate(specs)]
    if depths is not None:
        specs = [(s[0], s[1], depths[i % len(depths)], s[3]) for i, s in enumerate(specs)]
    rng = np.random.default_rng(subseed(seed, 5, 0, 2) & 0xFFFFFFFF)
    t = 0.0
    sessions = []
    for i, (w, h, bd, sseed) in enumerate(specs):
        if rate > 0:
            t += float(rng.exponential(1.0 / rate))
        s = Sess(i, w, h, bd, sseed, t)
        slot_of = {}
        anchor_k = 0
        b_k = 0
        for poc, kind, refs in gop_order(gop):
            if kind == "B":
                slot = B_SLOTS[b_k % 2]
                b_k += 1
            else:
                slot = ANCHOR_SLOTS[anchor_k % 3]
                anchor_k += 1
            slot_of[poc] = slot
            p = Pic(s, poc, kind, refs, slot, tuple(slot_of[r] for r in refs))
            p.luma = cf.visible_luma(w, h)
            s.pics.append(p)
        by_poc = {q.poc: q for q in s.pics}
        for i, q in enumerate(s.pics):
            q.ref_pics = [by_poc[r] for r in q.refs]
            for r in q.refs:
                by_poc[r].dependents.append(q)
            for e in s.pics[:i]:
                if e.slot == q.slot:
                    q.prior_writers.append(e)
                elif q.slot in e.ref_slots:
                    q.prior_readers.append(e)
                else:
                    continue
                e.slot_successors.append(q)
        sessions.append(s)

    def gen(p: Pic):
        if p.kind == "I":
            return None
        s = p.sess
        cfg = _pic_cfg(s.width, s.height, s.bit_depth, p.kind, p.poc, p.refs, p.ref_slots)
        return GenPicture(RV_CODEC_HEVC, cfg, subseed(s.seed, 7, p.poc, 0))

    todo = [p for s in sessions for p in s.pics if p.kind != "I"]
    with cf_.ThreadPoolExecutor(max_workers=threads) as ex:
        for p, g in zip(todo, ex.map(gen, todo)):
            p.gen = g
        for s, planes in zip(sessions, ex.map(i_planes, sessions)):
            s.iplanes = planes
    return sessions


def i_planes(s: Sess):
    return gen_planes(s.width, s.height, s.bit_depth, subseed(s.seed, 7, 0, 1))


def footprint(s: Sess) -> int:
    """DPB + per picture: reconstruction scratch and its work lists."""
    dpb = Session.footprint(RV_CODEC_HEVC, s.width, s.height, s.bit_depth, DPB_SLOTS)
    per = 0
    for p in s.pics:
        if p.gen is None:
            continue
        c = p.gen.pic.contents
        per += c.n_nz * 4 + c.n_tus * 16 + c.n_pus * 40 + c.n_cus * 12
    ss = 2 if s.bit_depth > 8 else 1
    per += (len(s.pics) - 1) * (s.width * s.height * 3 // 2 * ss + s.width * s.height // 2)
    return dpb + per


def run(sessions, mode: str = "batched", budget: int = 32 << 30, max_batch: int = 64, inflight_max: int = 2,
        record: bool = False, upload_threads: int = 8, preload: bool = False):
    """Serve every session; returns throughput and latency statistics.
    record=True downloads every reconstructed picture as it completes (into
    Pic.out); a later picture that writes the same DPB slot launches only
    after that download.
    Work-list uploads run on `upload_threads` host threads from admission
    on, overlapping the reconstruction of earlier pictures. preload=True
    opens every session and prepares every work list before the clock
    starts (arrivals and the budget are ignored): the loop then measures the
    device side alone, scheduling and launching."""
    ex = cf_.ThreadPoolExecutor(max_workers=upload_threads)
    try:
        return _run(sessions, mode, budget, max_batch, inflight_max, record, ex, preload)
    finally:
        ex.shutdown(wait=True)
        sync()


def _run(sessions, mode, budget, max_batch, inflight_max, record, ex, preload=False):
    """The loop is event-driven: finished uploads (posted by the upload
    threads) and launched references feed a heap of launchable pictures.
    Every launch goes to the library's one in-order stream, so a picture
    may launch as soon as its references have launched: the stream runs it
    after them, and the next batch is queued while the last one runs. It
    also waits until every earlier picture of its session that uses its
    DPB slot (writes it, or reads it as a reference) has launched in an
    earlier group, so no slot is overwritten before it is read, and two
    writers of a slot never share a launch. At most `inflight_max` launch
    groups (one loop iteration's launches) are outstanding in either mode.
    A picture's latency runs from when it was ready (references complete,
    work list uploaded) to its completion. With record=True a picture
    launches only after its references and the earlier writers of its slot
    have completed, and finished groups are downloaded before the next
    launch, so each download sees its picture's own output."""
    t0 = time.perf_counter()
    arrivals = collections.deque(sorted(sessions, key=lambda s: s.arrival))
    waiting = collections.deque()
    active = {}  # id(session) -> session
    used = 0
    inflight = collections.deque()
    total = sum(len(s.pics) - 1 for s in sessions)
    done = 0
    lat = []
    batches = 0
    batch_sizes = []
    launchable = []  # heap of (ready_t, seq, pic): ready and uploaded
    seq = itertools.count()
    uploaded = collections.deque()  # appended by upload threads
    wake = threading.Event()
    groups = collections.Counter()  # launch group -> its inflight entries
    gid = itertools.count()
    for s in sessions:
        s.footprint = footprint(s)
        s.remaining = len(s.pics) - 1
        for p in s.pics:
            p.launched = p.done = p.queued = p.started = False
            p.ready_t = p.done_t = math.nan
            p.job = p.fut = None
            p.on_start, p.on_done = [], []
    # What each picture waits for: its references and the earlier users of
    # its slot must have launched (references and earlier writers must have
    # completed when recording). Counted once here, then decremented as
    # pictures launch and complete.
    for s in sessions:
        for q in s.pics:
            need_done = set(q.ref_pics) | set(q.prior_writers) if record else set()
            waits = list(dict.fromkeys(q.ref_pics + q.prior_writers + q.prior_readers))
            q.pending = len(waits)
            q.refs_pending = len(q.ref_pics)
            for e in waits:
                (e.on_done if e in need_done else e.on_start).append(q)

    def now():
        return time.perf_counter() - t0

    def on_uploaded(p: Pic):
        def cb(_fut):
            uploaded.append(p)
            wake.set()
        return cb

    def eligible(q: Pic) -> bool:
        """References launched (completed when recording: the download at
        their completion must see their own output), and the slot free:
        every earlier picture that writes q's slot, or reads it as a
        reference, launched in an earlier group, so the stream runs it
        first. When recording, earlier writers must also have completed,
        since their download happens at completion."""
        return q.job is not None and q.pending == 0 and not q.launched and not q.queued

    def push(q: Pic, t: float):
        q.queued = True
        heapq.heappush(launchable, (t, next(seq), q))

    def release_waiters(waiters, t: float):
        for q in waiters:
            q.pending -= 1
            if eligible(q):
                push(q, t)

    def start(p: Pic, t: float):
        if not p.started:
            p.started = True
            release_waiters(p.on_start, t)

    def complete(p: Pic, t: float):
        nonlocal done, used
        p.done = True
        p.done_t = t
        start(p, t)
        release_waiters(p.on_done, t)
        s = p.sess
        if p.kind != "I":
            done += 1
            lat.append(p.done_t - p.ready_t)
            if record:
                p.out = s.session.download(p.slot)
            p.job.close()
            p.job = None
            s.remaining -= 1
        for q in p.dependents:
            q.refs_pending -= 1
            if q.refs_pending == 0:
                q.ready_t = t
        if s.remaining == 0 and id(s) in active:
            s.session.close()
            s.session = None
            used -= s.footprint
            del active[id(s)]

    def release(fin):
        for tc, pics, bs in fin:
            for b in bs:
                b.close()
            for p in pics:
                complete(p, tc)

    def admit(s: Sess, t: float):
        """Open the session and hand its uploads to the loader threads: the
        host-decoded I picture first, then every P/B work list. The I
        picture completes when its planes are resident."""
        nonlocal used
        s.session = Session(RV_CODEC_HEVC, s.width, s.height, s.bit_depth, DPB_SLOTS)
        for p in s.pics:
            if p.gen is None:
                p.fut = ex.submit(s.session.upload, ANCHOR_SLOTS[0], s.iplanes)
            else:
                p.fut = ex.submit(s.session.prepare, p.gen, p.slot)
            p.fut.add_done_callback(on_uploaded(p))
        s.admitted_t = t
        used += s.footprint
        active[id(s)] = s

    if preload:
        while arrivals:
            admit(arrivals.popleft(), 0.0)
        cf_.wait([p.fut for s in sessions for p in s.pics if p.fut is not None])
        t0 = time.perf_counter()

    while done < total:
        wake.clear()
        t = now()
        while arrivals and arrivals[0].arrival <= t:
            waiting.append(arrivals.popleft())
        while waiting and (used + waiting[0].footprint <= budget or not active):
            admit(waiting.popleft(), now())
        # Finished groups free their place in flight now; their batches and
        # jobs are released after the next launch, so that host work
        # overlaps the GPU work queued behind it.
        fin = []
        while inflight and mark_done(inflight[0][0]):
            _, pics, bs, g = inflight.popleft()
            groups[g] -= 1
            if groups[g] == 0:
                del groups[g]
            fin.append((now(), pics, bs))
        if record:
            # A recorded picture is downloaded before any later picture can
            # be launched into its slot.
            release(fin)
            fin = []
        while uploaded:
            p = uploaded.popleft()
            r = p.fut.result()
            p.fut = None
            if p.gen is None:
                p.ready_t = now()
                complete(p, p.ready_t)
                continue
            p.job = r
            if eligible(p):
                push(p, now())
        launched = False
        if launchable and len(groups) < inflight_max:
            group = [heapq.heappop(launchable)[2] for _ in range(min(max_batch, len(launchable)))]
            g = next(gid)
            if mode == "batched":
                # One batch per sample depth class, so each runs the kernel
                # variants tuned for its class.
                bs = [Batch([p.job for p in part]) for part in
                      ([p for p in group if p.sess.bit_depth == 8], [p for p in group if p.sess.bit_depth != 8]) if part]
                for b in bs:
                    b.launch(RV_LAUNCH_BATCHED)
                for p in group:
                    p.launched = True
                inflight.append((mark(), group, bs, g))
                groups[g] += 1
                batches += 1
                batch_sizes.append(len(group))
            else:
                for p in group:
                    b = Batch([p.job])
                    b.launch(RV_LAUNCH_PLAIN)
                    p.launched = True
                    inflight.append((mark(), [p], [b], g))
                    groups[g] += 1
                    batches += 1
                    batch_sizes.append(1)
            # Their dependents, and pictures waiting for their slots, may
            # launch next: the stream orders them.
            tl = now()
            for p in group:
                start(p, tl)
            launched = True
        release(fin)
        if launched or fin:
            continue
        if len(groups) >= inflight_max:
            mark_wait(inflight[0][0])
        elif inflight:
            wake.wait(0.0002)  # an upload, or soon a completion to collect
        elif arrivals and not waiting and not active:
            time.sleep(max(0.0, min(arrivals[0].arrival - now(), 0.01)))
        else:
            wake.wait(0.001 if not arrivals else max(0.0, min(arrivals[0].arrival - now(), 0.001)))
    wall = now()
    luma = sum(p.luma for s in sessions for p in s.pics if p.kind != "I")
    lat_ms = np.array(lat) * 1e3
    return {
        "mode": mode + (" (prepared in advance)" if preload else ""),
        "pictures": total,
        "wall_s": round(wall, 4),
        "luma_msamples_per_s": round(luma / wall / 1e6, 1),
        "latency_ms_p50": round(float(np.percentile(lat_ms, 50)), 3),
        "latency_ms_p99": round(float(np.percentile(lat_ms, 99)), 3),
        "latency_ms_max": round(float(lat_ms.max()), 3),
        "batches": batches,
        "mean_batch": round(float(np.mean(batch_sizes)), 2) if batch_sizes else 0.0,
    }


def experiment(n: int = 64, seed: int = cf.BASE_SEED, budget: int = 16 << 30, load: float = 0.8, threads: int = 16,
               log=print):
    """Config 5: capacity at saturation (every session arriving at t = 0),
    then Poisson arrivals at `load` x capacity, served batched and plain on
    the same trace."""
    t = time.perf_counter()
    sessions = build_sessions(n, seed, rate=0.0, threads=threads)
    log(f"[serve] generated {n} sessions ({sum(len(s.pics) - 1 for s in sessions)} P/B pictures) in "
        f"{time.perf_counter() - t:.1f} s")
    sat = run(sessions, "batched", budget=budget, upload_threads=threads)
    capacity = n / sat["wall_s"]
    rate = load * capacity
    rng = np.random.default_rng(subseed(seed, 5, 0, 3) & 0xFFFFFFFF)
    arr = np.cumsum(rng.exponential(1.0 / rate, size=n))
    for s, a in zip(sessions, arr):
        s.arrival = float(a)
    batched = run(sessions, "batched", budget=budget, upload_threads=threads)
    plain = run(sessions, "plain", budget=budget, upload_threads=threads)
    # The device side alone: every work list prepared before the clock.
    dev_b = run(sessions, "batched", upload_threads=threads, preload=True)
    dev_p = run(sessions, "plain", upload_threads=threads, preload=True)
    return {
        "sessions": n,
        "mix": "uniform over 854x480 (coded 856), 1280x720, 1920x1080, 3840x2160; 70% Main 8-bit / 30% Main10; "
               "GOP 30 IPBB",
        "hbm_budget_gib": budget / (1 << 30),
        "capacity_sessions_per_s": round(capacity, 2),
        "saturation": sat,
        "poisson_rate_sessions_per_s": round(rate, 2),
        "batched": batched,
        "plain": plain,
        "batched_over_plain": round(batched["luma_msamples_per_s"] / plain["luma_msamples_per_s"], 3),
        "device": {"batched": dev_b, "plain": dev_p,
                   "batched_over_plain": round(dev_b["luma_msamples_per_s"] / dev_p["luma_msamples_per_s"], 3)},
    }
