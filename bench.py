#!/usr/bin/env python3
"""Throughput of the HIP picture-reconstruction path on one MI355X.

Workload (default): 8 independent HEVC Main 8-bit 4:2:0 streams at
3840x2160, one P- or B-picture each (50/50 by stream seed), reconstructed
by one batched launch per stage, replayed from a hipGraph. A step is one
reconstruction of all 8 pictures (motion compensation, inverse transform +
residual, deblocking, SAO) with references, work lists and outputs
resident in HBM. The value is luma Msample/s = 8 * 3840 * 2160 / seconds
per step, reported only if every output plane equals the CPU oracle byte
for byte.

Also reported: per-stage kernel times (HIP events on the library's
stream) against the HBM roofline, the plain one-picture-per-launch path
on the same pictures, and the CPU oracle on every host core (pictures in
parallel, each split by CTB rows and block ranges) as the baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf_
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "rivet-h26x-codecs_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

import numpy as np  # noqa: E402

METRIC = "bit-exact luma Msample/s 1080p/4K batched; p99 picture latency, serve loop"
HBM_PEAK_GBS = 8000.0
STAGES = ["mc", "tx", "dbk_v", "dbk_h", "sao"]
LEVELS_NOTE = "; coefficients as levels with scaling lists, scaled on the device"
INTRA_NOTE = "; intra blocks predicted on the device"


# Slices and tiles that refuse every crossing, seeded per-slice deblocking
# switches and offsets, and a few exempt CUs.
SLICED = dict(n_slices=16, tile_cols=4, tile_rows=3, lf_across_slices=0, lf_across_tiles=0, slice_vary=1,
              exempt_percent=2)
SLICED_NOTE = "; 16 slices in 4 x 3 tiles per picture, no loop filtering across them, 2% exempt CUs"


def workload_note(extra_cfg: dict) -> str:
    return (LEVELS_NOTE * bool(extra_cfg.get("levels")) + INTRA_NOTE * bool(extra_cfg.get("intra_device")) +
            SLICED_NOTE * bool(extra_cfg.get("n_slices")))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FILE = ROOT / "profiles" / "hevc_c3_8x4k_stage0_traffic.json"


def pmc_traffic(stage: str, streams: int, variant: str = ""):
    """(HBM bytes per launch of the stage, its kernels) on this workload,
    from the committed rocprofv3 --pmc summary: FETCH_SIZE and WRITE_SIZE
    per dispatch of each kernel the stage variant `variant` launches, each
    scaled by what the counter reads for a copy kernel of known size, summed
    over the stage's kernels. (None, None) when the summary does not cover
    this stage, variant and workload."""
    try:
        d = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None, None
    if d.get("stage") != stage or streams != 8:
        return None, None
    v = d.get("variants", {}).get(variant)
    if not v:
        return None, None
    return int(v["hbm_bytes_per_launch"]), " + ".join(v["kernels"])


def dbk_segment_bytes(oracle, pics):
    """Algorithmic bytes of the vertical and horizontal deblocking passes
    of HEVC pictures, from the boundary strengths of their work lists (the
    oracle's derivation, outside any timed region): per 4-line luma segment
    with bS > 0, 8 samples read and 6 written per line; per 2-line chroma
    segment of each component (bS 2, edges on the 16-sample luma grid), 4
    read and 2 written per line."""
    vb = hb = 0
    for g in pics:
        if not g.pic.contents.deblock:
            continue
        s = 2 if g.bit_depth > 8 else 1
        ver, hor = oracle.hevc_bs_map(g.pic)
        on16_x = (np.arange(ver.shape[1]) * 4) % 16 == 0
        on16_y = (np.arange(ver.shape[0]) * 4) % 16 == 0
        vb += s * (int(np.count_nonzero(ver)) * 4 * 14 + int(np.count_nonzero((ver == 2) & on16_x[None, :])) * 2 * 2 * 6)
        hb += s * (int(np.count_nonzero(hor)) * 4 * 14 + int(np.count_nonzero((hor == 2) & on16_y[:, None])) * 2 * 2 * 6)
    return vb, hb


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def host_cores() -> int:
    """Cores this process may use: its CPU affinity, capped by the job's
    thread allowance (OMP_NUM_THREADS) where one is set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(cap))) if cap.isdigit() and int(cap) > 0 else n


def cpu_baseline(oracle, pics, refs, cores: int, target_core_s: float = 10.0):
    """The oracle on every host core: pictures in parallel, each picture's
    work split over cores // pictures threads by CTB rows and block ranges
    (the reference's frame and row threading). Repeated until about
    `target_core_s` core-seconds of work. Returns (outputs of the first
    round, seconds per round, rounds, threads per picture, seconds for one
    picture on one thread)."""
    t0 = time.perf_counter()
    oracle.reconstruct(0, pics[0].pic, refs[0] + [None], 0)
    one = time.perf_counter() - t0
    par = min(len(pics), cores)
    inner = max(1, cores // par)
    rounds = max(1, int(round(target_core_s / (one * len(pics)))))

    def run(i):
        return oracle.reconstruct(0, pics[i].pic, refs[i] + [None], 0, threads=inner)

    with cf_.ThreadPoolExecutor(max_workers=par) as ex:
        first = list(ex.map(run, range(len(pics))))  # untimed: warms the pool and the pages
        t0 = time.perf_counter()
        for _ in range(rounds):
            list(ex.map(run, range(len(pics))))
        wall = (time.perf_counter() - t0) / rounds
    return first, wall, rounds, inner, one


def build_workload(rh, cf, streams: int, extra_cfg: dict):
    pics, refs, sessions, jobs = [], [], [], []
    for k in range(streams):
        cfg, seed = cf.c3_stream(k)
        cfg = dict(cfg, **extra_cfg)
        g = rh.GenPicture(rh.RV_CODEC_HEVC, cfg, seed)
        r = [rh.gen_planes(cfg["width"], cfg["height"], cfg["bit_depth"], cf.ref_seed(seed, s)) for s in range(2)]
        s = rh.Session(rh.RV_CODEC_HEVC, cfg["width"], cfg["height"], cfg["bit_depth"], 3)
        for slot in range(2):
            s.upload(slot, r[slot])
        pics.append(g)
        refs.append(r)
        sessions.append(s)
        jobs.append(s.prepare(g, 2))
    return pics, refs, sessions, jobs


def measure_config(rh, cf, oracle, config: int, n: int, steps: int, warmup: int, extra_cfg: dict):
    """Configuration 2 (HEVC Main10 1080p B, explicit WP, intra with DST) or
    4 (H.264 High 1080p B, implicit weights) as a batched workload: n pictures
    of the configuration (the first is the configuration's own seed), one
    graph launch per step, every plane checked against the oracle."""
    codec, cfg, seed0 = cf.single(config)
    if codec != rh.RV_CODEC_HEVC:
        extra_cfg = {k: v for k, v in extra_cfg.items() if k not in SLICED}  # HEVC slices and tiles only
    cfg = dict(cfg, **extra_cfg)
    w, h, bd = cfg["width"], cfg["height"], cfg.get("bit_depth", 8)
    pics, refs, sessions, jobs = [], [], [], []
    for k in range(n):
        seed = seed0 if k == 0 else rh.subseed(seed0, 11, k, 0)
        g = rh.GenPicture(codec, cfg, seed)
        r = [rh.gen_planes(w, h, bd, cf.ref_seed(seed, slot)) for slot in range(2)]
        s = rh.Session(codec, w, h, bd, 3)
        for slot in range(2):
            s.upload(slot, r[slot])
        pics.append(g)
        refs.append(r)
        sessions.append(s)
        jobs.append(s.prepare(g, 2))
    batch = rh.Batch(jobs)
    batch.launch(rh.RV_LAUNCH_GRAPH)
    rh.sync()
    ok = True
    for g, r, s in zip(pics, refs, sessions):
        got, want = s.download(2), oracle.reconstruct(codec, g.pic, r + [None], 0)
        ok = ok and all(np.array_equal(got[c], want[c]) for c in range(3))
    t = time_steps(rh, batch, rh.RV_LAUNCH_GRAPH, steps, warmup)
    stage_ms = batch.profile(10)
    nb = {"mc": 0, "tx": 0, "dbk": 0, "sao": 0}
    for j in jobs:
        for key, v in j.bytes().items():
            nb[key] += v
    if codec == rh.RV_CODEC_HEVC:
        names = STAGES
        vb, hb = dbk_segment_bytes(oracle, pics)
        stage_bytes = [nb["mc"], nb["tx"], vb, hb, nb["sao"]]
    else:
        # H.264: stage 2 derives the strengths and edge parameters (side data
        # only); stage 3, the macroblock wavefront, moves every deblocking
        # byte; there is no stage 4.
        names = ["mc", "tx", "dbk_strength", "dbk_wavefront", "none"]
        stage_bytes = [nb["mc"], nb["tx"], 0, nb["dbk"], 0]
    luma = n * cf.visible_luma(w, h)
    out = {
        "workload": f"{n} x config {config} ({'HEVC' if codec == rh.RV_CODEC_HEVC else 'H.264'} {w}x{h}, "
                    f"{bd}-bit), batched, hipGraph replay" + workload_note(extra_cfg),
        "bit_exact": ok,
        "ms_per_step": round(t * 1e3, 4),
        "luma_msamples_per_s": round(luma / t / 1e6, 1) if ok else None,
        "stage_ms": {k: round(v, 4) for k, v in zip(names, stage_ms)},
        "stage_gbs": {k: round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 and b > 0 else None
                      for k, b, ms in zip(names, stage_bytes, stage_ms)},
    }
    batch.close()
    for j, s in zip(jobs, sessions):
        j.close()
        s.close()
    return out


def time_steps(rh, batch, mode, steps, warmup):
    for _ in range(warmup):
        batch.launch(mode)
    rh.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.launch(mode)
    rh.sync()
    return (time.perf_counter() - t0) / steps


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--streams", type=int, default=8)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host cores for the CPU baseline (default: this process's CPU share)")
    ap.add_argument("--probe-only", action="store_true",
                    help="run only the 256 MiB copy probe once (counter calibration) and exit")
    ap.add_argument("--no-serve", action="store_true", help="skip the serve-loop measurement (config 5)")
    ap.add_argument("--no-configs", action="store_true", help="skip the batched configs 2 and 4")
    ap.add_argument("--only-config", type=int, choices=(2, 4), default=0,
                    help="measure only 8 pictures of configuration 2 or 4, batched (for kernel profiling), "
                         "print its JSON and exit")
    ap.add_argument("--levels", action="store_true",
                    help="generate coefficients as levels with seeded scaling lists, scaled on the device in "
                         "stage 1 (same pictures' structure; the coefficient values differ)")
    ap.add_argument("--device-intra", action="store_true",
                    help="predict the intra blocks of configurations 2 and 4 on the device (RV_INTRA_DEVICE: HEVC "
                         "intra CUs of every size, H.264 intra macroblocks of every kind) instead of taking "
                         "host-supplied prediction")
    ap.add_argument("--sliced", action="store_true",
                    help="cut every picture into 16 slices in 4 x 3 tiles with no loop filtering across them, "
                         "seeded per-slice deblocking switches and offsets and 2%% exempt CUs (HEVC pictures)")
    ap.add_argument("--graph-only", action="store_true",
                    help="time only the graph replay (for kernel profiling: every stage launch is a batched one)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU path; "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import oracle
    import rivet_hip as rh
    from rivet_hip import configs as cf

    rh.init(0)
    if args.probe_only:
        gbs = rh.probe_copy_gbs(256 << 20, 1)
        log(f"[bench] copy probe: 256 MiB read + 256 MiB written, {gbs:.1f} GB/s")
        return 0
    extra_cfg = dict(levels=1, scaling=1) if args.levels else {}
    if args.device_intra:
        extra_cfg["intra_device"] = 1
    if args.sliced:
        extra_cfg.update(SLICED)
    if args.only_config:
        r = measure_config(rh, cf, oracle, args.only_config, 8, args.steps, args.warmup, extra_cfg)
        log(f"[bench] config {args.only_config}: {r['workload']}: bit-exact {r['bit_exact']}, {r['ms_per_step']} ms")
        print(json.dumps(r), flush=True)
        return 0 if r["bit_exact"] else 1
    log(f"[bench] tier {rh.tier_name()}; generating {args.streams} x 3840x2160 pictures")
    pics, refs, sessions, jobs = build_workload(rh, cf, args.streams, extra_cfg)
    batch = rh.Batch(jobs)
    luma = sum(cf.visible_luma(g.width, g.height) for g in pics)

    # Correctness first: one graph launch, every plane against the oracle.
    batch.launch(rh.RV_LAUNCH_GRAPH)
    rh.sync()
    got = [s.download(2) for s in sessions]
    cores = args.cpu_threads or host_cores()
    log(f"[bench] oracle on {len(pics)} pictures, {cores} host cores")
    want, cpu_wall, cpu_rounds, cpu_inner, cpu_one = cpu_baseline(oracle, pics, refs, cores)
    bit_exact = all(np.array_equal(got[i][c], want[i][c]) for i in range(len(pics)) for c in range(3))
    log(f"[bench] bit-exact: {bit_exact}")

    # Timed region: GRAPH replay of the batched launches (inputs resident).
    t_graph = time_steps(rh, batch, rh.RV_LAUNCH_GRAPH, args.steps, args.warmup)
    if args.graph_only:
        t_batched = t_plain = t_graph
        stage_ms = batch.profile(2)
    else:
        t_batched = time_steps(rh, batch, rh.RV_LAUNCH_BATCHED, max(10, args.steps // 4), 3)
        t_plain = time_steps(rh, batch, rh.RV_LAUNCH_PLAIN, max(10, args.steps // 4), 3)
        stage_ms = batch.profile(max(10, args.steps // 4))
    log(f"[bench] step {t_graph * 1e3:.3f} ms (graph), {t_batched * 1e3:.3f} (batched), {t_plain * 1e3:.3f} (plain); "
        f"stages ms {['%.3f' % v for v in stage_ms]}")

    copy_gbs = rh.probe_copy_gbs(1 << 30, 10)
    log(f"[bench] achievable copy bandwidth: {copy_gbs:.1f} GB/s")
    nb = {"mc": 0, "tx": 0, "dbk": 0, "sao": 0}
    for j in jobs:
        for k, v in j.bytes().items():
            nb[k] += v
    vb, hb = dbk_segment_bytes(oracle, pics)
    stage_bytes = [nb["mc"], nb["tx"], vb, hb, nb["sao"]]
    stage_gbs = [b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0 for b, ms in zip(stage_bytes, stage_ms)]
    dom = int(np.argmax(stage_ms))
    value = luma / t_graph / 1e6 if bit_exact else None
    chosen = {k: names[v] for k, names, v in zip(STAGES, rh.variant_names(), batch.variants())}
    traffic, traffic_kernel = pmc_traffic(STAGES[dom], args.streams, chosen[STAGES[dom]])
    # Stage 0 (motion compensation) is two kernels launched back to back:
    # the tile kernel (side data, intra copies, blocks whose windows cross a
    # plane edge) and the strip kernel (the other blocks).
    stage_kernels = {0: "hevc_stage0 + hevc_mc_strips"}.get(dom, f"hevc_stage{dom}")
    roofline = {
        "bound": "hbm",
        "kernel": f"{stage_kernels} ({STAGES[dom]}, one stage)",
        "achieved": round(stage_gbs[dom], 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(stage_gbs[dom] / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_source": (TRAFFIC_FILE.relative_to(ROOT).as_posix() + f" ({traffic_kernel}: rocprofv3 --pmc "
                           "FETCH_SIZE x2.0 + WRITE_SIZE x1.0 per dispatch, summed over the stage's kernels, "
                           "factors from the copy kernel)") if traffic else None,
        "algorithmic_bytes": int(stage_bytes[dom]),
        "avg_launch_ms": round(stage_ms[dom], 4),
        "achievable_copy_gbs": round(copy_gbs, 1),
    }
    cpu_value = luma / cpu_wall / 1e6
    out = {
        "metric": METRIC,
        "value": round(value, 1) if value is not None else None,
        "unit": "luma Msample/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_graph * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.streams} x HEVC Main 8-bit 4:2:0 3840x2160 P/B pictures (one per stream), "
                        "batched launch per stage, hipGraph replay" + workload_note(extra_cfg),
            "streams": args.streams, "width": 3840, "height": 2160, "bit_depth": 8,
            "pictures_per_step": len(pics),
        },
        "bit_exact": bit_exact,
        "roofline": roofline,
        "cpu_baseline": {
            "value": round(cpu_value, 2),
            "unit": "luma Msample/s",
            "cores": cores,
            "kind": "port",
            "sample": f"the same {len(pics)} 3840x2160 pictures, full reconstruction by the C++ oracle "
                      f"(restatement of the reference's scalar dsp), {cpu_rounds} rounds; {min(len(pics), cores)} "
                      f"pictures at once, each split over {cpu_inner} threads by CTB rows and block ranges, on "
                      f"{cores} cores of {cpu_model()} (os.cpu_count {os.cpu_count()}); "
                      f"1 thread, 1 picture: {luma / len(pics) / cpu_one / 1e6:.2f} Msample/s",
            "one_thread": round(luma / len(pics) / cpu_one / 1e6, 2),
        },
        "extra": {
            "msamples_per_s": {
                "graph": round(luma / t_graph / 1e6, 1),
                "batched": round(luma / t_batched / 1e6, 1),
                "plain": round(luma / t_plain / 1e6, 1),
            },
            "batched_over_plain": round(t_plain / t_graph, 3),
            "stage_ms": {k: round(v, 4) for k, v in zip(STAGES, stage_ms)},
            "stage_gbs": {k: round(v, 1) for k, v in zip(STAGES, stage_gbs)},
            "stage_bytes": {k: int(v) for k, v in zip(STAGES, stage_bytes)},
            "kernel_variants": {k: names[v] for k, names, v in zip(STAGES, rh.variant_names(), batch.variants())},
        },
    }
    if not args.graph_only and not args.no_configs:
        out["extra"]["configs"] = {}
        for config in (2, 4):
            r = measure_config(rh, cf, oracle, config, 8, max(10, args.steps // 4), 3, extra_cfg)
            log(f"[bench] config {config}: {r['workload']}: bit-exact {r['bit_exact']}, {r['ms_per_step']} ms, "
                f"{r['luma_msamples_per_s']} Msample/s; stages ms {r['stage_ms']}")
            out["extra"]["configs"][str(config)] = r
    if not args.graph_only and not args.no_serve:
        from rivet_hip import serve

        for j in jobs:
            j.close()
        jobs = []
        batch.close()
        sessions = None
        log("[bench] serve loop (config 5)")
        sv = serve.experiment(log=log)
        log(f"[bench] serve: batched {sv['batched']['luma_msamples_per_s']} Msample/s p50 "
            f"{sv['batched']['latency_ms_p50']} ms p99 {sv['batched']['latency_ms_p99']} ms; plain "
            f"{sv['plain']['luma_msamples_per_s']} Msample/s p99 {sv['plain']['latency_ms_p99']} ms")
        dv = sv["device"]
        log(f"[bench] serve, work lists prepared in advance: batched {dv['batched']['luma_msamples_per_s']} "
            f"Msample/s, plain {dv['plain']['luma_msamples_per_s']} Msample/s ({dv['batched_over_plain']}x)")
        out["serve_p99_ms"] = sv["batched"]["latency_ms_p99"]
        out["extra"]["serve"] = sv
    print(json.dumps(out), flush=True)
    for j in jobs:
        j.close()
    if jobs:
        batch.close()
    return 0 if bit_exact else 1


if __name__ == "__main__":
    sys.exit(main())
