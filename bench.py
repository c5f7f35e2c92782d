#!/usr/bin/env python3
"""Benchmark of the MI355X VP9 reconstruction backend.

Default run (N=1):
  * single stream, 1920x1080 preset 1 (25% dense coefficients, mixed 4..32
    transforms, 30% intra blocks on inter frames), a 100-frame sequence
    prepared in HBM; a step reconstructs one frame through the plain
    per-stage kernels. `value` = frames/s over exactly K timed steps.
  * 64-stream mix (16 each of 640x360, 854x480, 1280x720, 1920x1080), one
    batch of 64 frames per step, one launch per stage: aggregate frames/s
    and 1080p-equivalent frames/s (sum of luma pixels / 1920*1080).
  * roofline of the dominant kernel from HIP events on the engine's stream
    over the timed region, against MI355X HBM peak.
  * cpu_baseline: the reference's own CPU reconstruction (oracle/_ref, one
    thread) on a bounded sample of the same 1080p sequence, with per-frame
    latency percentiles; multi_stream.cpu_baseline_all_cores: the same code
    on the first frames of every stream of the mix, streams dealt over every
    host core the process may use (at most 16).
Prints ONE JSON line.
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
from _pkgload import load_package  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TOPS = 78.6   # MI355X vector FP64 (FMA counted as 2); mul+add separately: half
METRIC = "recon frames/s @1080p (1 vs 64 streams), p50/p99 frame latency, HBM GB/s"
MIX = [(640, 360), (854, 480), (1280, 720), (1920, 1080)]


KERNEL_OF = {"intra": "k_intra", "neighbours": "k_nbr", "mc": "k_mc", "deblock": "k_lf",
             "owner": "k_owner",
             "fill": "k_fill", "nv12": "k_nv12", "residual4": "k_residual<4>",
             "residual8": "k_residual<8>", "residual16": "k_residual<16>",
             "residual32": "k_residual<32>"}
SINGLE_WORKLOAD = ("1920x1080 single stream, 100-frame preset-1 sequence, "
                   "plain per-stage kernels")


def kernel_of(stage, tune):
    """Kernel that runs a stage under a launch shape, named as rocprofv3
    names it (templates on threads per intra worker, transform size and
    waves per transform group)."""
    if stage == "intra" and tune:
        return f"k_intra<{tune['intra_lanes']}>"
    if stage.startswith("residual") and tune:
        n = int(stage[8:])
        return f"k_residual<{n}, {tune['res_split'][(4, 8, 16, 32).index(n)]}>"
    return KERNEL_OF.get(stage, stage)


def measured_traffic(workload, kernel):
    """HBM-side bytes per launch of a kernel from the committed rocprofv3
    counter summary (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            entry = json.load(f)[workload]
        return entry["kernels"][kernel]["traffic_bytes"], entry["profile"]
    except (OSError, KeyError, ValueError):
        return None, None


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def gen_frames(P, specs, workers=16):
    """specs: list of (preset, w, h, seed, index). Generated in parallel
    threads (ctypes releases the GIL)."""
    def one(s):
        preset, w, h, seed, idx = s
        return P.SynthFrame(P.synth_params(preset, w, h, seed), idx)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, specs))


def stage_bytes(P, frame):
    """Algorithmic HBM bytes per stage for one frame (numpy over blocks)."""
    b = frame.blocks()
    W, H = frame.pf.hdr.width, frame.pf.hdr.height
    x = b["x"].astype(np.int64)
    y = b["y"].astype(np.int64)
    bw = np.minimum(b["width"].astype(np.int64), W - x)
    bh = np.minimum(b["height"].astype(np.int64), H - y)
    ok = (bw > 0) & (bh > 0)
    intra = (b["is_intra"] != 0) & ok
    inter = (b["is_intra"] == 0) & ok
    T = 4 << b["tx_size"].astype(np.int64)
    area = np.where(ok & (b["skip"] == 0), (bw // T) * (bh // T) * T * T, 0)
    cbw, cbh = (bw + 1) // 2, (bh + 1) // 2
    out = {}
    out["mc"] = int(((bw + 1) * (bh + 1) + 2 * (cbw + 1) * (cbh + 1) + bw * bh + 2 * cbw * cbh)[inter].sum())
    for k, n in enumerate((4, 8, 16, 32)):
        sel = inter & (b["tx_size"] == k)
        out[f"residual{n}"] = int(4 * area[sel].sum())
    out["intra"] = int((bw * bh + 2 * cbw * cbh)[intra].sum() + 4 * area[intra].sum())
    # neighbour samples gathered and the neighbour record written per intra block
    ncell = (bw + 3) // 4 + (bh + 3) // 4 + 1
    out["neighbours"] = int((bw + bh + 1 + 2 * (cbw + cbh + 1) + 8 * ncell + 40)[intra].sum())
    out["deblock"] = 2 * W * H if frame.pf.hdr.filter_level else 0
    out["fill"] = 0
    out["owner"] = int((32 + ((bw + 3) // 4) * ((bh + 3) // 4) * 4)[ok].sum())
    out["nv12"] = 0
    # FP64 operations of the IDCT (multiply + add per MAC, dense bound)
    out["fp64_ops"] = {n: int(2 * 2 * n * area[inter & (b["tx_size"] == k)].sum())
                       for k, n in enumerate((4, 8, 16, 32))}
    return out


def single_stream(P, args):
    W, H = 1920, 1080
    F = args.frames
    log(f"generating {F} frames {W}x{H} preset 1")
    frames = gen_frames(P, [(1, W, H, 20250, i) for i in range(F)])
    eng = P.Engine(0)
    preps = [eng.prepare(f) for f in frames]
    tune = None
    if not args.no_tune:  # launch shape for this workload, on a tuning stream
        tsid = eng.open_stream()
        eng.run([tsid], [preps[0]])
        tune = eng.autotune([tsid], [preps[1]], iters=5).as_dict()
        eng.close_stream(tsid)
        log("single-stream tune:", tune)
    else:
        tune = eng.get_tune().as_dict()
    sid = eng.open_stream()
    order = [(args.warmup + k) % F for k in range(args.steps)]
    for k in range(args.warmup):
        eng.run([sid], [preps[k % F]])
    eng.sync()
    # untimed-events pass of the same steps: the clean throughput figure
    eng.set_timing(False)
    t0 = time.perf_counter()
    for i in order:
        eng.run([sid], [preps[i]])
    eng.sync()
    clean_s = time.perf_counter() - t0
    # timed region with per-stage HIP events on the engine's stream
    for k in range(args.warmup):
        eng.run([sid], [preps[k % F]])
    eng.sync()
    eng.stage_stats(reset=True)
    eng.set_timing(True)
    t0 = time.perf_counter()
    for i in order:
        eng.run([sid], [preps[i]])
    eng.sync()
    timed_s = time.perf_counter() - t0
    stats = eng.stage_stats(reset=True)
    eng.set_timing(False)
    per_stage = {}
    sb = [stage_bytes(P, frames[i]) for i in range(F)]
    for name, (ms, n) in stats.items():
        if n == 0:
            continue
        tot_bytes = sum(sb[i].get(name, 0) for i in order)
        per_stage[name] = {"ms_total": round(ms, 4), "launches": n,
                           "us_per_launch": round(1e3 * ms / n, 3),
                           "algo_bytes_per_launch": round(tot_bytes / n),
                           "algo_GBps": round(tot_bytes / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
        if name.startswith("residual"):
            N = int(name[8:])
            ops = sum(sb[i]["fp64_ops"][N] for i in order)
            per_stage[name]["fp64_dense_TOPs"] = round(ops / (ms * 1e-3) / 1e12, 2)
    algo = sum(frames[i].algo_bytes() for i in order)
    eng.close()
    lv = [P.intra_levels(f) for f in frames]
    dropin_fps = dropin(P, frames, args.steps)
    return {"clean_s": clean_s, "timed_s": timed_s, "stats": per_stage, "algo_bytes": algo,
            "frames": frames, "tune": tune, "depth": [max(len(x) - 1, 0) for x in lv],
            "width0": [int(x[1] - x[0]) if len(x) > 1 else 0 for x in lv],
            "dropin_fps": dropin_fps}


def dropin(P, frames, steps):
    """The drop-in boundary (hip_decode_frame / hip_get_frame) on the same
    1080p sequence: host-side validation and staging of each parsed frame,
    PCIe upload, reconstruction, I420 copy back to pinned host memory. Host
    and PCIe inclusive, so not `value` (which times inputs resident in HBM)."""
    import ctypes as C
    lib = P.lib()
    with P.HipBackend() as be:
        info = P.FrameInfo()
        done = 0

        def submit(f, i):
            nonlocal done
            while True:
                rc = lib.hip_decode_frame(be.ctx, f.ptr, i)
                if rc != P.CVP9_ERR_AGAIN:
                    P.check(rc, "hip_decode_frame")
                    return
                P.check(lib.hip_get_frame(be.ctx, C.byref(info), 1), "hip_get_frame")
                done += 1

        for i in range(10):
            submit(frames[i % len(frames)], i)
        while lib.hip_get_frame(be.ctx, C.byref(info), 1) == P.CVP9_OK:
            pass
        done = 0
        t0 = time.perf_counter()
        for i in range(steps):
            submit(frames[i % len(frames)], i)
        while lib.hip_get_frame(be.ctx, C.byref(info), 1) == P.CVP9_OK:
            done += 1
        dt = time.perf_counter() - t0
    return steps / dt


def multi_stream(P, args):
    per_res = args.streams // 4
    F = args.mframes
    specs, layout = [], []
    for r, (w, h) in enumerate(MIX):
        for s in range(per_res):
            sid = r * per_res + s
            layout.append((w, h))
            for i in range(F):
                specs.append((3, w, h, 7000 + sid, i))
    log(f"generating {len(specs)} frames for {len(layout)} streams")
    frames = gen_frames(P, specs)
    eng = P.Engine(0)
    preps = [eng.prepare(f) for f in frames]
    del frames
    tune = None
    if args.mtune:  # a fixed launch shape (profiling runs)
        a = args.mtune
        eng.set_tune(a[:4], a[4], a[5], a[6])
        tune = eng.get_tune().as_dict()
    elif not args.no_tune:  # tune on streams opened for it, then drop them
        tsids = [eng.open_stream() for _ in layout]
        eng.run(tsids, [preps[s * F] for s in range(len(tsids))])
        tune = eng.autotune(tsids, [preps[s * F + 1] for s in range(len(tsids))], iters=3).as_dict()
        for t in tsids:
            eng.close_stream(t)
        log("multi-stream tune:", tune)
    sids = [eng.open_stream() for _ in layout]
    steps = args.msteps

    G = max(1, args.mgroups)

    def batch(k, groups):
        """One step: every stream's next frame, as `groups` batches of
        interleaved streams (each group mixes every resolution); the engine
        runs disjoint groups on its two queues, overlapping them."""
        idx = k if k < F else 1 + (k - 1) % (F - 1)   # frame 0 (KEY) once, then cycle
        for g in range(groups):
            mine = [s for s in range(len(sids)) if s % groups == g]
            eng.run([sids[s] for s in mine], [preps[s * F + idx] for s in mine])

    def timed(groups, k0):
        t0 = time.perf_counter()
        for k in range(k0, k0 + steps):
            batch(k, groups)
        eng.sync()
        return time.perf_counter() - t0

    for k in range(args.mwarmup):
        batch(k, G)
    eng.sync()
    dt = timed(G, args.mwarmup)
    dt1 = timed(1, args.mwarmup + steps) if G > 1 else dt
    px = sum(w * h for w, h in layout) * steps
    # per-stage view of one-batch steps, HIP events around each launch
    eng.stage_stats(reset=True)
    eng.set_timing(True)
    for k in range(args.mwarmup, args.mwarmup + steps):
        batch(k, 1)
    stats = {n: {"us_per_launch": round(1e3 * ms / c, 2), "launches": c}
             for n, (ms, c) in eng.stage_stats(reset=True).items() if c}
    eng.set_timing(False)
    eng.close()
    best = min(dt, dt1)  # the faster schedule of the two is the aggregate figure
    return {"streams": len(layout), "steps": steps, "seconds": best, "stages": stats, "tune": tune,
            "groups": G, "schedule_groups": G if dt <= dt1 else 1,
            "frames_per_s": len(layout) * steps / best,
            "p1080_equiv_per_s": px / (1920 * 1080) / best,
            "ms_per_batch": 1e3 * best / steps,
            "groups_p1080_equiv_per_s": px / (1920 * 1080) / dt,
            "one_batch_p1080_equiv_per_s": px / (1920 * 1080) / dt1}


def serving(P, args):
    """Serving loop: Poisson arrivals, HBM-budget admission, 30 fps pacing,
    continuous batching; submit -> NV12-ready latency per frame."""
    rng = np.random.default_rng(4040)
    n = args.sstreams
    res = [MIX[i] for i in rng.integers(0, 4, n)]
    lengths = rng.integers(30, 301, n)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, n))
    pool_n = args.pool
    log(f"serving: {n} streams, {int(lengths.sum())} frames, pools of {pool_n} frames")
    specs = [(4, w, h, 9000 + r, i) for r, (w, h) in enumerate(MIX) for i in range(pool_n)]
    frames = gen_frames(P, specs)
    eng = P.Engine(0)
    if args.serve_tune:
        eng.set_tune(*args.serve_tune)
    pool = {}
    for r, (w, h) in enumerate(MIX):
        pool[(w, h)] = [eng.prepare(frames[r * pool_n + i]) for i in range(pool_n)]
    del frames
    budget = int(args.budget_gib * (1 << 30))
    srv = P.Server(eng, budget, max_batch=4096, max_inflight=2)
    period_ns = int(1e9 / 30)
    handle, start_ns, next_k, done = {}, {}, [0] * n, [0] * n
    submit_ns = {}
    lat = []
    wait_ms = []
    waiting = []
    arrived = 0
    finished = 0
    waited = 0
    max_active = 0
    batches = 0
    t0 = time.monotonic_ns()
    while finished < n:
        now = time.monotonic_ns() - t0
        while arrived < n and arrivals[arrived] * 1e9 <= now:
            waiting.append(arrived)
            arrived += 1
        while waiting:
            s = waiting[0]
            h = srv.open(*res[s])
            if h is None:
                break
            waiting.pop(0)
            handle[s] = h
            start_ns[s] = now
            wait_ms.append((now - arrivals[s] * 1e9) / 1e6)
            if wait_ms[-1] > 1.0:
                waited += 1
        max_active = max(max_active, len(handle))
        for s, h in handle.items():
            while next_k[s] < lengths[s] and start_ns[s] + next_k[s] * period_ns <= now:
                k = next_k[s]
                pl = pool[res[s]]
                prep = pl[0] if k == 0 else pl[1 + (k - 1 + s) % (pool_n - 1)]
                tag = s * 1000 + k
                submit_ns[tag] = time.monotonic_ns()
                srv.submit(h, prep, tag)
                next_k[s] += 1
        if srv.pump() > 0:
            batches += 1
        for tag, dns in srv.poll():
            tag = int(tag)
            lat.append((dns - submit_ns.pop(tag)) / 1e6)
            s = tag // 1000
            done[s] += 1
            if done[s] == lengths[s]:
                srv.close(handle.pop(s))
                finished += 1
        if not srv.pending():
            time.sleep(0.0002)
    wall = (time.monotonic_ns() - t0) / 1e9
    srv.drain()
    srv.close_server()
    eng.close()
    lat = np.array(lat)
    px = sum(int(lengths[s]) * res[s][0] * res[s][1] for s in range(n))
    return {"streams": n, "frames": int(lat.size), "seconds": round(wall, 3),
            "frames_per_s": round(lat.size / wall, 1),
            "p1080_equiv_frames_per_s": round(px / (1920 * 1080) / wall, 1),
            "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
            "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
            "latency_ms_max": round(float(lat.max()), 3),
            "max_concurrent_streams": max_active, "streams_that_waited": waited,
            "admission_wait_ms_p50": round(float(np.percentile(wait_ms, 50)), 3),
            "admission_wait_ms_p99": round(float(np.percentile(wait_ms, 99)), 3),
            "batches": batches, "hbm_budget_GiB": args.budget_gib,
            "workload": f"{n} streams, Poisson {args.rate}/s, 25% each 360p/480p/720p/1080p, "
                        f"30-300 frames at 30 fps, {pool_n}-frame pool per resolution"}


def config2(P, args):
    """3840x2160 preset 2, 60 frames: plain per-stage kernels, the fused
    persistent kernel, and the fused kernel with the sequence captured as
    one HIP graph. Inputs prepared in HBM; frames/s of each."""
    W, H, F = 3840, 2160, args.c2frames
    log(f"config 2: generating {F} frames {W}x{H} preset 2")
    frames = gen_frames(P, [(2, W, H, 2222, i) for i in range(F)])
    eng = P.Engine(0)
    preps = [eng.prepare(f) for f in frames]
    del frames
    tune = None
    if not args.no_tune:
        tsid = eng.open_stream()
        eng.run([tsid], [preps[0]])
        tune = eng.autotune([tsid], [preps[1]], iters=3).as_dict()
        eng.close_stream(tsid)
    sid = eng.open_stream()
    out = {"workload": f"{W}x{H} single stream, {F}-frame preset-2 sequence "
                       "(32x32 transforms at p=0.5, MV +-256 px, level per frame)",
           "tune": tune}

    def timed(run):
        run()                      # warm: sizes scratch and frame pool
        eng.sync()
        t0 = time.perf_counter()
        run()
        eng.sync()
        return time.perf_counter() - t0

    def seq():
        for p in preps:
            eng.run([sid], [p])

    dt = timed(seq)
    out["plain_fps"] = round(F / dt, 2)
    eng.set_fused(True)
    dt = timed(seq)
    out["fused_fps"] = round(F / dt, 2)
    eng.capture_begin()
    seq()
    g = eng.capture_end()
    dt = timed(lambda: eng.graph_launch(g))
    out["fused_graph_fps"] = round(F / dt, 2)
    eng.graph_free(g)
    eng.set_fused(False)
    # per-stage view of the plain path (HIP events)
    eng.stage_stats(reset=True)
    eng.set_timing(True)
    seq()
    out["plain_stages_us"] = {n: round(1e3 * ms / c, 1)
                              for n, (ms, c) in eng.stage_stats(reset=True).items() if c}
    eng.close()
    return out


def cpu_baseline(P, frames, n):
    """The reference's own C (oracle/_ref, else the restatement) on the first
    n frames of the 1080p sequence, one thread, one call per frame on one
    decoder (its reference pool carried over): frames/s and the per-frame
    latency percentiles."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    kind = "reference" if pyoracle.have_ref() else "port"

    def timed(sample):
        lat = pyoracle.time_per_frame(sample, kind == "reference")
        dt = float(np.sum(lat))
        return dt, {"frames_per_s": round(len(sample) / dt, 3),
                    "latency_ms_p50": round(1e3 * float(np.percentile(lat, 50)), 2),
                    "latency_ms_p99": round(1e3 * float(np.percentile(lat, 99)), 2)}

    dt, one = timed(frames[:n])
    out = {"value": one["frames_per_s"], "unit": "frames/s", "cores": 1, "kind": kind,
           "latency_ms_p50": one["latency_ms_p50"], "latency_ms_p99": one["latency_ms_p99"],
           "sample": f"first {min(n, len(frames))} frames of the 1920x1080 preset-1 sequence, "
                     f"1 thread, {dt:.1f} s"}
    # the other single-stream configurations, shortened
    for key, (preset, w, h, seed, cnt) in {"config0_720p": (0, 1280, 720, 10, 30),
                                           "config2_2160p": (2, 3840, 2160, 2222, 8)}.items():
        dt, r = timed(gen_frames(P, [(preset, w, h, seed, i) for i in range(cnt)]))
        r["sample"] = f"first {cnt} frames of the {w}x{h} preset-{preset} sequence, 1 thread, {dt:.1f} s"
        out[key] = r
    return out


def cpu_baseline_mix(P, args):
    """The 64-stream mix on every host core the process may use (at most 16):
    streams dealt round robin over threads, each thread decoding its streams
    one after another with the reference's own C (one decoder per stream; the
    reference has no parallelism inside a frame). A bounded sample: the first
    args.cpu_mix_frames frames of every stream."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    per_res = args.streams // 4
    F = args.cpu_mix_frames
    layout = [(w, h) for (w, h) in MIX for _ in range(per_res)]
    specs = [(3, w, h, 7000 + sid, i) for sid, (w, h) in enumerate(layout) for i in range(F)]
    frames = gen_frames(P, specs)
    streams = [frames[k * F:(k + 1) * F] for k in range(len(layout))]
    kind = "reference" if pyoracle.have_ref() else "port"
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    run = pyoracle.run_reference if kind == "reference" else pyoracle.run_oracle

    def worker(t):
        for k in range(t, len(streams), cores):
            run(streams[k])

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=cores) as ex:
        list(ex.map(worker, range(cores)))
    dt = time.perf_counter() - t0
    px = sum(w * h for w, h in layout) * F
    return {"value": round(px / (1920 * 1080) / dt, 2), "unit": "1080p-equivalent frames/s",
            "cores": cores, "kind": kind,
            "sample": f"first {F} frames of each of the {len(layout)} streams, "
                      f"{cores} threads, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--frames", type=int, default=100)
    ap.add_argument("--streams", type=int, default=64)
    ap.add_argument("--mframes", type=int, default=60)
    ap.add_argument("--msteps", type=int, default=50)
    ap.add_argument("--mwarmup", type=int, default=5)
    ap.add_argument("--mgroups", type=int, default=2, choices=(1, 2),
                    help="batches per multi-stream step (two disjoint stream groups overlap on "
                         "the engine's two queues; more groups are not supported)")
    ap.add_argument("--cpu-frames", type=int, default=100)
    ap.add_argument("--cpu-mix-frames", type=int, default=20,
                    help="frames per stream of the all-core CPU baseline of the 64-stream mix")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the single-stream phase (for profiling the others); value is null")
    ap.add_argument("--mtune", type=int, nargs=7, default=None,
                    help="fixed multi-stream launch shape: res_split x4, intra_lanes, "
                         "intra_workers, intra_level_min (no autotuning)")
    ap.add_argument("--no-multi", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--c2frames", type=int, default=60)
    ap.add_argument("--sstreams", type=int, default=256)
    ap.add_argument("--rate", type=float, default=40.0)
    ap.add_argument("--pool", type=int, default=24)
    ap.add_argument("--budget-gib", type=float, default=3.0)
    ap.add_argument("--serve-tune", type=int, nargs=6, default=None,
                    help="res_split x4, intra_lanes, intra_workers for the serving run")
    ap.add_argument("--copy-probe-only", action="store_true",
                    help="only run the calibration copy kernel (for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(streams would shard across GPUs, one engine per GPU, with no collectives)",
              file=sys.stderr)
        sys.exit(2)
    P = load_package()
    if P.lib().cvp9_hip_device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)

    copy_bytes = 1 << 30
    copy_gbps = P.lib().cvp9_hip_copy_probe(copy_bytes, 5)
    if args.copy_probe_only:
        print(json.dumps({"copy_probe_bytes_per_launch": 2 * copy_bytes,
                          "copy_GBps": round(copy_gbps, 1)}), flush=True)
        return
    log(f"copy probe: {copy_gbps:.1f} GB/s")

    value, stats, roof, ss = None, {}, None, None
    if not args.no_single:  # (skipped only for profiling the other workloads)
        ss = single_stream(P, args)
        value = args.steps / ss["clean_s"]
        stats = ss["stats"]
        dom = max(stats, key=lambda k: stats[k]["ms_total"])
        d = stats[dom]
        kern = kernel_of(dom, ss["tune"])
        traffic, tsrc = measured_traffic(SINGLE_WORKLOAD, kern)
        roof = {"kernel": kern, "bound": "hbm", "achieved": d["algo_GBps"],
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(d["algo_GBps"] / HBM_PEAK_GBS, 4) if d["algo_GBps"] else None,
                "traffic": traffic, "traffic_source": tsrc,
                "algo_bytes_per_launch": d.get("algo_bytes_per_launch"),
                "us_per_launch": d["us_per_launch"]}
        if dom == "intra":
            # the ordered stage is bounded by its dependency chain, not by HBM:
            # levels per frame and the time per level
            order = [(args.warmup + k) % args.frames for k in range(args.steps)]
            depth = [ss["depth"][i] for i in order]
            lm = (ss["tune"] or {}).get("intra_level_min", 0)
            heads = sum(1 for i in order if lm and ss["width0"][i] >= lm)
            roof["dependency_chain"] = {
                "levels_per_frame": round(float(np.mean(depth)), 2),
                "max_levels": int(max(depth)),
                "us_per_level": round(d["us_per_launch"] / float(np.mean(depth)), 3),
                "frames_with_level_launches": heads}
        if "fp64_dense_TOPs" in d:
            roof["fp64_dense_TOPs"] = d["fp64_dense_TOPs"]
            roof["fp64_peak_no_fma_TOPs"] = FP64_PEAK_TOPS / 2
        log("stages:", json.dumps(stats))

    multi = None if args.no_multi else multi_stream(P, args)
    if multi:
        log("multi:", json.dumps(multi))
    c2 = None if args.no_c2 else config2(P, args)
    if c2:
        log("config 2:", json.dumps(c2))
    if args.serve_tune:
        a = args.serve_tune
        args.serve_tune = (a[:4], a[4], a[5])
    serve = None if args.no_serve else serving(P, args)
    if serve:
        log("serving:", json.dumps(serve))
    cpu = None if args.no_cpu or ss is None else cpu_baseline(P, ss["frames"], args.cpu_frames)
    cpu_mix = None if args.no_cpu or not multi else cpu_baseline_mix(P, args)

    line = {
        "metric": METRIC,
        "value": round(value, 2) if value is not None else None,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * ss["clean_s"] / args.steps, 4) if ss else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/i16 (fp64 transform)",
        "data": "synthetic parsed frames (seeded generator), inputs resident in HBM",
        "config": {"workload": SINGLE_WORKLOAD, "frames": args.frames,
                   "step": "one frame"},
        "dropin_host_pcie_inclusive_fps": round(ss["dropin_fps"], 1) if ss else None,
        "timed_with_events_fps": round(args.steps / ss["timed_s"], 2) if ss else None,
        "algo_GBps": round(ss["algo_bytes"] / ss["clean_s"] / 1e9, 1) if ss else None,
        "stages": stats,
        "tune": ss["tune"] if ss else None,
        "roofline": roof,
        "copy_peak_GBps": round(copy_gbps, 1),
        "multi_stream": None if not multi else {
            "workload": f"{multi['streams']} streams (16 each 640x360/854x480/1280x720/1920x1080), "
                        "each step one frame per stream, as one batch or as "
                        f"{multi['groups']} batches of interleaved streams (the faster)",
            "schedule": "one batch per step" if multi["schedule_groups"] == 1 else
                        f"{multi['schedule_groups']} interleaved batches per step",
            "frames_per_s": round(multi["frames_per_s"], 1),
            "p1080_equiv_frames_per_s": round(multi["p1080_equiv_per_s"], 1),
            "ms_per_step": round(multi["ms_per_batch"], 4),
            "one_batch_per_step_p1080_equiv_frames_per_s": round(multi["one_batch_p1080_equiv_per_s"], 1),
            f"{multi['groups']}_batches_per_step_p1080_equiv_frames_per_s":
                round(multi["groups_p1080_equiv_per_s"], 1),
            "stages_one_batch": multi["stages"],
            "tune": multi["tune"],
            "vs_single_stream": round(multi["p1080_equiv_per_s"] / value, 2) if value else None,
            "cpu_baseline_all_cores": cpu_mix},
        "config2_2160p": c2,
        "serving": serve,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
