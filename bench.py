#!/usr/bin/env python3
"""Benchmark: the 256-frame 1080p turntable sweep of a displaced icosphere.

Workload (a "step" = one pass of the raster path over the whole sweep):
  displaced icosphere, 7 subdivisions = 327,680 triangles (seed 2), 256 frames
  of the `-j 256` turntable schedule (pitch pi + k * fl(2 pi / 256)), each
  1920x1080 image mode (2,074,680 cells + 2,073,600 z slots), default shader.
  Inputs are resident in device memory before the timed region; outputs stay
  there.  The batch's tile size and raster variant come from the autotuner,
  run before the warmup.

Forms timed:
  batched  all 256 frames by one launch sequence captured in a HIP graph
           (sloth_batch_run)                                  -> `value`
  plain    one launch sequence per frame, device sync after each
  cpu      the oracle (C restatement of the reference) on a bounded sample of
           frames, one thread and frame-parallel over the host's cores; median
           of 5 runs after a warm-up; the sampled frames are compared with the
           GPU output bit for bit

Roofline: HBM bytes (algorithmic bytes of the dominant kernel / its time,
HIP events) and atomics (the raster's LDS atomicMin rate against the
measured peak of rust-sloth_amd/bench/atomic_peak.hip).

Prints ONE JSON line.  Torch is not imported: it bundles a second HIP runtime.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "rust-sloth_amd"))

import numpy as np  # noqa: E402

METRIC = "frames/s + Gtri/s at 1080p (batched vs plain GPU vs CPU); fwd+bwd views/s"
HBM_PEAK_GBS = 8000.0
ATOMIC_PEAK = os.path.join(ROOT, "rust-sloth_amd", "sloth_amd", "atomic_peak")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(n, 16))


def profile_value(name: str, key: str):
    """A `key: value` line of a committed profile summary under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1].split()[0])
    return None


def stage_times(batch, runs: int) -> dict:
    """Median per-stage milliseconds over `runs` profiled renders (HIP events
    between the stages on the batch's stream)."""
    acc = {}
    for _ in range(runs):
        for k, v in batch.profile().items():
            acc.setdefault(k, []).append(v)
    return {k: statistics.median(v) for k, v in acc.items()}


def atomic_peak():
    if not os.path.exists(ATOMIC_PEAK):
        return None
    r = subprocess.run([ATOMIC_PEAK], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        log(f"bench: atomic_peak failed: {r.stderr.strip()}")
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=256)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--profile-runs", type=int, default=5)
    ap.add_argument("--repeats", type=int, default=5, help="timed regions for the medians")
    ap.add_argument("--cpu-frames", type=int, default=2, help="frames per CPU-baseline run (0 = skip)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile size (64x32 cells)")
    ap.add_argument("--tile", default=None, help="WxH: the sweep's tile size in cells, instead of the autotuner's")
    ap.add_argument("--walk", type=int, default=None, help="0/1: the tile raster variant without/with the "
                    "wavefront walk of big boxes (with --tile)")
    ap.add_argument("--shading-only", action="store_true", help="only the shading leg (for profilers)")
    ap.add_argument("--frame4k-only", default=None, metavar="WxH,WALK,TPT",
                    help="only the 4K-frame leg, at this plan instead of the autotuner's (for profilers)")
    ap.add_argument("--uneven-only", default=None, metavar="WxH,WALK,TPT",
                    help="only the uneven-batch leg, at this plan instead of the autotuner's (for profilers)")
    ap.add_argument("--graph-only", action="store_true",
                    help="only the batched graph renders (for profilers): no plain, stage, CPU or other legs")
    ap.add_argument("--no-extra", action="store_true", help="skip the 4K, uneven-batch and shading legs")
    args = ap.parse_args()
    if args.graph_only:
        args.plain_steps = args.profile_runs = args.cpu_frames = 0
        args.repeats = 1
        args.no_extra = True

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(json.dumps({"error": "one GPU only: the frames of a batch are independent and would shard "
                          "across GPUs without any exchange, but multi-GPU runs are not part of this library"}))
        return 2

    import sloth_amd
    from sloth_amd import meshgen

    if args.shading_only:
        sloth_amd.init(0)
        print(json.dumps({"fwd_bwd": leg_shading(sloth_amd, meshgen, args)}), flush=True)
        return 0
    if args.frame4k_only:
        sloth_amd.init(0)
        print(json.dumps({"frame_4k": leg_4k(sloth_amd, meshgen, args)}), flush=True)
        return 0
    if args.uneven_only:
        sloth_amd.init(0)
        print(json.dumps({"uneven_batch": leg_uneven(sloth_amd, meshgen, args)}), flush=True)
        return 0

    F = args.frames
    log(f"bench: building sweep workload ({F} frames)")
    m, W, H, U, Rs = meshgen.workload_sweep_1080p(F)
    N = m.n
    ncell = W * H + H
    nz = W * H

    sloth_amd.init(0)
    mesh = sloth_amd.SimpleMesh.from_mesh(m)
    frames = [sloth_amd.FrameSpec(W, H, U, True) for _ in range(F)]
    draws = [sloth_amd.DrawSpec(f, mesh, Rs[f]) for f in range(F)]
    batch = sloth_amd.Batch(frames, draws)
    tune_s = None
    if args.tile:
        tw, th = (int(v) for v in args.tile.lower().split("x"))
        batch.set_tile(tw, th)
        if args.walk is not None:
            batch.set_walk(args.walk)
    elif not args.no_tune:
        log("bench: autotuning the sweep's tile size")
        t0 = time.perf_counter()
        batch.autotune()
        tune_s = time.perf_counter() - t0
    plan = batch.plan()

    log("bench: batched (graph) timing")
    ms = batch.time(args.steps, args.warmup, per_frame=False)
    ms_step = ms / args.steps
    fps = F / (ms_step / 1e3)
    reps = [ms_step] + [batch.time(args.steps, 1, per_frame=False) / args.steps for _ in range(args.repeats - 1)]
    ms_median = statistics.median(reps)

    prof = stage_times(batch, args.profile_runs) if args.profile_runs > 0 else {}
    dominant = max(prof, key=prof.get) if prof else None

    plain_fps = None
    if args.plain_steps > 0:
        log("bench: plain (per-frame) timing")
        pms = batch.time(args.plain_steps, 1, per_frame=True)
        plain_fps = F / (pms / args.plain_steps / 1e3)

    # leave the batch holding one graph render for the parity spot check
    batch.run()
    batch.sync()

    # algorithmic bytes per sweep: per frame glyph u8 + RGB u8x3 per cell and
    # the f32 depth buffer; per launch the mesh once (xyz f32 x 3 vertices +
    # RGB u8 x 3 per triangle)
    alg_bytes = F * (4 * ncell + 4 * nz) + 39 * N
    roof = None
    if dominant:
        t = prof[dominant] / 1e3
        achieved = alg_bytes / t / 1e9
        prof_name = f"{dominant}_sweep_1080p.md"
        rp_us = profile_value(prof_name, "kernel_avg_us")
        roof = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": profile_value(prof_name, "hbm_bytes_per_launch"),
                "alg_bytes_per_launch": alg_bytes, "kernel_ms": round(prof[dominant], 4),
                "kernel_ms_source": f"HIP events around the stage, median of {args.profile_runs} profiled renders",
                "rocprof_avg_ms": round(rp_us / 1e3, 4) if rp_us else None,
                "rocprof_frac": round(alg_bytes / (rp_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4) if rp_us else None,
                "stages_ms": {k: round(v, 4) for k, v in prof.items()}}

    cpu = None
    verified = 0
    frags = None
    if args.cpu_frames > 0:
        import oracle

        runs = args.repeats
        idx_all = np.linspace(0, F - 1, runs * args.cpu_frames).round().astype(int)
        log(f"bench: CPU baseline (oracle, 1 thread), {runs} runs of {args.cpu_frames} frames")
        oracle.render_frames(W, H, U, Rs[idx_all[:1]], m.xyz, m.rgb, image=True, nthreads=1, keep=False)  # warm-up
        rates = []
        for r in range(runs):
            idx = idx_all[r::runs]
            t0 = time.perf_counter()
            cells, z = oracle.render_frames(W, H, U, Rs[idx], m.xyz, m.rgb, image=True, nthreads=1)
            rates.append(len(idx) / (time.perf_counter() - t0))
            for j, f in enumerate(idx):
                gc, gz, _ = batch.read_frame(int(f))
                if np.array_equal(gc, cells[j]) and np.array_equal(gz.view(np.uint32), z[j].view(np.uint32)):
                    verified += 1
        cpu = {"value": round(statistics.median(rates), 4), "unit": "frames/s", "cores": 1, "kind": "port",
               "cpu": cpu_model(),
               "sample": f"median of {runs} runs of {args.cpu_frames} sweep frames each ({idx_all.tolist()}), "
                         f"after one warm-up frame; update+clear+draw_mesh; C restatement of the Rust reference "
                         f"(-O2 -ffp-contract=off), no Rust toolchain available",
               "runs": [round(x, 4) for x in rates]}
        # the same C path frame-parallel over the host's cores (frames are
        # independent): the fairer CPU comparison
        ncpu = host_cores()
        prates = []
        for r in range(runs):
            idx2 = (np.arange(ncpu) * (F // ncpu) + r) % F
            t0 = time.perf_counter()
            oracle.render_frames(W, H, U, Rs[idx2], m.xyz, m.rgb, image=True, nthreads=ncpu, keep=False)
            prates.append(len(idx2) / (time.perf_counter() - t0))
        cpu["parallel"] = {"value": round(statistics.median(prates), 4), "cores": ncpu,
                           "sample": f"median of {runs} runs of {ncpu} sweep frames, one per thread"}
        # LDS atomics of the raster = fragments passing the edge test, from
        # the oracle's count on 4 frames of the sweep
        counts = []
        for f in np.linspace(0, F - 1, 4).round().astype(int):
            fr = oracle.Frame(W, H, True)
            fr.draw_mesh(U, m.xyz, m.rgb, Rs[f])
            counts.append(int(fr.stats[1]))
        frags = statistics.mean(counts)

    atom = atomic_peak() if not args.graph_only else None
    atomic_roof = None
    if atom and frags and dominant:
        per_sweep = frags * F
        rate = per_sweep / (prof[dominant] / 1e3)
        atomic_roof = {"kernel": dominant, "achieved": round(rate, 1), "peak": atom["lds_min_u64_per_s"],
                       "unit": "LDS atomicMin u64/s", "frac": round(rate / atom["lds_min_u64_per_s"], 4),
                       "atomics_per_launch": round(per_sweep),
                       "atomics_source": "derived: the oracle's count of fragments passing the edge test (one LDS "
                                         "atomicMin each) on 4 sweep frames x frames per sweep",
                       "peak_source": "measured: rust-sloth_amd/bench/atomic_peak.hip", "microbench": atom}

    extra = {}
    if not args.no_extra:
        extra["frame_4k"] = leg_4k(sloth_amd, meshgen, args)
        extra["uneven_batch"] = leg_uneven(sloth_amd, meshgen, args)
        extra["fwd_bwd"] = leg_shading(sloth_amd, meshgen, args)

    out = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"sweep_1080p: displaced icosphere s7 ({N} tris, seed 2), {F} frames "
                               f"1920x1080 image mode, turntable schedule, one HIP graph per sweep",
                   "frames": F, "triangles": N, "width": W, "height": H},
        "plan": dict(plan, autotune_s=round(tune_s, 2) if tune_s else None),
        "ms_per_step_median": round(ms_median, 4),
        "ms_per_step_runs": [round(x, 4) for x in reps],
        "gtri_per_s": round(N * fps / 1e9, 3),
        "mpix_per_s": round(ncell * fps / 1e6, 1),
        "plain_frames_per_s": round(plain_fps, 2) if plain_fps else None,
        "batched_over_plain": round(fps / plain_fps, 2) if plain_fps else None,
        "batched_over_cpu": round(fps / cpu["value"], 1) if cpu else None,
        "batched_over_cpu_parallel": round(fps / cpu["parallel"]["value"], 1) if cpu else None,
        "fragments_per_frame": round(frags) if frags else None,
        "atomics_per_s": atomic_roof["achieved"] if atomic_roof else None,
        "verified_frames": verified,
        "fwd_bwd_views_per_s": extra.get("fwd_bwd", {}).get("views_per_s"),
        "roofline": roof,
        "atomic_roofline": atomic_roof,
        "cpu_baseline": cpu,
        "legs": extra,
    }
    print(json.dumps(out), flush=True)
    return 0


def _frac(alg_bytes, ms):
    gbs = alg_bytes / (ms / 1e3) / 1e9
    return round(gbs, 1), round(gbs / HBM_PEAK_GBS, 4)


def leg_4k(sloth_amd, meshgen, args):
    """One 3840x2160 frame of the 8-subdivision displaced icosphere
    (1,310,720 triangles, seed 1): plain form (launch sequence + sync) and
    graph replay, median of 5 timings of 10 frames each."""
    log("bench: 4K frame leg")
    m, W, H, U, R = meshgen.workload_icosphere_4k()
    mesh = sloth_amd.SimpleMesh.from_mesh(m)
    b = sloth_amd.Batch([sloth_amd.FrameSpec(W, H, U, True)], [sloth_amd.DrawSpec(0, mesh, R)])
    if args.frame4k_only:
        tile, walk, tpt = args.frame4k_only.lower().split(",")
        b.set_tile(*(int(v) for v in tile.split("x")))
        b.set_walk(int(walk))
        b.set_bin(int(tpt))
    elif not args.no_tune:
        b.autotune()
    ms = statistics.median(b.time(10, 2, per_frame=True) / 10 for _ in range(5))
    msg = statistics.median(b.time(10, 2) / 10 for _ in range(5))
    alg = 4 * (W * H + H) + 4 * W * H + 39 * m.n
    gbs, frac = _frac(alg, msg)
    return {"workload": f"displaced icosphere s8 ({m.n} tris, seed 1), 3840x2160 image mode, one frame",
            "plan": b.plan(), "ms_per_frame": round(ms, 4), "frames_per_s": round(1e3 / ms, 1),
            "ms_per_frame_graph": round(msg, 4), "gtri_per_s": round(m.n / ms / 1e6, 3),
            "alg_bytes": alg, "achieved_GBps_graph": gbs, "frac_graph": frac}


def leg_uneven(sloth_amd, meshgen, args):
    """64 meshes, triangle counts log-uniform in 1k..1M (seed 3), each with its
    own resolution and rotation, rendered by one batch launch sequence."""
    log("bench: uneven-batch leg (generating 64 meshes)")
    items = meshgen.workload_uneven_meshes()
    frames, draws, tris = [], [], 0
    keep = []
    alg = 0
    for i, (m, W, H, U, R) in enumerate(items):
        sm = sloth_amd.SimpleMesh.from_mesh(m)
        keep.append(sm)
        frames.append(sloth_amd.FrameSpec(W, H, U, True))
        draws.append(sloth_amd.DrawSpec(i, sm, R))
        tris += m.n
        alg += 4 * (W * H + H) + 4 * W * H + 39 * m.n
    b = sloth_amd.Batch(frames, draws)
    if args.uneven_only:
        tile, walk, tpt = args.uneven_only.lower().split(",")
        b.set_tile(*(int(v) for v in tile.split("x")))
        b.set_walk(int(walk))
        b.set_bin(int(tpt))
    elif not args.no_tune:
        b.autotune()
    ms = statistics.median(b.time(5, 2) / 5 for _ in range(5))
    # (profiling runs keep to the batched launches)
    pms = None if args.uneven_only else b.time(1, 1, per_frame=True)
    gbs, frac = _frac(alg, ms)
    return {"workload": f"64 meshes, {tris} triangles total (1k..1M log-uniform, seed 3), resolutions from "
                        f"{sorted(set((f.width, f.height) for f in frames))}",
            "plan": b.plan(), "ms_per_batch": round(ms, 4), "meshes_per_s": round(64e3 / ms, 1),
            "gtri_per_s": round(tris / ms / 1e6, 3), "plain_ms": round(pms, 4) if pms else None,
            "batched_over_plain": round(pms / ms, 2) if pms else None, "alg_bytes": alg, "achieved_GBps": gbs,
            "frac": frac}


def leg_shading(sloth_amd, meshgen, args):
    """Differentiable shading: UV sphere 224 x 224 (99,904 triangles), 32
    seeded views at 1024x1024, L2 loss against a target rendered with colours
    + N(0, 0.1^2) and the light rotated 10 degrees; one step = forward +
    backward over all 32 views with visibility fixed (k_shade_bwd computes
    both).  Algorithmic bytes per step: per z slot the visibility word (4 B)
    and the target RGB (12 B); per triangle its vertex ids (12 B); per vertex
    the colour and normal read and their gradients written (4 x 12 B)."""
    log("bench: shading fwd+bwd leg")
    wl = meshgen.workload_shading_views()
    m, W, H, U, Rs = wl["mesh"], wl["W"], wl["H"], wl["U"], wl["Rs"]
    V = len(Rs)
    mesh = sloth_amd.SimpleMesh.from_mesh(m)
    b = sloth_amd.Batch([sloth_amd.FrameSpec(W, H, U, True) for _ in range(V)],
                        [sloth_amd.DrawSpec(f, mesh, Rs[f]) for f in range(V)])
    sh = sloth_amd.Shading(b, m.faces, m.verts.shape[0])
    vc, vn, L = wl["vcolor"], wl["vnormal"], wl["light"]
    _, target = sh.forward(wl["vcolor_target"], vn, wl["light_target"], want_out=True)
    sh.set_target(target)
    loss, _ = sh.forward(vc, vn, L)
    ms = statistics.median(sh.time(10, 2) / 10 for _ in range(5))
    alg = V * W * H * 16 + m.n * 12 + m.verts.shape[0] * 48
    gbs, frac = _frac(alg, ms)
    return {"workload": f"UV sphere 224x224 ({m.n} tris), {V} views 1024x1024, fwd+bwd, visibility fixed",
            "ms_per_step": round(ms, 4), "views_per_s": round(V * 1e3 / ms, 1), "loss": loss,
            "alg_bytes": alg, "achieved_GBps": gbs, "frac": frac,
            "rocprof_avg_ms": (lambda v: round(v / 1e3, 4) if v else None)(
                profile_value("k_shade_bwd_shading_32x1024.md", "kernel_avg_us"))}


if __name__ == "__main__":
    sys.exit(main())
