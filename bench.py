#!/usr/bin/env python3
"""Benchmark: the 256-frame 1080p turntable sweep of a displaced icosphere.

Workload (a "step" = one pass of the raster path over the whole sweep):
  displaced icosphere, 7 subdivisions = 327,680 triangles (seed 2), 256 frames
  of the `-j 256` turntable schedule (pitch pi + k * fl(2 pi / 256)), each
  1920x1080 image mode (2,074,680 cells + 2,073,600 z slots), default shader.
  Inputs are resident in device memory before the timed region; outputs stay
  there.

Forms timed:
  batched  all 256 frames by one launch sequence captured in a HIP graph
           (sloth_batch_run)                                  -> `value`
  plain    one launch sequence per frame, device sync after each
  cpu      the oracle (C restatement of the reference, 1 thread) on a bounded
           sample of frames; the same frames are compared against the GPU
           output bit for bit

Prints ONE JSON line.  Torch is not imported: it bundles a second HIP runtime.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "rust-sloth_amd"))

import numpy as np  # noqa: E402

METRIC = "frames/s + Gtri/s at 1080p (batched vs plain GPU vs CPU); fwd+bwd views/s"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=256)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--profile-runs", type=int, default=5)
    ap.add_argument("--cpu-frames", type=int, default=8, help="frames in the CPU-baseline sample (0 = skip)")
    ap.add_argument("--graph-only", action="store_true",
                    help="only the batched graph renders (for profilers): no plain, stage or CPU legs")
    args = ap.parse_args()
    if args.graph_only:
        args.plain_steps = args.profile_runs = args.cpu_frames = 0

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(json.dumps({"error": "this build runs on one GPU; frames shard across GPUs without "
                          "exchange (see DESIGN.md), but multi-GPU runs are not part of it"}))
        return 2

    import sloth_amd
    from sloth_amd import meshgen

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

    log("bench: batched (graph) timing")
    ms = batch.time(args.steps, args.warmup, per_frame=False)
    ms_step = ms / args.steps
    fps = F / (ms_step / 1e3)

    log("bench: per-stage profile")
    prof = {}
    for _ in range(args.profile_runs):
        for k, v in batch.profile().items():
            prof[k] = prof.get(k, 0.0) + v / args.profile_runs
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
        traffic = None
        tf = os.path.join(ROOT, "profiles", f"traffic_{dominant}_sweep_1080p.json")
        if os.path.exists(tf):
            with open(tf) as fh:
                traffic = json.load(fh).get("hbm_bytes_per_launch")
        roof = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "alg_bytes_per_launch": alg_bytes, "kernel_ms": round(prof[dominant], 4),
                "stages_ms": {k: round(v, 4) for k, v in prof.items()}}

    cpu = None
    verified = 0
    frags = None
    if args.cpu_frames > 0:
        import oracle

        idx = np.linspace(0, F - 1, min(args.cpu_frames, F)).round().astype(int)
        log(f"bench: CPU baseline (oracle, 1 thread) on frames {idx.tolist()}")
        t0 = time.perf_counter()
        cells, z = oracle.render_frames(W, H, U, Rs[idx], m.xyz, m.rgb, image=True, nthreads=1)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(len(idx) / cpu_s, 4), "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": f"{len(idx)} of the {F} sweep frames ({idx.tolist()}), update+clear+draw_mesh, "
                         f"C restatement of the Rust reference (-O2 -ffp-contract=off), {cpu_s:.1f} s"}
        for j, f in enumerate(idx):
            gc, gz, _ = batch.read_frame(int(f))
            if np.array_equal(gc, cells[j]) and np.array_equal(gz.view(np.uint32), z[j].view(np.uint32)):
                verified += 1
        fr = oracle.Frame(W, H, True)
        fr.draw_mesh(U, m.xyz, m.rgb, Rs[0])
        frags = int(fr.stats[1])

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
        "gtri_per_s": round(N * fps / 1e9, 3),
        "mpix_per_s": round(ncell * fps / 1e6, 1),
        "plain_frames_per_s": round(plain_fps, 2) if plain_fps else None,
        "batched_over_plain": round(fps / plain_fps, 2) if plain_fps else None,
        "batched_over_cpu": round(fps / cpu["value"], 1) if cpu else None,
        "fragments_per_frame": frags,
        "atomics_per_s": round(frags * fps, 1) if frags else None,
        "verified_frames": verified,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
