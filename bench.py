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
           output bit for bit.  Also frame-parallel over the host's cores
           (cpu_baseline.parallel), frames being independent

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
                    help="only the batched graph renders (for profilers): no plain, stage, CPU or other legs")
    ap.add_argument("--no-extra", action="store_true", help="skip the 4K, uneven-batch and shading legs")
    args = ap.parse_args()
    if args.graph_only:
        args.plain_steps = args.profile_runs = args.cpu_frames = 0
        args.no_extra = True

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(json.dumps({"error": "one GPU only: the frames of a batch are independent and would shard "
                          "across GPUs without any exchange, but multi-GPU runs are not part of this library"}))
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
        # HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE
        # summary of this kernel
        tf = os.path.join(ROOT, "profiles", f"{dominant}_sweep_1080p.md")
        if os.path.exists(tf):
            with open(tf) as fh:
                for line in fh:
                    if line.startswith("hbm_bytes_per_launch:"):
                        traffic = int(line.split(":")[1])
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
        # the same C path frame-parallel over the host's cores (frames are
        # independent): the fairer CPU comparison, reported beside the
        # single-thread baseline
        ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        ncpu = max(1, min(ncpu, 16))
        idx2 = np.linspace(0, F - 1, 2 * ncpu).round().astype(int)
        log(f"bench: CPU baseline frame-parallel on {ncpu} threads, {len(idx2)} frames")
        t0 = time.perf_counter()
        oracle.render_frames(W, H, U, Rs[idx2], m.xyz, m.rgb, image=True, nthreads=ncpu, keep=False)
        par_s = time.perf_counter() - t0
        cpu["parallel"] = {"value": round(len(idx2) / par_s, 4), "cores": ncpu,
                           "sample": f"{len(idx2)} sweep frames, one per thread at a time, {par_s:.1f} s"}

    extra = {}
    if not args.no_extra:
        extra["frame_4k"] = leg_4k(sloth_amd, meshgen)
        extra["uneven_batch"] = leg_uneven(sloth_amd, meshgen)
        extra["fwd_bwd"] = leg_shading(sloth_amd, meshgen)

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
        "fwd_bwd_views_per_s": extra.get("fwd_bwd", {}).get("views_per_s"),
        "roofline": roof,
        "cpu_baseline": cpu,
        "legs": extra,
    }
    print(json.dumps(out), flush=True)
    return 0


def leg_4k(sloth_amd, meshgen):
    """One 3840x2160 frame of the 8-subdivision displaced icosphere
    (1,310,720 triangles, seed 1), plain form (launch sequence + sync)."""
    log("bench: 4K frame leg")
    m, W, H, U, R = meshgen.workload_icosphere_4k()
    mesh = sloth_amd.SimpleMesh.from_mesh(m)
    b = sloth_amd.Batch([sloth_amd.FrameSpec(W, H, U, True)], [sloth_amd.DrawSpec(0, mesh, R)])
    ms = b.time(10, 2, per_frame=True) / 10
    return {"workload": f"displaced icosphere s8 ({m.n} tris, seed 1), 3840x2160 image mode, one frame",
            "ms_per_frame": round(ms, 4), "frames_per_s": round(1e3 / ms, 1),
            "gtri_per_s": round(m.n / ms / 1e6, 3)}


def leg_uneven(sloth_amd, meshgen):
    """64 meshes, triangle counts log-uniform in 1k..1M (seed 3), each with its
    own resolution and rotation, rendered by one batch launch sequence."""
    log("bench: uneven-batch leg (generating 64 meshes)")
    items = meshgen.workload_uneven_meshes()
    frames, draws, tris = [], [], 0
    keep = []
    for i, (m, W, H, U, R) in enumerate(items):
        sm = sloth_amd.SimpleMesh.from_mesh(m)
        keep.append(sm)
        frames.append(sloth_amd.FrameSpec(W, H, U, True))
        draws.append(sloth_amd.DrawSpec(i, sm, R))
        tris += m.n
    b = sloth_amd.Batch(frames, draws)
    ms = b.time(5, 2) / 5
    pms = b.time(1, 1, per_frame=True)
    return {"workload": f"64 meshes, {tris} triangles total (1k..1M log-uniform, seed 3), resolutions from "
                        f"{sorted(set((f.width, f.height) for f in frames))}",
            "ms_per_batch": round(ms, 4), "meshes_per_s": round(64e3 / ms, 1),
            "gtri_per_s": round(tris / ms / 1e6, 3), "plain_ms": round(pms, 4),
            "batched_over_plain": round(pms / ms, 2)}


def leg_shading(sloth_amd, meshgen):
    """Differentiable shading: UV sphere 224 x 224 (99,904 triangles), 32
    seeded views at 1024x1024, L2 loss against a target rendered with colours
    + N(0, 0.1^2) and the light rotated 10 degrees; one step = forward +
    backward over all 32 views with visibility fixed."""
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
    ms = sh.time(10, 2) / 10
    return {"workload": f"UV sphere 224x224 ({m.n} tris), {V} views 1024x1024, fwd+bwd, visibility fixed",
            "ms_per_step": round(ms, 4), "views_per_s": round(V * 1e3 / ms, 1), "loss": loss}


if __name__ == "__main__":
    sys.exit(main())
