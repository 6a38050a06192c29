"""Seeded random batches against the oracle (GPU).

Each case draws a batch of 1-6 frames of random size (1-320 x 1-100 cells,
either parity, image or terminal mode) with 0-3 draws each (random soups from
sub-pixel to frame-filling triangles, slivers, triangles far off screen,
closed meshes, empty meshes), a batch-wide shader, and a random plan: tile
size, big-box walk and the binning kernel's triangles per thread; half the
closed meshes are indexed device meshes.  Every
frame's cells and depths must equal the oracle's (rasterizer.rs:39-93), and
the winners of the last frame too; the graph is replayed once and checked
again, with each frame's encoded byte streams (plain, colour, HTML); then
every draw takes a new transform (some zoomed, so that tile lists can
outgrow their sizing and the batch re-plans) and every frame is checked once
more.  The cases mix the plan's paths: flat and 3-D raster grids, one and
several draws per frame, both raster parities; each case also runs on the
general global-key path."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from _helpers import assert_frame_equal, oracle_render, random_soup, say

pytestmark = pytest.mark.gpu

TILES = [(8, 4), (16, 8), (32, 8), (64, 16), (64, 32), (128, 16), (256, 8)]


def _mesh(rng, kind, W, H, U):
    from sloth_amd import meshgen

    if kind == "soup":
        return random_soup(rng, int(rng.integers(1, 2000)), W, H, U, size=float(rng.choice([0.01, 0.05, 0.2])))
    if kind == "big":
        return random_soup(rng, int(rng.integers(1, 40)), W, H, U, size=1.0)
    if kind == "far":
        xyz, rgb = random_soup(rng, int(rng.integers(1, 300)), W, H, U, spread=40.0, size=2.0)
        return xyz, rgb
    if kind == "sliver":
        n = int(rng.integers(1, 400))
        a = rng.uniform(-1, 1, size=(n, 3))
        d = rng.normal(0.0, 0.4, size=(n, 3))
        perp = rng.normal(0.0, 1.0, size=(n, 3)) * rng.choice([1e-3, 1e-5, 0.0], size=(n, 1))
        xyz = np.stack([a, a + d, a + 0.5 * d + perp], 1).astype(np.float32)
        return xyz, rng.integers(0, 256, size=(n, 3), dtype=np.uint8)
    if kind == "empty":
        return np.zeros((0, 3, 3), np.float32), np.zeros((0, 3), np.uint8)
    m = meshgen.displaced_icosphere(int(rng.integers(1, 5)), seed=int(rng.integers(0, 1000)))
    return m.xyz, m.rgb, m


@pytest.mark.parametrize("path", ["tiled", "global"])
@pytest.mark.parametrize("seed", range(64))
def test_random_batch(sloth, seed, path, monkeypatch):
    from sloth_amd import Shader, meshgen

    # the general global-key path (SLOTH_FORCE_GLOBAL=1) has no tile plan
    if path == "global":
        monkeypatch.setenv("SLOTH_FORCE_GLOBAL", "1")
    else:
        monkeypatch.delenv("SLOTH_FORCE_GLOBAL", raising=False)

    rng = np.random.default_rng(1000 + seed)
    shaders = [None, Shader((-0.1, 0.25, 0.75, 1.0, 2.0), "01234", "#"),
               Shader((0.9, 0.3, 0.6, 0.3, 0.95), "vwxyz", "q"),
               Shader(tuple(-0.05 + 0.0625 * k for k in range(16)), "ABCDEFGHIJKLMNOP", "!")]
    sh = shaders[int(rng.integers(0, len(shaders)))]
    n_frames = int(rng.integers(1, 7))
    frames, draws, plan = [], [], []
    for f in range(n_frames):
        W, H = int(rng.integers(1, 321)), int(rng.integers(1, 101))
        image = bool(rng.integers(0, 2))
        U = meshgen.screen_transform(W, H, [np.array([1.0, 1.0, 1.0], np.float32)])
        frames.append(sloth.FrameSpec(W, H, U, image))
        fd = []
        for _ in range(int(rng.integers(0, 4))):
            kind = str(rng.choice(["soup", "soup", "big", "far", "sliver", "closed", "empty"]))
            xyz, rgb, *closed = _mesh(rng, kind, W, H, U)
            R = meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3))
            # closed meshes: a soup or (half of them) an indexed device mesh
            sm = sloth.SimpleMesh.from_mesh(closed[0], indexed=True) if closed and rng.integers(0, 2) else \
                sloth.SimpleMesh(xyz, rgb)
            draws.append(sloth.DrawSpec(f, sm, R))
            fd.append((xyz, rgb, R, sh))
        plan.append((W, H, U, image, fd))
    tile = TILES[int(rng.integers(0, len(TILES)))]
    walk = int(rng.integers(0, 2))
    tpt = int(rng.choice([1, 2, 4]))
    say(f"case random batch seed {seed} [{path}]: {n_frames} frames {[(p[0], p[1], p[3], len(p[4])) for p in plan]}, "
        f"tile {tile}, walk {walk}, {tpt} per thread, shader {sh is not None}")
    b = sloth.Batch(frames, draws, sh) if sh is not None else sloth.Batch(frames, draws)
    if path == "tiled":
        b.set_tile(*tile)
        b.set_walk(walk)
        b.set_bin(tpt)
    for rep in range(2):
        b.run()
        b.sync()
        for f, (W, H, U, image, fd) in enumerate(plan):
            ref = oracle_render(W, H, U, fd, image=image)
            want = rep == 1 and f == n_frames - 1
            cells, z, win = b.read_frame(f, want_winners=want)
            assert_frame_equal(cells, z, win if want else None, ref, W)
            if rep == 1:
                # Context::flush's byte streams (context.rs:50-92) of the frame
                for mode in (sloth.ENCODE_TEXT, sloth.ENCODE_ANSI, sloth.ENCODE_HTML):
                    assert b.encode_frame(f, mode) == oracle.encode(ref.cells, mode, image), (f, mode)
    # new transforms for every draw (some zoomed, so that tile lists can
    # outgrow their sizing and the batch re-plans), then render and check again
    if draws:
        news = []
        for fd in (p[4] for p in plan):
            for j, (xyz, rgb, R, dsh) in enumerate(fd):
                R2 = np.asarray(meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3)), np.float32).reshape(4, 4).copy()
                if rng.integers(0, 2):
                    R2[:3, :3] *= np.float32(rng.uniform(1.0, 3.0))
                fd[j] = (xyz, rgb, R2.reshape(-1), dsh)
        # the order of the draw array: draws were appended frame by frame
        for fd in (p[4] for p in plan):
            news.extend(d[2] for d in fd)
        b.set_transforms(np.stack(news))
        b.run()
        b.sync()
        for f, (W, H, U, image, fd) in enumerate(plan):
            cells, z, _ = b.read_frame(f)
            assert_frame_equal(cells, z, None, oracle_render(W, H, U, fd, image=image), W)


def _random_shader(rng):
    from sloth_amd import Shader

    kind = int(rng.integers(0, 5))
    n = int(rng.integers(0, 17))
    thr = np.sort(rng.uniform(-0.2, 1.2, n)) if kind < 3 else rng.uniform(-0.2, 1.2, n)
    thr = [float(t) for t in thr]
    if kind == 1 and n > 1:
        thr[-1] = float("inf")
    if kind == 2 and n > 2:
        thr[1] = thr[0]
    glyphs = "".join(chr(int(c)) for c in rng.integers(33, 127, n))
    return Shader(tuple(thr), glyphs, chr(int(rng.integers(33, 127))))


@pytest.mark.parametrize("path", ["tiled", "global"])
@pytest.mark.parametrize("seed", range(32))
def test_random_context_shaders(sloth, seed, path, monkeypatch):
    """One frame through the Context API (draw_mesh per draw,
    rasterizer.rs:39-46) with a random shader per draw: sorted and unsorted
    step tables, repeated and infinite thresholds, from no step to all 16;
    cells, depths and winners against the oracle."""
    from sloth_amd import meshgen

    if path == "global":
        monkeypatch.setenv("SLOTH_FORCE_GLOBAL", "1")
    else:
        monkeypatch.delenv("SLOTH_FORCE_GLOBAL", raising=False)
    rng = np.random.default_rng(3000 + seed)
    W, H = int(rng.integers(1, 200)), int(rng.integers(1, 70))
    image = bool(rng.integers(0, 2))
    U = meshgen.screen_transform(W, H, [np.array([1.0, 1.0, 1.0], np.float32)])
    ctx = sloth.Context(W, H, image=image)
    ctx.utransform = U
    ctx.clear()
    draws, keep = [], []
    for _ in range(int(rng.integers(1, 6))):
        kind = str(rng.choice(["soup", "big", "sliver", "closed"]))
        xyz, rgb, *_ = _mesh(rng, kind, W, H, U)
        R = meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3))
        sh = _random_shader(rng)
        m = sloth.SimpleMesh(xyz, rgb)
        keep.append(m)
        sloth.draw_mesh(ctx, m, R, sh)
        draws.append((xyz, rgb, R, sh))
    say(f"case random context seed {seed} [{path}]: {W}x{H} image={image} draws={len(draws)}")
    cells, z, win = ctx.read(want_winners=True)
    assert_frame_equal(cells, z, win, oracle_render(W, H, U, draws, image=image), W)
