"""Parity at the benchmark configurations' full sizes (GPU).

* the 1080p turntable sweep (327,680 triangles, 256 frames) as one graph
  launch: three frames against the oracle; every frame through properties
  that do not depend on size (a slot's two cells agree, odd z slots stay at
  f32::MAX, newline cells only in column 1, glyphs from the shader's
  alphabet); and a second replay byte-identical to the first;
* the 4K frame (1,310,720 triangles, 3840x2160) against the oracle;
* the uneven 64-mesh batch: every fourth mesh against the oracle, with the
  raster variant the plan picks and with the other one.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from _helpers import assert_frame_equal, say

pytestmark = pytest.mark.gpu

ALPHABET = set(b".:-=+*#%@ ")


def _frame_properties(cells, z, W, H):
    grid = cells[: W * H].reshape(H, W)
    zz = z.reshape(H, W)
    # odd z slots are never touched (ids are y*W + 2x)
    assert (zz[:, 1::2] == np.float32(np.finfo(np.float32).max)).all()
    # a slot writes cells id and id+1 alike, except the newline in column 1
    body = grid[:, 2::2] == grid[:, 3::2]
    assert body.all()
    nl = (grid & 0xFF) == 0x0A
    assert not nl[:, 2:].any() and not nl[:, 0].any()
    glyphs = set(np.unique(grid & 0xFF).tolist()) - {0x0A}
    assert glyphs <= ALPHABET, glyphs - ALPHABET
    # the H tail cells of image mode stay blank
    assert (cells[W * H:] == 0x20).all()


def test_sweep_1080p_full(sloth):
    from sloth_amd import meshgen

    m, W, H, U, Rs = meshgen.workload_sweep_1080p(256)
    mesh = sloth.SimpleMesh.from_mesh(m)
    b = sloth.Batch([sloth.FrameSpec(W, H, U, True) for _ in Rs],
                    [sloth.DrawSpec(f, mesh, R) for f, R in enumerate(Rs)])
    say("case full sweep 256 x 1920x1080, 327680 triangles: graph run")
    b.run()
    b.sync()
    idx = [0, 127, 255]
    ref_cells, ref_z = oracle.render_frames(W, H, U, Rs[idx], m.xyz, m.rgb, image=True, nthreads=3)
    first = {}
    for f in range(256):
        cells, z, _ = b.read_frame(f)
        _frame_properties(cells, z, W, H)
        if f in idx:
            j = idx.index(f)
            assert np.array_equal(cells, ref_cells[j]), f"frame {f}: cells differ"
            assert np.array_equal(z.view(np.uint32), ref_z[j].view(np.uint32)), f"frame {f}: z differs"
        if f % 32 == 0:
            first[f] = (cells.copy(), z.copy())
    say("case full sweep: second replay")
    b.run()
    b.sync()
    for f, (c0, z0) in first.items():
        cells, z, _ = b.read_frame(f)
        assert np.array_equal(cells, c0) and np.array_equal(z.view(np.uint32), z0.view(np.uint32)), f


def test_frame_4k_full(sloth):
    from sloth_amd import meshgen

    m, W, H, U, R = meshgen.workload_icosphere_4k()
    say(f"case 4K frame: {m.n} triangles, {W}x{H}")
    mesh = sloth.SimpleMesh.from_mesh(m)
    b = sloth.Batch([sloth.FrameSpec(W, H, U, True)], [sloth.DrawSpec(0, mesh, R)])
    b.run()
    b.sync()
    cells, z, win = b.read_frame(0, want_winners=True)
    ref = oracle.render(W, H, U, m.xyz, m.rgb, R)
    assert_frame_equal(cells, z, win, ref, W)
    _frame_properties(cells, z, W, H)


@pytest.mark.parametrize("big_walk", ["auto", "0"])
def test_uneven_batch_full(sloth, big_walk, monkeypatch):
    """The plan picks the big-box raster variant for this batch (about 3% of
    its triangles have boxes of more than 128 pixels); "0" forces the other."""
    from sloth_amd import meshgen

    if big_walk == "auto":
        monkeypatch.delenv("SLOTH_BIG_WALK", raising=False)
    else:
        monkeypatch.setenv("SLOTH_BIG_WALK", big_walk)

    items = meshgen.workload_uneven_meshes()
    frames, draws, keep = [], [], []
    for i, (m, W, H, U, R) in enumerate(items):
        sm = sloth.SimpleMesh.from_mesh(m)
        keep.append(sm)
        frames.append(sloth.FrameSpec(W, H, U, True))
        draws.append(sloth.DrawSpec(i, sm, R))
    say(f"case uneven batch (big walk {big_walk}): 64 meshes, {sum(it[0].n for it in items)} triangles")
    b = sloth.Batch(frames, draws)
    b.run()
    b.sync()
    for i in range(0, 64, 4):
        m, W, H, U, R = items[i]
        cells, z, _ = b.read_frame(i)
        ref = oracle.render(W, H, U, m.xyz, m.rgb, R)
        assert_frame_equal(cells, z, None, ref, W)


def test_uneven_batch_autotuned(sloth):
    """The uneven 64-mesh batch after the autotuner picked its tile size and
    raster variant: every fourth mesh against the oracle, on two replays."""
    from sloth_amd import meshgen

    items = meshgen.workload_uneven_meshes()
    frames, draws, keep = [], [], []
    for i, (m, W, H, U, R) in enumerate(items):
        sm = sloth.SimpleMesh.from_mesh(m)
        keep.append(sm)
        frames.append(sloth.FrameSpec(W, H, U, True))
        draws.append(sloth.DrawSpec(i, sm, R))
    b = sloth.Batch(frames, draws)
    tw, th = b.autotune()
    plan = b.plan()
    say(f"case uneven batch autotuned: tile {tw}x{th}, big walk {plan['big_walk']}")
    assert (plan["tile_w"], plan["tile_h"]) == (tw, th)
    for rep in range(2):
        b.run()
        b.sync()
        for i in range(0, 64, 4):
            m, W, H, U, R = items[i]
            cells, z, _ = b.read_frame(i)
            ref = oracle.render(W, H, U, m.xyz, m.rgb, R)
            assert_frame_equal(cells, z, None, ref, W)


def test_frame_past_the_tile_key(sloth):
    """A frame of 4 x 1,310,720 = 5,242,880 triangles (more than the tiled
    path's 2^22 - 1 per frame) takes the general global-key path by itself
    (the plan has no tiles), beside a small frame in the same batch, and
    both match the oracle."""
    from sloth_amd import meshgen

    m, _, _, _, R = meshgen.workload_icosphere_4k()
    small = meshgen.uv_sphere(16, 8, seed=50)
    W, H = 160, 48
    U = meshgen.screen_transform(W, H, [m.bbox_max()])
    U2 = meshgen.screen_transform(64, 30, [small.bbox_max()])
    Rs = [meshgen.rotation_euler(0.3 * k, 0.7 * k, 0.1) for k in range(4)]
    sm, ss = sloth.SimpleMesh.from_mesh(m), sloth.SimpleMesh.from_mesh(small)
    say(f"case frame of {4 * m.n} triangles at {W}x{H} beside a small frame")
    b = sloth.Batch([sloth.FrameSpec(W, H, U, True), sloth.FrameSpec(64, 30, U2, True)],
                    [sloth.DrawSpec(0, sm, R_) for R_ in Rs] + [sloth.DrawSpec(1, ss, R)])
    b.run()
    b.sync()
    assert b.plan()["tile_w"] == 0
    ref = oracle_render_draws(W, H, U, [(m.xyz, m.rgb, R_) for R_ in Rs])
    cells, z, win = b.read_frame(0, want_winners=True)
    assert_frame_equal(cells, z, win, ref, W)
    ref2 = oracle_render_draws(64, 30, U2, [(small.xyz, small.rgb, R)])
    cells, z, win = b.read_frame(1, want_winners=True)
    assert_frame_equal(cells, z, win, ref2, 64)


def oracle_render_draws(W, H, U, draws):
    f = oracle.Frame(W, H, True)
    for xyz, rgb, R in draws:
        f.draw_mesh(U, xyz, rgb, R)
    return f
