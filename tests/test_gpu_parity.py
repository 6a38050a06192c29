"""GPU raster vs the CPU oracle, bit for bit: frame-buffer cells (glyph +
RGB), z buffer bits and the winning triangle of every z slot.

Every case is small enough for the oracle to finish in seconds; the
full-size workloads are covered in test_gpu_fullsize.py."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from _helpers import assert_frame_equal, gpu_render, oracle_render, random_soup, say

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["tiled", "tiled-big", "global"])
def raster_path(request, monkeypatch):
    """Every case runs on the tiled path with each tile raster variant
    (SLOTH_BIG_WALK=0: every lane walks its own box; 1: wavefronts walk boxes
    of more than 128 pixels together; frames of odd width take the odd-width
    instance either way) and on the general global-key path (forced by
    SLOTH_FORCE_GLOBAL=1)."""
    if request.param == "global":
        monkeypatch.setenv("SLOTH_FORCE_GLOBAL", "1")
    else:
        monkeypatch.delenv("SLOTH_FORCE_GLOBAL", raising=False)
    monkeypatch.setenv("SLOTH_BIG_WALK", "1" if request.param == "tiled-big" else "0")
    return request.param


def _check(sloth, name, W, H, U, draws, image=True):
    import os

    path = "global" if os.environ.get("SLOTH_FORCE_GLOBAL") == "1" else ("tiled-odd" if W % 2 else "tiled")
    if path == "tiled" and os.environ.get("SLOTH_BIG_WALK") == "1":
        path = "tiled-big"
    say(f"case {name} [{path}]: {W}x{H} image={image} draws={len(draws)} "
        f"tris={sum(d[0].shape[0] for d in draws)}")
    cells, z, win = gpu_render(sloth, W, H, U, draws, image=image)
    ref = oracle_render(W, H, U, draws, image=image)
    assert_frame_equal(cells, z, win, ref, W)
    return ref


def test_sphere_160x48(sloth):
    from sloth_amd import meshgen

    m, W, H, U, R = meshgen.workload_sphere_160x48()
    ref = _check(sloth, "uv-sphere 960", W, H, U, [(m.xyz, m.rgb, R, None)])
    assert ref.stats[2] > 0


def test_sphere_terminal_mode(sloth):
    from sloth_amd import meshgen

    m, W, H, U, R = meshgen.workload_sphere_160x48()
    _check(sloth, "uv-sphere terminal", W, H, U, [(m.xyz, m.rgb, R, None)], image=False)


@pytest.mark.parametrize("W,H", [(160, 48), (37, 23), (101, 64), (64, 64), (2, 2), (1, 5), (5, 1), (3, 3)])
def test_random_soups(sloth, W, H):
    rng = np.random.default_rng(W * 1000 + H)
    from sloth_amd import meshgen

    U = meshgen.screen_transform(W, H, [np.array([1.0, 1.0, 1.0], np.float32)])
    draws = []
    for k in range(3):
        xyz, rgb = random_soup(rng, 400, W, H, U, spread=1.4, size=0.08 * (k + 1))
        R = meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3))
        draws.append((xyz, rgb, R, None))
    _check(sloth, f"random soups {W}x{H}", W, H, U, draws)


@pytest.mark.parametrize("bin_tree", ["auto", "0", "1"])
@pytest.mark.parametrize("W,H", [(301, 600), (300, 600)])
def test_tall_frames_newline_rows(sloth, W, H, bin_tree, monkeypatch):
    """Frames taller than a block of 256 triangles can span: a UV sphere
    (each block's triangles share a narrow band of rows, away from row 0),
    scattered small triangles and a few that cover most of the frame (a
    block's rows span more than 256).  The newline rows and the frame's
    cells must match the oracle on each path, with the binning kernel's
    newline rows per row or through its segment tree (SLOTH_BIN_TREE)."""
    if bin_tree != "auto":
        monkeypatch.setenv("SLOTH_BIN_TREE", bin_tree)
    rng = np.random.default_rng(W + H)
    from sloth_amd import meshgen

    U = meshgen.screen_transform(W, H, [np.array([1.0, 1.0, 1.0], np.float32)])
    sph = meshgen.uv_sphere(48, 40, 0.9, seed=9)
    R0 = meshgen.rotation_euler(0.3, 1.1, 0.2)
    draws = [(sph.xyz, sph.rgb, R0, None)]
    # stretched 3.5x in y: the frame's rows hold about +-3.7 model units
    S = np.diag([1.0, 3.5, 1.0, 1.0]).astype(np.float32).reshape(16)
    xyz, rgb = random_soup(rng, 700, W, H, U, spread=1.0, size=0.03)
    draws.append((xyz, rgb, S, None))
    big = rng.uniform(-1.0, 1.0, size=(40, 3, 3)).astype(np.float32)
    draws.append((big, rng.integers(0, 256, size=(40, 3), dtype=np.uint8), S, None))
    _check(sloth, f"tall frame {W}x{H}", W, H, U, draws)


def test_custom_shaders_per_draw(sloth):
    from sloth_amd import meshgen, Shader

    rng = np.random.default_rng(7)
    W, H = 120, 40
    U = meshgen.screen_transform(W, H, [np.array([1.0, 1.0, 1.0], np.float32)])
    sh1 = Shader((0.5, 0.9), "ab", "z")
    sh2 = Shader((-0.1, 0.25, 0.75, 1.0, 2.0), "01234", "#")
    # not non-decreasing, and with a repeated threshold: the first-match scan
    sh3 = Shader((0.9, 0.3, 0.6, 0.3, 0.95), "vwxyz", "q")
    # non-decreasing with repeats and an infinite step
    sh4 = Shader((0.25, 0.25, 0.5, float("inf")), "abcd", "e")
    # the full 16 steps (shades above the last one take the fallback), no
    # step at all, and one step; with 8 shaders in the job, the ones past the
    # first four are read from global memory rather than LDS
    sh5 = Shader(tuple(-0.05 + 0.0625 * k for k in range(16)), "ABCDEFGHIJKLMNOP", "!")
    sh6 = Shader((), "", "0")
    sh7 = Shader((0.5,), "1", "2")
    draws = []
    for sh in (sh1, sh2, sh3, sh4, sh5, sh6, sh7, None):
        xyz, rgb = random_soup(rng, 300, W, H, U, size=0.1)
        draws.append((xyz, rgb, meshgen.rotation_euler(0.2, 1.1, 0.3), sh))
    _check(sloth, "per-draw shaders", W, H, U, draws)


def test_wraparound_aliasing(sloth):
    """Logical x >= W/2 maps into the next row (id = y*W + 2x,
    rasterizer.rs:80); triangles there and the newline cells they overwrite."""
    from sloth_amd import meshgen

    rng = np.random.default_rng(11)
    for W, H in ((80, 30), (81, 30)):
        U = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1], np.float32)
        # triangles directly in screen space, right half of the logical range
        n = 600
        cx = rng.uniform(W / 2 - 4, W - 1, n)
        cy = rng.uniform(0, H, n)
        xyz = np.stack([np.stack([cx + rng.normal(0, 2, n), cy + rng.normal(0, 2, n),
                                  rng.uniform(-5, 5, n)], 1) for _ in range(3)], 1).astype(np.float32)
        rgb = rng.integers(0, 256, (n, 3), dtype=np.uint8)
        _check(sloth, f"wrap {W}x{H}", W, H, U, [(xyz, rgb, U, None)])


def test_degenerate_and_extreme(sloth):
    """Points, segments, collinear slivers, horizontal lines on integer rows,
    NaN and infinite vertices, huge coordinates, full-screen triangles."""
    W, H = 64, 32
    I = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1], np.float32)
    tris = [
        [[10, 10, 1], [10, 10, 1], [10, 10, 1]],            # point
        [[5, 7, 2], [20, 7, 2], [12, 7, 2]],                # horizontal line on a row
        [[5, 8, 2], [5, 8, 2], [30, 8, 3]],                 # segment with repeated vertex
        [[3, 3, 0], [13, 9, 0], [23, 15, 0]],               # collinear
        [[0, 0, -1], [0, 40, -1], [80, 0, -1]],             # covers the screen, front
        [[0, 0, -2], [80, 0, -2], [0, 40, -2]],             # same, back-facing
        [[1e30, 1, 0], [2, 1e30, 0], [3, 3, 0]],            # huge
        [[np.nan, 4, 0], [10, 20, 0], [20, 4, 0]],          # NaN vertex
        [[np.inf, 4, 0], [10, 20, 0], [20, 4, 0]],          # inf vertex
        [[-np.inf, 4, 0], [10, 20, 0], [20, 4, 0]],
        [[10, 4, np.nan], [10, 20, 0], [20, 4, 0]],         # NaN depth
        [[30, 10, 0], [31, 10.5, 0], [30.5, 11, 0]],        # sub-pixel
        [[40, 5, 5], [60, 5, 5], [40, 25, 5]],              # right-angle, integer
        [[40, 5, 5], [40, 25, 5], [60, 5, 5]],              # same, other winding
        [[40, 5, 5], [60, 5, 5], [40, 25, 5]],              # exact duplicate (tie)
        [[-10, -10, 0], [-5, -10, 0], [-10, -5, 0]],        # off screen
        [[1, 1, 3.4e38], [30, 1, 3.4e38], [1, 30, 3.4e38]], # depth at f32::MAX
        [[2, 2, -0.0], [20, 2, 0.0], [2, 20, -0.0]],        # signed zero depth
        [[2, 2, 0.0], [20, 2, 0.0], [2, 20, 0.0]],
        [[31.999999, 3, 1], [32.000001, 3, 1], [32, 29, 1]],  # needle at the wrap column
    ]
    xyz = np.array(tris, np.float32)
    rgb = (np.arange(len(tris) * 3).reshape(-1, 3) * 37 % 256).astype(np.uint8)
    _check(sloth, "degenerate/extreme", W, H, I, [(xyz, rgb, I, None)])
    _check(sloth, "degenerate/extreme reversed", W, H, I, [(xyz[::-1].copy(), rgb[::-1].copy(), I, None)])


def test_slivers_near_silhouette(sloth):
    """Nearly edge-on triangles (orient close to 0), where the FP32 edge test
    can disagree with exact arithmetic."""
    rng = np.random.default_rng(5)
    W, H = 200, 60
    I = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1], np.float32)
    n = 2000
    p = rng.uniform([2, 2], [98, 58], (n, 2))
    d = rng.normal(0, 6, (n, 2))
    t = rng.uniform(0.2, 0.8, n)
    eps = rng.normal(0, 1e-4, (n, 2))
    v1 = p
    v2 = p + d
    v3 = p + t[:, None] * d + eps
    z = rng.uniform(-10, 10, (n, 3, 1))
    xyz = np.concatenate([np.stack([v1, v2, v3], 1), z], 2).astype(np.float32)
    rgb = rng.integers(0, 256, (n, 3), dtype=np.uint8)
    _check(sloth, "slivers", W, H, I, [(xyz, rgb, I, None)])


@pytest.mark.parametrize("kind", ["flat-edges", "steep-edges", "wide", "offscreen"])
def test_row_span_limits(sloth, kind):
    """Triangles whose boxes are at least 6 pixels wide, so that the tiled
    walk cuts each row to its span: edges within 1e-3 of horizontal or
    vertical, wide slivers, large triangles reaching far off screen.  Every
    pixel the FP32 test passes must lie inside the computed spans."""
    rng = np.random.default_rng({"flat-edges": 11, "steep-edges": 12, "wide": 13, "offscreen": 14}[kind])
    W, H = 240, 90
    I = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1], np.float32)
    n = 1500
    c = rng.uniform([0, 0], [W / 2, H], (n, 2))
    if kind == "flat-edges":
        # one edge nearly horizontal, the third vertex well above or below
        a = c + rng.uniform(-20, 20, (n, 2)) * [1, 0] + rng.normal(0, 1e-3, (n, 2)) * [0, 1]
        b = c + rng.uniform(-20, 20, (n, 2)) * [1, 0] + rng.normal(0, 1e-3, (n, 2)) * [0, 1]
        d = c + rng.uniform(-15, 15, (n, 2))
    elif kind == "steep-edges":
        a = c + rng.uniform(-20, 20, (n, 2)) * [0, 1] + rng.normal(0, 1e-3, (n, 2)) * [1, 0]
        b = c + rng.uniform(-20, 20, (n, 2)) * [0, 1] + rng.normal(0, 1e-3, (n, 2)) * [1, 0]
        d = c + rng.uniform(-15, 15, (n, 2))
    elif kind == "wide":
        # long thin triangles in every direction
        ang = rng.uniform(0, 2 * math.pi, n)
        L = rng.uniform(8, 60, n)
        u = np.stack([np.cos(ang), np.sin(ang)], 1)
        a = c - u * L[:, None] / 2
        b = c + u * L[:, None] / 2
        d = c + np.stack([-u[:, 1], u[:, 0]], 1) * rng.normal(0, 0.8, (n, 1))
    else:
        # large triangles with vertices up to 5000 pixels away
        a = c + rng.normal(0, 1500, (n, 2))
        b = c + rng.uniform(-30, 30, (n, 2))
        d = c + rng.uniform(-30, 30, (n, 2))
    z = rng.uniform(-10, 10, (n, 3, 1))
    xyz = np.concatenate([np.stack([a, b, d], 1), z], 2).astype(np.float32)
    rgb = rng.integers(0, 256, (n, 3), dtype=np.uint8)
    _check(sloth, f"row spans {kind}", W, H, I, [(xyz, rgb, I, None)])


def test_coplanar_ties(sloth):
    """Identical triangles drawn twice: the first one drawn keeps every z slot
    (strict `<`, rasterizer.rs:81)."""
    from sloth_amd import meshgen

    m = meshgen.uv_sphere(16, 8, seed=4)
    W, H = 90, 40
    U = meshgen.screen_transform(W, H, [m.bbox_max()])
    R = meshgen.rotation_euler(0.0, 3.5, 0.0)
    rgb2 = (255 - m.rgb).astype(np.uint8)
    ref = _check(sloth, "ties", W, H, U, [(m.xyz, m.rgb, R, None), (m.xyz, rgb2, R, None)])
    assert (ref.win[ref.win >= 0] < m.n).all()


def test_skull_frames(sloth):
    """The reference's skull.obj (4 meshes, 3185 triangles) at 100x100 for a
    sample of the `-j 100` turntable frames, every mesh a separate draw."""
    import os

    from sloth_amd import meshgen

    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "skull_datajs.npz"))
    xyz, rgb, counts = d["xyz"], d["rgb"], d["mesh_counts"]
    offs = np.concatenate([[0], np.cumsum(counts)])
    W, H = int(d["width"]), int(d["height"])
    U = meshgen.screen_transform(W, H, [oracle.mesh_bbox(xyz[offs[i]:offs[i + 1]])[1]
                                        for i in range(len(counts))])
    pitches = meshgen.turntable_pitches(100)
    for k in (0, 17, 50, 99):
        R = oracle.rotation_euler(0.0, pitches[k], 0.0)
        draws = [(xyz[offs[i]:offs[i + 1]], rgb[offs[i]:offs[i + 1]], R, None) for i in range(len(counts))]
        _check(sloth, f"skull frame {k}", W, H, U, draws)


def test_batch_matches_oracle(sloth):
    """A batch of frames of different sizes, draws interleaved across frames,
    rendered by the graph launch and by the per-frame launches."""
    from sloth_amd import meshgen

    rng = np.random.default_rng(21)
    sizes = [(80, 24), (64, 64), (37, 19), (160, 48), (2, 2)]
    meshes = [meshgen.uv_sphere(12, 6, seed=1), meshgen.displaced_icosphere(2, seed=2),
              meshgen.subdivided_box(3, seed=3)]
    smeshes = [sloth.SimpleMesh.from_mesh(m) for m in meshes]
    frames, draws, plan = [], [], []
    for f, (W, H) in enumerate(sizes):
        U = meshgen.screen_transform(W, H, [m.bbox_max() for m in meshes])
        frames.append(sloth.FrameSpec(W, H, U, image=(f % 2 == 0)))
    for k in range(9):
        f = int(rng.integers(0, len(sizes)))
        mi = int(rng.integers(0, len(meshes)))
        R = meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3))
        draws.append(sloth.DrawSpec(f, smeshes[mi], R))
        plan.append((f, mi, R))
    b = sloth.Batch(frames, draws)
    for mode in ("graph", "per-frame", "graph-again"):
        say(f"case batch {mode}")
        if mode == "per-frame":
            b.run_per_frame()
        else:
            b.run()
        b.sync()
        for f, fs in enumerate(frames):
            fd = [(meshes[mi].xyz, meshes[mi].rgb, R, None) for (ff, mi, R) in plan if ff == f]
            ref = oracle_render(fs.width, fs.height, fs.utransform, fd, image=fs.image)
            # cells and depths as the launch left them (asking for winners re-renders the frame)
            cells, z, _ = b.read_frame(f)
            assert_frame_equal(cells, z, None, ref, fs.width)
            cells, z, win = b.read_frame(f, want_winners=True)
            assert_frame_equal(cells, z, win, ref, fs.width)


def _slivers(rng, n, spread=1.0):
    """Nearly collinear triangles: the box bound cannot place them, so their
    candidate boxes come from exact per-row spans."""
    a = rng.uniform(-spread, spread, size=(n, 3))
    d = rng.normal(0.0, 0.3, size=(n, 3))
    perp = rng.normal(0.0, 1.0, size=(n, 3)) * rng.choice([1e-3, 1e-4, 1e-5, 0.0], size=(n, 1))
    xyz = np.stack([a, a + d, a + 0.5 * d + perp], 1).astype(np.float32)
    return xyz, rng.integers(0, 256, size=(n, 3), dtype=np.uint8)


def test_binning_state_across_launches(sloth, raster_path):
    """Renders of even-width frames launch no clear kernel: each raster
    block leaves its tile cursor, its newline rows and the careful list's
    length as the clear sets them.  Graph replays, per-frame launches, a
    re-planned tile size, a one-frame re-render for winners and new
    transforms follow one another here (image frames with newline rows,
    slivers that take the careful pass), every frame against the oracle
    after each."""
    from sloth_amd import meshgen

    if raster_path == "global":
        pytest.skip("the binning state belongs to the tiled path")
    rng = np.random.default_rng(77)
    m = meshgen.displaced_icosphere(3, seed=5)
    sm = sloth.SimpleMesh.from_mesh(m)
    sxyz, srgb = _slivers(rng, 60)
    sliv = sloth.SimpleMesh(sxyz, srgb)
    sizes = [(96, 40, True), (128, 33, True), (64, 64, False), (200, 20, True)]
    frames, draws, plan = [], [], []
    for f, (W, H, image) in enumerate(sizes):
        U = meshgen.screen_transform(W, H, [m.bbox_max(), np.array([1.0, 1.0, 1.0], np.float32)])
        frames.append(sloth.FrameSpec(W, H, U, image))
        for mesh, xyz, rgb in ((sm, m.xyz, m.rgb), (sliv, sxyz, srgb)):
            R = meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3))
            draws.append(sloth.DrawSpec(f, mesh, R))
            plan.append((f, xyz, rgb, R))
    b = sloth.Batch(frames, draws)

    def check(step, winners_of=None):
        say(f"case binning state: {step}")
        for f, (W, H, image) in enumerate(sizes):
            U = frames[f].utransform
            ref = oracle_render(W, H, U, [(xyz, rgb, R, None) for (ff, xyz, rgb, R) in plan if ff == f], image=image)
            cells, z, win = b.read_frame(f, want_winners=(f == winners_of))
            assert_frame_equal(cells, z, win if f == winners_of else None, ref, W)

    for step in ("graph", "per-frame", "graph again"):
        b.run_per_frame() if step == "per-frame" else b.run()
        b.sync()
        check(step, winners_of=1 if step == "per-frame" else None)
    b.set_tile(64, 16)
    b.run()
    b.sync()
    check("re-planned 64x16 tiles", winners_of=2)
    new = [meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3)) for _ in plan]
    plan = [(f, xyz, rgb, R) for (f, xyz, rgb, _), R in zip(plan, new)]
    b.set_transforms(np.stack([R.reshape(16) for R in new]))
    for step in ("new transforms", "new transforms, replayed"):
        b.run()
        b.sync()
        check(step)


@pytest.mark.parametrize("n_frames,pipeline", [(40, True), (136, True), (136, False)])
def test_batch_pipelined_groups(sloth, n_frames, pipeline, monkeypatch):
    """With SLOTH_PIPELINE=1, graph launches of 32+ frames bin one frame group
    while rastering the previous one; each group's careful triangles (slivers)
    come from its own range of the shared careful list.  Every frame against
    the oracle, on two replays of the graph."""
    from sloth_amd import meshgen

    monkeypatch.setenv("SLOTH_PIPELINE", "1" if pipeline else "0")
    rng = np.random.default_rng(100 + n_frames)
    W, H = 96, 40
    U = meshgen.screen_transform(W, H, [np.array([1.0, 1.0, 1.0], np.float32)])
    soup = sloth.SimpleMesh(*random_soup(rng, 300, W, H, U, size=0.06))
    sliv = sloth.SimpleMesh(*_slivers(rng, 40))
    frames = [sloth.FrameSpec(W, H, U, image=(f % 3 != 1)) for f in range(n_frames)]
    draws, plan = [], []
    for f in range(n_frames):
        if f % 11 == 5:
            continue   # a frame with no draw
        for mesh in ((soup, sliv) if f % 2 == 0 else (sliv, soup, sliv)):
            R = meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3))
            draws.append(sloth.DrawSpec(f, mesh, R))
            plan.append((f, mesh, R))
    b = sloth.Batch(frames, draws)
    refs = []
    for f, fs in enumerate(frames):
        fd = [(m.positions, m.colors, R, None) for (ff, m, R) in plan if ff == f]
        refs.append(oracle_render(W, H, U, fd, image=fs.image))
    for rep in range(2):
        say(f"case pipelined batch {n_frames} frames, replay {rep}")
        b.run()
        b.sync()
        for f in range(n_frames):
            cells, z, _ = b.read_frame(f)
            assert_frame_equal(cells, z, None, refs[f], W)


def test_context_redraw_after_read(sloth):
    """draw, read, draw more, read: the second read covers every draw since
    the clear, in draw order."""
    from sloth_amd import meshgen

    m1 = meshgen.uv_sphere(16, 8, seed=1)
    m2 = meshgen.subdivided_box(2, seed=2)
    W, H = 70, 30
    U = meshgen.screen_transform(W, H, [m1.bbox_max(), m2.bbox_max()])
    R1 = meshgen.rotation_euler(0.1, 0.7, 0.0)
    R2 = meshgen.rotation_euler(0.0, 2.0, 0.4)
    ctx = sloth.Context(W, H, image=True)
    ctx.utransform = U
    ctx.clear()
    g1, g2 = sloth.SimpleMesh.from_mesh(m1), sloth.SimpleMesh.from_mesh(m2)
    sloth.draw_mesh(ctx, g1, R1)
    c1, z1, w1 = ctx.read(want_winners=True)
    ref = oracle_render(W, H, U, [(m1.xyz, m1.rgb, R1, None)])
    assert_frame_equal(c1, z1, w1, ref, W)
    sloth.draw_mesh(ctx, g2, R2)
    c2, z2, w2 = ctx.read(want_winners=True)
    ref = oracle_render(W, H, U, [(m1.xyz, m1.rgb, R1, None), (m2.xyz, m2.rgb, R2, None)])
    assert_frame_equal(c2, z2, w2, ref, W)
    ctx.clear()
    c3, z3, _ = ctx.read()
    assert (c3 == oracle.CELL_BLANK).all() and (z3 == oracle.F32_MAX).all()


def test_empty_mesh_and_frame(sloth):
    W, H = 10, 6
    I = np.eye(4, dtype=np.float32).reshape(16)
    _check(sloth, "empty mesh", W, H, I, [(np.zeros((0, 3, 3), np.float32), np.zeros((0, 3), np.uint8), I, None)])
    _check(sloth, "no draws", W, H, I, [])


@pytest.mark.parametrize("tile", [(32, 8), (128, 32), (8, 4), (64, 64), (256, 8), (512, 4), (512, 16)])
def test_tile_sizes_agree(sloth, tile, raster_path):
    """Every tile size gives the oracle's frames (the wrap and newline cells
    land in different tiles for different sizes)."""
    from sloth_amd import meshgen

    if raster_path == "global":
        pytest.skip("tile sizes only exist on the tiled path")
    m = meshgen.displaced_icosphere(3, seed=9)
    frames, draws, plan = [], [], []
    sm = sloth.SimpleMesh.from_mesh(m)
    for f, (W, H) in enumerate([(200, 64), (130, 40), (64, 30)]):
        U = meshgen.screen_transform(W, H, [m.bbox_max()])
        R = meshgen.rotation_euler(0.3 * f, 1.0 + f, 0.2)
        frames.append(sloth.FrameSpec(W, H, U, True))
        draws.append(sloth.DrawSpec(f, sm, R))
        plan.append((W, H, U, R))
    b = sloth.Batch(frames, draws)
    b.set_tile(*tile)
    say(f"case tile {tile}")
    b.run()
    b.sync()
    for f, (W, H, U, R) in enumerate(plan):
        ref = oracle_render(W, H, U, [(m.xyz, m.rgb, R, None)])
        cells, z, win = b.read_frame(f, want_winners=True)
        assert_frame_equal(cells, z, win, ref, W)


def test_autotune_keeps_results(sloth, raster_path):
    from sloth_amd import meshgen

    if raster_path == "global":
        pytest.skip("tile sizes only exist on the tiled path")
    m = meshgen.displaced_icosphere(4, seed=4)
    W, H = 320, 120
    U = meshgen.screen_transform(W, H, [m.bbox_max()])
    sm = sloth.SimpleMesh.from_mesh(m)
    Rs = [meshgen.rotation_euler(0.0, 0.5 * k, 0.0) for k in range(4)]
    b = sloth.Batch([sloth.FrameSpec(W, H, U, True) for _ in Rs], [sloth.DrawSpec(k, sm, R) for k, R in enumerate(Rs)])
    tw, th = b.autotune()
    say(f"case autotune -> {tw}x{th}")
    assert tw in (16, 32, 64, 128, 256, 512) and th in (8, 16, 32, 64)
    assert b.plan()["tile_w"] == tw and b.plan()["tile_h"] == th
    assert b.plan()["bin_tris_per_thread"] in (1, 2, 4)
    b.run()
    b.sync()
    for k, R in enumerate(Rs):
        ref = oracle_render(W, H, U, [(m.xyz, m.rgb, R, None)])
        cells, z, win = b.read_frame(k, want_winners=True)
        assert_frame_equal(cells, z, win, ref, W)


def test_model_files_render(sloth):
    """Meshes from the model loader (the reference's cube.obj and cube.stl)
    uploaded, queued in file order and drawn after update, against the oracle
    on the loader's own arrays; uploaded STL meshes keep the STL bounding box."""
    import os
    from sloth_amd import meshgen

    here = os.path.dirname(os.path.abspath(__file__))
    models = os.path.join(here, "golden", "models")
    queue = sloth.load_meshes([os.path.join(models, "cube.obj"), os.path.join(models, "cube.stl")])
    assert len(queue) == 7
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    sloth.load_library().sloth_mesh_bbox(queue[-1].handle, lo.ctypes.data_as(sloth._fp), hi.ctypes.data_as(sloth._fp))
    assert np.array_equal(hi, queue[-1].bounding_box.max) and np.array_equal(lo, queue[-1].bounding_box.min)
    assert lo.min() < 0 and hi.max() > 30
    W, H = 120, 40
    U = meshgen.screen_transform(W, H, [m.bounding_box.max for m in queue])
    R = meshgen.rotation_euler(0.3, 2.2, 0.1)
    ctx = sloth.Context(W, H, image=True)
    ctx.utransform = U
    ctx.clear()
    for m in queue:
        sloth.draw_mesh(ctx, m, R)
    cells, z, win = ctx.read(want_winners=True)
    say("case model files cube.obj + cube.stl")
    ref = oracle_render(W, H, U, [(m.positions, m.colors, R, None) for m in queue])
    assert_frame_equal(cells, z, win, ref, W)
    assert ref.stats[2] > 0


def test_near_threshold_shades(sloth):
    """Small triangles whose unit normal z lies within a few ulps of
    default_shader's thresholds, so that the glyph can change from pixel to
    pixel of one triangle: the tile raster takes the cell from the entry only
    when it has proved the glyph constant over the box, and re-evaluates the
    pixel otherwise.  Drawn in screen space (identity transforms)."""
    import os
    import sys

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_glyph_bound import _triangles

    rng = np.random.default_rng(77)
    t, _, _, _ = _triangles(6000, rng, W=1280, H=720)
    W, H = 1280, 720
    I = np.eye(4, dtype=np.float32).reshape(16)
    rgb = rng.integers(0, 256, (len(t), 3), dtype=np.uint8)
    _check(sloth, "near-threshold shades", W, H, I, [(t, rgb, I, None)])


def test_indexed_meshes(sloth):
    """Indexed device meshes (vertex rows + face ids) render as the soups
    vertices[faces]: a frame of two indexed draws and one soup draw against
    the oracle on the soups, and a batch of indexed draws against the same
    batch of soups, cell for cell."""
    from sloth_amd import meshgen

    ico = meshgen.displaced_icosphere(3, seed=31)
    box = meshgen.subdivided_box(4, seed=32)
    sph = meshgen.uv_sphere(20, 10, seed=33)
    assert sloth.SimpleMesh.from_mesh(ico, indexed=True).is_indexed
    W, H = 180, 70
    U = meshgen.screen_transform(W, H, [m.bbox_max() for m in (ico, box, sph)])
    Rs = [meshgen.rotation_euler(0.3, 1.3, 0.2), meshgen.rotation_euler(1.0, 2.0, 0.5),
          meshgen.rotation_euler(0.0, 0.4, 0.0)]
    say(f"case indexed meshes {W}x{H}")
    ctx = sloth.Context(W, H, image=True)
    ctx.utransform = U
    ctx.clear()
    gm = [sloth.SimpleMesh.from_mesh(ico, indexed=True), sloth.SimpleMesh.from_mesh(box),
          sloth.SimpleMesh.from_mesh(sph, indexed=True)]
    for m, R in zip(gm, Rs):
        sloth.draw_mesh(ctx, m, R)
    cells, z, win = ctx.read(want_winners=True)
    ref = oracle_render(W, H, U, [(m.xyz, m.rgb, R, None) for m, R in zip((ico, box, sph), Rs)])
    assert_frame_equal(cells, z, win, ref, W)

    frames = [sloth.FrameSpec(W, H, U, True) for _ in range(4)]
    for indexed in (True, False):
        mesh = sloth.SimpleMesh.from_mesh(ico, indexed=indexed)
        b = sloth.Batch(frames, [sloth.DrawSpec(f, mesh, meshgen.rotation_euler(0.1 * f, 0.7 * f, 0.0))
                                 for f in range(4)])
        b.run()
        b.sync()
        out = [b.read_frame(f, want_winners=True) for f in range(4)]
        if indexed:
            first = out
        else:
            for (c0, z0, w0), (c1, z1, w1) in zip(first, out):
                assert np.array_equal(c0, c1) and np.array_equal(z0.view(np.uint32), z1.view(np.uint32))
                assert np.array_equal(w0, w1)


@pytest.mark.parametrize("tile", [(8, 4), (16, 8), (64, 16), (128, 32), (256, 8)])
def test_odd_width_tiles(sloth, tile, raster_path):
    """Frames of odd width on the tiled path: every cell column a z slot,
    cells resolved from two slots inside a tile, first-column cells finished
    across tile edges and across rows (cell (r, 0) takes the last slot of
    row r - 1).  201 and 129 columns leave a last tile one column wide at
    the 64- and 128-column tiles; terminal-mode frames have no newline
    cells.  Against the oracle, winners included."""
    from sloth_amd import meshgen

    if raster_path == "global":
        pytest.skip("tile sizes only exist on the tiled path")
    m = meshgen.displaced_icosphere(3, seed=19)
    sph = meshgen.uv_sphere(16, 8, seed=20)
    frames, draws, plan = [], [], []
    sms = [sloth.SimpleMesh.from_mesh(m), sloth.SimpleMesh.from_mesh(sph)]
    for f, (W, H, image) in enumerate([(201, 64, True), (129, 40, True), (63, 30, False), (81, 25, True)]):
        U = meshgen.screen_transform(W, H, [m.bbox_max(), sph.bbox_max()])
        Rs = [meshgen.rotation_euler(0.3 * f, 1.0 + f, 0.2), meshgen.rotation_euler(0.1, 2.0 + f, 0.4)]
        frames.append(sloth.FrameSpec(W, H, U, image))
        for sm, R in zip(sms, Rs):
            draws.append(sloth.DrawSpec(f, sm, R))
        plan.append((W, H, U, Rs, image))
    b = sloth.Batch(frames, draws)
    b.set_tile(*tile)
    say(f"case odd-width tiles {tile}")
    b.run()
    b.sync()
    for f, (W, H, U, Rs, image) in enumerate(plan):
        ref = oracle_render(W, H, U, [(m.xyz, m.rgb, Rs[0], None), (sph.xyz, sph.rgb, Rs[1], None)], image=image)
        cells, z, win = b.read_frame(f, want_winners=True)
        assert_frame_equal(cells, z, win, ref, W)


@pytest.mark.parametrize("tpt", [1, 2, 4])
def test_bin_tris_per_thread(sloth, tpt, raster_path):
    """The binning kernel's triangles per thread (1, 2 or 4) change only how
    the work is spread: dense meshes in small frames (many triangles per
    tile), a mesh wider than the frame's logical width (whole patches of
    triangles in the wrapped region), two draws in one frame and draws whose
    sizes are not multiples of a block's triangles, against the oracle."""
    from sloth_amd import meshgen

    if raster_path == "global":
        pytest.skip("the binning kernel belongs to the tiled path")
    dense = meshgen.displaced_icosphere(5, seed=31)      # 20,480 triangles
    box = meshgen.subdivided_box(9, seed=32)             # 972 triangles
    ball = meshgen.uv_sphere(40, 21, seed=33)
    spec = [((80, 24), [(dense, 1.0)]), ((200, 64), [(box, 1.7)]), ((64, 64), [(ball, 1.0), (dense, 0.6)])]
    frames, draws, keep, plan = [], [], [], []
    rng = np.random.default_rng(30)
    for f, ((W, H), items) in enumerate(spec):
        U = meshgen.screen_transform(W, H, [m.bbox_max() for m, _ in items])
        frames.append(sloth.FrameSpec(W, H, U, True))
        fd = []
        for m, zoom in items:
            R = meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3))
            R = (R.reshape(4, 4) * np.float32(zoom)).astype(np.float32).reshape(16)
            R[15] = 1.0
            sm = sloth.SimpleMesh.from_mesh(m)
            keep.append(sm)
            draws.append(sloth.DrawSpec(f, sm, R))
            fd.append((m.xyz, m.rgb, R, None))
        plan.append((W, H, U, fd))
    b = sloth.Batch(frames, draws)
    b.set_bin(tpt)
    assert b.plan()["bin_tris_per_thread"] == tpt
    say(f"case bin triangles per thread {tpt}")
    for rep in range(2):
        b.run()
        b.sync()
        for f, (W, H, U, fd) in enumerate(plan):
            ref = oracle_render(W, H, U, fd)
            cells, z, _ = b.read_frame(f)
            assert_frame_equal(cells, z, None, ref, W)
    cells, z, win = b.read_frame(1, want_winners=True)
    assert_frame_equal(cells, z, win, oracle_render(*plan[1][:3], plan[1][3]), plan[1][0])


def test_tile_that_does_not_fit(sloth, raster_path):
    """A tile the raster's workgroup cannot take (512 slots a row on an
    odd-width frame, or more than 64 KiB of LDS) is refused at the render
    with SLOTH_EINVAL before any launch; a fitting size then renders."""
    from sloth_amd import meshgen

    if raster_path == "global":
        pytest.skip("tile sizes only exist on the tiled path")
    m = meshgen.uv_sphere(16, 8, seed=40)
    W, H = 101, 40
    U = meshgen.screen_transform(W, H, [m.bbox_max()])
    R = meshgen.rotation_euler(0.2, 0.7, 0.1)
    b = sloth.Batch([sloth.FrameSpec(W, H, U, True)], [sloth.DrawSpec(0, sloth.SimpleMesh.from_mesh(m), R)])
    for tile in [(512, 8), (256, 64)]:
        b.set_tile(*tile)
        with pytest.raises(sloth.SlothError) as e:
            b.run()
        assert e.value.code == sloth.SLOTH_EINVAL
    b.set_tile(64, 16)
    b.run()
    b.sync()
    cells, z, win = b.read_frame(0, want_winners=True)
    assert_frame_equal(cells, z, win, oracle_render(W, H, U, [(m.xyz, m.rgb, R, None)]), W)


@pytest.mark.parametrize("tile", [(16, 8), (64, 16), (256, 8)])
def test_mixed_parity_batch(sloth, tile, raster_path):
    """A batch that mixes even and odd widths (1 and 2 columns included)
    stays on the tiled path, one raster instance per parity, and matches
    the oracle frame by frame, winners included."""
    from sloth_amd import meshgen

    if raster_path == "global":
        pytest.skip("tile sizes only exist on the tiled path")
    m = meshgen.displaced_icosphere(3, seed=41)
    sm = sloth.SimpleMesh.from_mesh(m)
    sizes = [(200, 64, True), (201, 64, True), (81, 25, False), (64, 30, True), (1, 5, True), (2, 3, True),
             (33, 12, True)]
    frames, draws, plan = [], [], []
    for f, (W, H, image) in enumerate(sizes):
        U = meshgen.screen_transform(W, H, [m.bbox_max()])
        R = meshgen.rotation_euler(0.2 * f, 0.9 + f, 0.1)
        frames.append(sloth.FrameSpec(W, H, U, image))
        draws.append(sloth.DrawSpec(f, sm, R))
        plan.append((W, H, U, R, image))
    b = sloth.Batch(frames, draws)
    b.set_tile(*tile)
    say(f"case mixed-parity batch, tile {tile}")
    for rep in range(2):
        b.run()
        b.sync()
        assert b.plan()["tile_w"] == tile[0]
        for f, (W, H, U, R, image) in enumerate(plan):
            ref = oracle_render(W, H, U, [(m.xyz, m.rgb, R, None)], image=image)
            cells, z, win = b.read_frame(f, want_winners=(rep == 1))
            assert_frame_equal(cells, z, win if rep == 1 else None, ref, W)


def test_long_tiles_first(sloth, raster_path):
    """A batch whose frames differ in tile grid takes a flat raster grid; a
    dense mesh in a small frame puts thousands of entries (more than four
    chunks) on one tile, which the whole-batch launch moves to the front of
    the grid.  Every frame matches the oracle through the graph launch (that
    order), the per-frame launches (tile order) and the re-render that
    returns winners."""
    from sloth_amd import meshgen

    if raster_path == "global":
        pytest.skip("tile order only exists on the tiled path")
    dense = meshgen.displaced_icosphere(5, seed=61)    # 20,480 triangles
    light = meshgen.displaced_icosphere(2, seed=62)
    sizes = [(160, 48, light), (80, 24, dense), (64, 30, light), (81, 25, dense), (200, 64, dense)]
    frames, draws, plan = [], [], []
    for f, (W, H, m) in enumerate(sizes):
        sm = sloth.SimpleMesh.from_mesh(m)
        U = meshgen.screen_transform(W, H, [m.bbox_max()])
        R = meshgen.rotation_euler(0.3 * f, 1.1 + f, 0.2)
        frames.append(sloth.FrameSpec(W, H, U, True))
        draws.append(sloth.DrawSpec(f, sm, R))
        plan.append((W, H, U, R, m, sm))
    b = sloth.Batch(frames, draws)
    b.set_tile(64, 32)
    for mode in ("graph", "per-frame", "graph-again"):
        say(f"case long tiles first, {mode}")
        if mode == "per-frame":
            b.run_per_frame()
        else:
            b.run()
        b.sync()
        for f, (W, H, U, R, m, _) in enumerate(plan):
            ref = oracle_render(W, H, U, [(m.xyz, m.rgb, R, None)], image=True)
            cells, z, _ = b.read_frame(f)
            assert_frame_equal(cells, z, None, ref, W)
            if mode == "graph-again":
                cells, z, win = b.read_frame(f, want_winners=True)
                assert_frame_equal(cells, z, win, ref, W)
