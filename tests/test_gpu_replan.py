"""Transforms replaced between renders of a batch (the turntable schedule of
main.rs:57, 92-97 without a rebuild): the graph replay after
sloth_batch_set_transforms, the tile-list overflow that new transforms can
cause (lists sized by an earlier counting run), the re-size and re-render
that recover from it, the shading visibility after new transforms, and
renders launched on a caller's stream.  Every frame against the oracle."""
from __future__ import annotations

import math

import numpy as np
import pytest

from _helpers import assert_frame_equal, oracle_render, say

pytestmark = pytest.mark.gpu


def _scaled(R, s, tx=0.0):
    """R * diag(s, s, s, 1) followed by a model-space x translation."""
    M = np.asarray(R, np.float32).reshape(4, 4).T.astype(np.float64)    # row-major view
    S = np.diag([s, s, s, 1.0])
    S[0, 3] = tx
    return np.ascontiguousarray((M @ S).T.astype(np.float32).reshape(16))


def _check_frames(b, m, frames, Rs, W, winners=True):
    for f, (fs, R) in enumerate(zip(frames, Rs)):
        ref = oracle_render(fs.width, fs.height, fs.utransform, [(m.xyz, m.rgb, R, None)], image=fs.image)
        cells, z, win = b.read_frame(f, want_winners=winners)
        assert_frame_equal(cells, z, win, ref, W)


def test_set_transforms_overflow_replan(sloth):
    """Frames first drawn with the mesh off screen (every tile list sized
    for 0 entries plus headroom), then given transforms that fill them: the
    replayed graph overflows its tile lists, the synchronising call re-sizes
    them, renders again and drops the graph; the next run captures a new one.
    Reading between set_transforms and the next run is refused."""
    from sloth_amd import meshgen

    m = meshgen.displaced_icosphere(3, seed=21)
    W, H, F = 192, 96, 5
    U = meshgen.screen_transform(W, H, [m.bbox_max()])
    frames = [sloth.FrameSpec(W, H, U, image=(f % 2 == 0)) for f in range(F)]
    mesh = sloth.SimpleMesh.from_mesh(m)
    base = [meshgen.rotation_euler(0.1 * f, 1.0 + 0.4 * f, 0.0) for f in range(F)]
    off = [_scaled(R, 1.0, tx=50.0) for R in base]           # entirely off screen
    b = sloth.Batch(frames, [sloth.DrawSpec(f, mesh, off[f]) for f in range(F)])
    say("case replan: off-screen first render")
    b.run()
    b.sync()
    _check_frames(b, m, frames, off, W, winners=False)

    big = [_scaled(R, 1.7) for R in base]                    # fills and overflows the frame
    b.set_transforms(np.stack(big))
    with pytest.raises(sloth.SlothError) as e:
        b.read_frame(0)
    assert e.value.code == sloth.SLOTH_ESTATE
    for rep in range(2):
        say(f"case replan: zoomed transforms, run {rep}")
        b.run()
        b.sync()
        _check_frames(b, m, frames, big, W, winners=(rep == 1))

    # winners asked for right after a run (read_frame synchronises and
    # checks the run first)
    small = [_scaled(R, 0.6) for R in reversed(base)]
    b.set_transforms(np.stack(small))
    say("case replan: smaller transforms, winners without an explicit sync")
    b.run()
    _check_frames(b, m, frames, small, W, winners=True)

    with pytest.raises(ValueError):
        b.set_transforms(np.stack(small[:-1]))


def test_shading_visibility_after_new_transforms(sloth):
    """sloth_shading_visibility after set_transforms: the winners and the
    forward shading follow the new transforms (the zoom also overflows the
    tile lists sized at creation)."""
    import oracle
    from sloth_amd import meshgen

    m = meshgen.uv_sphere(32, 16, 1.0, seed=8)
    W, H, V = 128, 64, 3
    rng = np.random.default_rng(8)
    U = meshgen.screen_transform(W, H, [m.bbox_max()])
    mesh = sloth.SimpleMesh.from_mesh(m)
    R0 = [_scaled(meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3)), 0.3) for _ in range(V)]
    b = sloth.Batch([sloth.FrameSpec(W, H, U, True) for _ in range(V)],
                    [sloth.DrawSpec(f, mesh, R0[f]) for f in range(V)])
    sh = sloth.Shading(b, m.faces, m.verts.shape[0])
    vc = rng.uniform(0, 1, (m.verts.shape[0], 3)).astype(np.float32)
    vn = (m.verts / np.linalg.norm(m.verts, axis=1, keepdims=True)).astype(np.float32)
    L = np.array([0.577, 0.577, -0.577], np.float32)
    R1 = [_scaled(meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3)), 1.4) for _ in range(V)]
    b.set_transforms(np.stack(R1))
    say("case shading visibility after set_transforms")
    sh.visibility()
    _, out = sh.forward(vc, vn, L, want_out=True)
    nz = W * H
    for f, R in enumerate(R1):
        fr = oracle.Frame(W, H, True)
        fr.draw_mesh(U, m.xyz, m.rgb, R)
        cells, z, win = b.read_frame(f, want_winners=True)
        assert_frame_equal(cells, z, win, fr, W)
        _, o, _ = oracle.shade_f64(fr, oracle.mat4_mul(U, R), m.xyz, m.faces, vc, vn, L, want_out=True)
        assert np.abs(out[f * nz:(f + 1) * nz] - o).max() < 1e-5


def _hip():
    """The HIP runtime the library runs on (same soname, so the one already
    loaded with it), for a caller-owned stream.  Torch is not used: it
    bundles a second HIP runtime whose stream handles this one cannot read."""
    import ctypes as C

    sloth_lib = __import__("sloth_amd").load_library()   # noqa: F841  (loads the runtime first)
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


def test_render_on_caller_stream(sloth):
    """Batch and context renders launched on a caller's stream, read back
    with no synchronisation of that stream by the caller: the handle's own
    stream waits for the launch."""
    import ctypes as C

    from sloth_amd import meshgen

    hip = _hip()
    s = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(s), 1) == 0   # hipStreamNonBlocking
    try:
        m = meshgen.displaced_icosphere(4, seed=5)
        W, H, F = 640, 240, 8
        U = meshgen.screen_transform(W, H, [m.bbox_max()])
        mesh = sloth.SimpleMesh.from_mesh(m)
        Rs = [meshgen.rotation_euler(0.0, 0.7 * f, 0.2) for f in range(F)]
        frames = [sloth.FrameSpec(W, H, U, True) for _ in range(F)]
        b = sloth.Batch(frames, [sloth.DrawSpec(f, mesh, Rs[f]) for f in range(F)])
        say("case batch on a caller's stream")
        for _ in range(3):
            b.run(s.value)
        _check_frames(b, m, frames, Rs, W, winners=False)

        say("case context on a caller's stream")
        ctx = sloth.Context(W, H, image=True)
        ctx.utransform = U
        ctx.clear()
        sloth.draw_mesh(ctx, mesh, Rs[3])
        ctx.render(s.value)
        cells, z, _ = ctx.read()
        ref = oracle_render(W, H, U, [(m.xyz, m.rgb, Rs[3], None)])
        assert_frame_equal(cells, z, None, ref, W)
        del b, ctx
    finally:
        hip.hipStreamDestroy(s)
