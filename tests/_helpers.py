"""Shared helpers for the parity tests: render a frame on the GPU through the
C-ABI and on the CPU oracle, and compare bit for bit."""
from __future__ import annotations

import numpy as np

import oracle


def say(msg: str) -> None:
    print(msg, flush=True)


def gpu_render(sloth, W, H, U, draws, image=True, want_winners=True):
    """draws: list of (xyz, rgb, R, shader or None)."""
    ctx = sloth.Context(W, H, image=image)
    ctx.utransform = U
    ctx.clear()
    meshes = []
    for xyz, rgb, R, sh in draws:
        m = sloth.SimpleMesh(xyz, rgb)
        meshes.append(m)
        sloth.draw_mesh(ctx, m, R, sh if sh is not None else sloth.default_shader)
    cells, z, win = ctx.read(want_winners=want_winners)
    return cells, z, win


def oracle_render(W, H, U, draws, image=True):
    f = oracle.Frame(W, H, image)
    for xyz, rgb, R, sh in draws:
        ramp = None
        if sh is not None:
            ramp = oracle.make_ramp(sh.thresholds, sh.glyphs, sh.fallback)
        f.draw_mesh(U, xyz, rgb, R, ramp)
    return f


def describe_diff(name, got, want, W):
    bad = np.nonzero(got != want)[0]
    if len(bad) == 0:
        return ""
    i = bad[0]
    return (f"{name}: {len(bad)} differ; first at {i} (row {i // W}, col {i % W}): "
            f"got {got[i]!r} want {want[i]!r}")


def assert_frame_equal(cells, z, win, ref: "oracle.Frame", W, check_win=True):
    msgs = []
    m = describe_diff("cells", cells, ref.cells, W)
    if m:
        msgs.append(m)
    m = describe_diff("z", z.view(np.uint32), ref.z.view(np.uint32), W)
    if m:
        msgs.append(m)
    if check_win and win is not None:
        m = describe_diff("winners", win, ref.win, W)
        if m:
            msgs.append(m)
    assert not msgs, "; ".join(msgs)


def random_soup(rng, n, W, H, U, spread=1.0, size=0.05):
    """Random small triangles in model space [-spread, spread]^3 with random
    winding; sizes from sub-pixel to a few pixels at W x H."""
    c = rng.uniform(-spread, spread, size=(n, 1, 3))
    d = rng.normal(0.0, size, size=(n, 3, 3))
    xyz = (c + d).astype(np.float32)
    rgb = rng.integers(0, 256, size=(n, 3), dtype=np.uint8)
    return xyz, rgb
