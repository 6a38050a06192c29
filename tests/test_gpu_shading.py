"""Differentiable shading (the build's added form): GPU forward/backward
against the FP64 CPU model and adjoint, the adjoint against central finite
differences, and the reduction to the reference's flat shade.

The error bound stated for the GPU gradients: norm-wise relative error
<= 1e-5 against the FP64 adjoint (the GPU evaluates per-fragment terms in
binary32 and sums them in binary64)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from _helpers import say

pytestmark = pytest.mark.gpu

REL = 1e-5


def rot_y(deg):
    a = math.radians(deg)
    return np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])


def setup_views(sloth, n_lon, n_lat, views, W, H, seed):
    from sloth_amd import meshgen

    m = meshgen.uv_sphere(n_lon, n_lat, 1.0, seed=seed)
    rng = np.random.default_rng(seed)
    U = meshgen.screen_transform(W, H, [m.bbox_max()])
    Rs = [meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3)) for _ in range(views)]
    mesh = sloth.SimpleMesh.from_mesh(m)
    batch = sloth.Batch([sloth.FrameSpec(W, H, U, True) for _ in range(views)],
                        [sloth.DrawSpec(f, mesh, Rs[f]) for f in range(views)])
    vcolor = rng.uniform(0, 1, (m.verts.shape[0], 3)).astype(np.float32)
    vnormal = (m.verts / np.linalg.norm(m.verts, axis=1, keepdims=True)).astype(np.float32)
    light = np.array([0.577, 0.577, -0.577], np.float32)
    return m, U, Rs, batch, vcolor, vnormal, light, rng


def oracle_frames(m, U, Rs, W, H):
    frames = []
    for R in Rs:
        f = oracle.Frame(W, H, True)
        f.draw_mesh(U, m.xyz, m.rgb, R)
        frames.append(f)
    return frames


def test_visibility_matches_oracle(sloth):
    m, U, Rs, batch, vc, vn, L, _ = setup_views(sloth, 24, 12, 3, 96, 64, seed=3)
    say("case shading visibility: uv 24x12, 3 views 96x64")
    sh = sloth.Shading(batch, m.faces, m.verts.shape[0])
    frames = oracle_frames(m, U, Rs, 96, 64)
    for f, fr in enumerate(frames):
        cells, z, win = batch.read_frame(f, want_winners=True)
        assert np.array_equal(win, fr.win)
        assert np.array_equal(cells, fr.cells)
    assert sh.n_slots == 3 * 96 * 64


# 100 x 70 and 101 x 45: backward tiles and 4-entry runs cut by the frame's
# right and bottom edges (101 is odd: its visibility comes from the odd-width
# tile raster)
@pytest.mark.parametrize("views,W,H", [(4, 128, 96), (2, 256, 256), (3, 100, 70), (2, 101, 45)])
def test_gradients_match_fp64_adjoint(sloth, views, W, H):
    m, U, Rs, batch, vc, vn, L, rng = setup_views(sloth, 48, 24, views, W, H, seed=11)
    say(f"case shading adjoint: uv 48x24 ({m.n} tris), {views} views {W}x{H}")
    sh = sloth.Shading(batch, m.faces, m.verts.shape[0])
    # target: colours + N(0, 0.1^2), light rotated 10 degrees about y
    vc_t = (vc + rng.normal(0, 0.1, vc.shape)).astype(np.float32)
    L_t = (rot_y(10.0) @ L.astype(np.float64)).astype(np.float32)
    _, target = sh.forward(vc_t, vn, L_t, want_out=True)
    sh.set_target(target)
    loss_gpu, out = sh.forward(vc, vn, L, want_out=True)
    dc, dn, dl = sh.backward()

    frames = oracle_frames(m, U, Rs, W, H)
    loss_cpu = 0.0
    gc = np.zeros_like(vc, dtype=np.float64)
    gn = np.zeros_like(vn, dtype=np.float64)
    gl = np.zeros(3)
    outs = []
    nz = W * H
    for f, (R, fr) in enumerate(zip(Rs, frames)):
        M = oracle.mat4_mul(U, R)
        lo, o, (a, b, c) = oracle.shade_f64(fr, M, m.xyz, m.faces, vc, vn, L, target[f * nz:(f + 1) * nz],
                                           want_out=True, want_grad=True)
        loss_cpu += lo
        gc += a
        gn += b
        gl += c
        outs.append(o)
    out_cpu = np.concatenate(outs)
    assert np.abs(out - out_cpu).max() < 1e-5
    assert abs(loss_gpu - loss_cpu) <= REL * abs(loss_cpu)
    for name, g, ref in (("color", dc, gc), ("normal", dn, gn), ("light", dl, gl)):
        err = np.linalg.norm(g.astype(np.float64) - ref) / np.linalg.norm(ref)
        say(f"  d{name}: norm-relative error vs FP64 adjoint {err:.2e}")
        assert err <= REL, (name, err)
    # the one-pass training step gives the same loss and gradients
    loss_s, dc_s, dn_s, dl_s = sh.step(vc, vn, L)
    assert abs(loss_s - loss_cpu) <= REL * abs(loss_cpu)
    for name, g, ref in (("color", dc_s, gc), ("normal", dn_s, gn), ("light", dl_s, gl)):
        err = np.linalg.norm(g.astype(np.float64) - ref) / np.linalg.norm(ref)
        assert err <= REL, ("step", name, err)

    # the FP64 adjoint against central differences, h = 1e-3, in float64
    h = 1e-3

    def loss64(vc64, vn64, L64):
        tot = 0.0
        for f, (R, fr) in enumerate(zip(Rs, frames)):
            M = oracle.mat4_mul(U, R)
            tot += oracle.shade_f64(fr, M, m.xyz, m.faces, vc64, vn64, L64, target[f * nz:(f + 1) * nz])[0]
        return tot

    vc64, vn64, L64 = vc.astype(np.float64), vn.astype(np.float64), L.astype(np.float64)
    picks = rng.choice(m.verts.shape[0], 4, replace=False)
    for v in picks:
        for ch in range(3):
            for which, base, grad in (("c", vc64, gc), ("n", vn64, gn)):
                p_, q_ = base.copy(), base.copy()
                p_[v, ch] += h
                q_[v, ch] -= h
                args_p = (p_, vn64, L64) if which == "c" else (vc64, p_, L64)
                args_q = (q_, vn64, L64) if which == "c" else (vc64, q_, L64)
                fd = (loss64(*args_p) - loss64(*args_q)) / (2 * h)
                assert abs(fd - grad[v, ch]) <= 1e-6 * max(1.0, abs(grad[v, ch])), (which, v, ch, fd, grad[v, ch])
    for ch in range(3):
        p_, q_ = L64.copy(), L64.copy()
        p_[ch] += h
        q_[ch] -= h
        fd = (loss64(vc64, vn64, p_) - loss64(vc64, vn64, q_)) / (2 * h)
        assert abs(fd - gl[ch]) <= 1e-6 * max(1.0, abs(gl[ch])), (ch, fd, gl[ch])


def test_reduces_to_flat_shade(sloth):
    """Per-corner normals = the face's unit normal, L = (0,0,1), colours = 1:
    the intensity is the reference's shade nz_unit * a * (w0 + w1 + w2)
    (rasterizer.rs:76) to 1e-6."""
    from sloth_amd import meshgen

    base = meshgen.uv_sphere(24, 12, seed=5)
    n = base.n
    xyz = base.xyz
    faces = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    e1 = xyz[:, 1] - xyz[:, 0]
    e2 = xyz[:, 2] - xyz[:, 0]
    fn = np.cross(e1.astype(np.float64), e2.astype(np.float64))
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    vnormal = np.repeat(fn, 3, axis=0).astype(np.float32)
    vcolor = np.ones((3 * n, 3), np.float32)
    W, H = 120, 60
    U = meshgen.screen_transform(W, H, [base.bbox_max()])
    R = meshgen.rotation_euler(0.4, 2.0, 0.1)
    say("case shading reduction to the flat shade")
    mesh = sloth.SimpleMesh(xyz, base.rgb)
    batch = sloth.Batch([sloth.FrameSpec(W, H, U, True)], [sloth.DrawSpec(0, mesh, R)])
    sh = sloth.Shading(batch, faces, 3 * n)
    _, out = sh.forward(vcolor, vnormal, np.array([0, 0, 1], np.float32), want_out=True)
    fr = oracle.Frame(W, H, True)
    fr.draw_mesh(U, xyz, base.rgb, R)
    M = oracle.mat4_mul(U, R)
    checked = 0
    for idx in np.nonzero(fr.win >= 0)[0][::7]:
        T = fr.win[idx]
        x, y = fr.winxy[2 * idx], fr.winxy[2 * idx + 1]
        t = oracle.tri_mul(M, np.concatenate([xyz[T], np.ones((3, 1), np.float32)], 1))
        nzu = oracle.tri_normal(t)[2]
        v1, v2, v3 = t[0], t[1], t[2]

        def orient(a, b, c):
            return np.float32(np.float32(np.float32(b[0] - a[0]) * np.float32(c[1] - a[1]))
                              - np.float32(np.float32(b[1] - a[1]) * np.float32(c[0] - a[0])))

        p = np.array([x, y], np.float32)
        a = np.float32(1.0) / orient(v1, v2, v3)
        w0, w1, w2 = orient(v2, v3, p), orient(v3, v1, p), orient(v1, v2, p)
        shade = np.float32(np.float32(nzu * a) * np.float32(np.float32(w0 + w1) + w2))
        assert abs(float(out[idx, 0]) - float(shade)) <= 1e-6, (idx, out[idx], shade)
        checked += 1
    assert checked > 50


def test_shading_on_indexed_mesh(sloth):
    """The shading pass reads the batch's triangles through an indexed mesh
    as through a soup: the same outputs and loss bit for bit, and the same
    gradients up to the order of their atomic sums."""
    from sloth_amd import meshgen

    m = meshgen.uv_sphere(24, 12, 1.0, seed=21)
    rng = np.random.default_rng(21)
    W, H, V = 96, 64, 2
    U = meshgen.screen_transform(W, H, [m.bbox_max()])
    Rs = [meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3)) for _ in range(V)]
    vc = rng.uniform(0, 1, (m.verts.shape[0], 3)).astype(np.float32)
    vn = (m.verts / np.linalg.norm(m.verts, axis=1, keepdims=True)).astype(np.float32)
    L = np.array([0.577, 0.577, -0.577], np.float32)
    say("case shading on an indexed mesh")
    res = []
    for indexed in (False, True):
        mesh = sloth.SimpleMesh.from_mesh(m, indexed=indexed)
        b = sloth.Batch([sloth.FrameSpec(W, H, U, True) for _ in range(V)],
                        [sloth.DrawSpec(f, mesh, Rs[f]) for f in range(V)])
        sh = sloth.Shading(b, m.faces, m.verts.shape[0])
        _, out = sh.forward(vc, vn, L, want_out=True)
        sh.set_target(np.zeros_like(out))
        res.append((out, sh.step(vc, vn, L)))
    (o0, s0), (o1, s1) = res
    assert np.array_equal(o0, o1) and s0[0] == s1[0]
    # the gradients are binary64 atomic sums, whose order varies from run to
    # run: equal to the last few binary32 bits
    for a, b in zip(s0[1:], s1[1:]):
        assert np.linalg.norm(a.astype(np.float64) - b) <= 1e-6 * np.linalg.norm(a.astype(np.float64))


@pytest.mark.parametrize("seed", range(12))
def test_random_views_match_fp64_adjoint(sloth, seed):
    """Seeded random shading jobs: frame sizes from 1 x 1 up (odd and even
    widths, widths below one 4-entry run, a single row), 1-3 views, spheres
    of random resolution.  Loss and gradients against the FP64 adjoint; a
    job whose views show no triangle has zero loss and gradients."""
    from sloth_amd import meshgen

    rng = np.random.default_rng(500 + seed)
    W = int(rng.choice([1, 2, 3, 5, 7, 31, 64, 97, 130]))
    H = int(rng.choice([1, 2, 3, 17, 40, 75]))
    views = int(rng.integers(1, 4))
    n_lon, n_lat = int(rng.integers(6, 40)), int(rng.integers(4, 24))
    m, U, Rs, batch, vc, vn, L, rng2 = setup_views(sloth, n_lon, n_lat, views, W, H, seed=600 + seed)
    say(f"case random shading seed {seed}: uv {n_lon}x{n_lat} ({m.n} tris), {views} views {W}x{H}")
    sh = sloth.Shading(batch, m.faces, m.verts.shape[0])
    vc_t = (vc + rng2.normal(0, 0.1, vc.shape)).astype(np.float32)
    L_t = (rot_y(10.0) @ L.astype(np.float64)).astype(np.float32)
    _, target = sh.forward(vc_t, vn, L_t, want_out=True)
    sh.set_target(target)
    loss_s, dc, dn, dl = sh.step(vc, vn, L)

    frames = oracle_frames(m, U, Rs, W, H)
    loss_cpu = 0.0
    gc = np.zeros_like(vc, dtype=np.float64)
    gn = np.zeros_like(vn, dtype=np.float64)
    gl = np.zeros(3)
    nz = W * H
    for f, (R, fr) in enumerate(zip(Rs, frames)):
        M = oracle.mat4_mul(U, R)
        lo, _, (a, b, c) = oracle.shade_f64(fr, M, m.xyz, m.faces, vc, vn, L, target[f * nz:(f + 1) * nz],
                                           want_out=False, want_grad=True)
        loss_cpu += lo
        gc += a
        gn += b
        gl += c
    assert abs(loss_s - loss_cpu) <= REL * abs(loss_cpu) + 1e-30
    for name, g, ref in (("color", dc, gc), ("normal", dn, gn), ("light", dl, gl)):
        nref = np.linalg.norm(ref)
        if nref == 0.0:
            assert np.linalg.norm(g.astype(np.float64)) == 0.0, name
            continue
        err = np.linalg.norm(g.astype(np.float64) - ref) / nref
        assert err <= REL, (name, err)
