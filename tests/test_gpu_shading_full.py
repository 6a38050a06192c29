"""The differentiable-shading configuration at its full size (GPU): the UV
sphere of 99,904 triangles, 32 seeded views at 1024x1024, the target from
colours + N(0, 0.1^2) and the light rotated 10 degrees (the workload the
benchmark times).

* visibility: the winners, cells and depths of 4 views bit for bit against
  the oracle (all 32 views' oracle frames feed the FP64 adjoint below);
* loss and gradients of the one-pass training step and of the
  forward/backward pair against the FP64 CPU adjoint summed over all 32
  views: norm-wise relative error <= 1e-5 per parameter group; the
  element-wise maximum relative error (over entries whose FP64 gradient is
  at least 1e-3 of the group's largest) is printed;
* the FP64 adjoint against central differences (h = 1e-3, float64) on a few
  vertices and on the light."""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from _helpers import assert_frame_equal, say

pytestmark = pytest.mark.gpu

REL = 1e-5


def test_shading_full_config(sloth):
    from sloth_amd import meshgen

    wl = meshgen.workload_shading_views()
    m, W, H, U, Rs = wl["mesh"], wl["W"], wl["H"], wl["U"], wl["Rs"]
    vc, vn, L = wl["vcolor"], wl["vnormal"], wl["light"]
    V = len(Rs)
    nz = W * H
    say(f"case shading full config: uv 224x224 ({m.n} tris), {V} views {W}x{H}")
    mesh = sloth.SimpleMesh.from_mesh(m)
    b = sloth.Batch([sloth.FrameSpec(W, H, U, True) for _ in range(V)],
                    [sloth.DrawSpec(f, mesh, Rs[f]) for f in range(V)])
    sh = sloth.Shading(b, m.faces, m.verts.shape[0])
    _, target = sh.forward(wl["vcolor_target"], vn, wl["light_target"], want_out=True)
    sh.set_target(target)
    loss_s, dc_s, dn_s, dl_s = sh.step(vc, vn, L)
    loss_f, _ = sh.forward(vc, vn, L)
    dc_f, dn_f, dl_f = sh.backward()

    say("  oracle frames of the 32 views (threads)")

    def render(R):
        f = oracle.Frame(W, H, True)
        f.draw_mesh(U, m.xyz, m.rgb, R)
        return f

    with ThreadPoolExecutor(8) as ex:
        frames = list(ex.map(render, Rs))
    for f in (0, 9, 18, 31):
        cells, z, win = b.read_frame(f, want_winners=True)
        assert_frame_equal(cells, z, win, frames[f], W)

    say("  FP64 adjoint over the 32 views")
    Ms = [oracle.mat4_mul(U, R) for R in Rs]

    def adjoint(f):
        return oracle.shade_f64(frames[f], Ms[f], m.xyz, m.faces, vc, vn, L, target[f * nz:(f + 1) * nz],
                                want_grad=True)

    with ThreadPoolExecutor(8) as ex:
        parts = list(ex.map(adjoint, range(V)))
    loss_cpu = sum(p[0] for p in parts)
    gc = sum(p[2][0] for p in parts)
    gn = sum(p[2][1] for p in parts)
    gl = sum(p[2][2] for p in parts)
    for tag, loss in (("step", loss_s), ("forward", loss_f)):
        err = abs(loss - loss_cpu) / abs(loss_cpu)
        say(f"  loss ({tag}) {loss:.10g} vs FP64 {loss_cpu:.10g}: relative error {err:.2e}")
        assert err <= REL
    for tag, grads in (("step", (dc_s, dn_s, dl_s)), ("backward", (dc_f, dn_f, dl_f))):
        for name, g, ref in zip(("color", "normal", "light"), grads, (gc, gn, gl)):
            d = g.astype(np.float64) - ref
            err = np.linalg.norm(d) / np.linalg.norm(ref)
            big = np.abs(ref) >= 1e-3 * np.abs(ref).max()
            emax = (np.abs(d[big]) / np.abs(ref[big])).max()
            say(f"  d{name} ({tag}): norm-relative {err:.2e}, element-wise max relative {emax:.2e} "
                f"over {int(big.sum())} entries")
            assert err <= REL, (tag, name, err)

    say("  central differences of the FP64 loss, h = 1e-3")
    vc64, vn64, L64 = vc.astype(np.float64), vn.astype(np.float64), L.astype(np.float64)
    h = 1e-3
    # a vertex only changes the loss of the views that see it
    seen = [set(np.unique(m.faces[fr.win[fr.win >= 0]]).tolist()) for fr in frames]

    def loss64(vc_, vn_, L_, views):
        return sum(oracle.shade_f64(frames[f], Ms[f], m.xyz, m.faces, vc_, vn_, L_,
                                    target[f * nz:(f + 1) * nz])[0] for f in views)

    def fd_tol(grad, views, loss_views):
        # the loss is quadratic in each parameter group, so central
        # differences are exact up to the rounding of the two sequential
        # float64 sums: about sqrt(terms) * eps * loss each, over 2h
        terms = 3 * sum(int((frames[f].win >= 0).sum()) for f in views)
        return 1e-6 * max(1.0, abs(grad)) + 8.0 * np.sqrt(terms) * np.finfo(np.float64).eps * loss_views / (2 * h)

    rng = np.random.default_rng(17)
    visible = np.nonzero(np.array([any(v in s for s in seen) for v in range(m.verts.shape[0])]))[0]
    for v in rng.choice(visible, 3, replace=False):
        views = [f for f in range(V) if v in seen[f]]
        loss_v = sum(parts[f][0] for f in views)
        for ch in range(3):
            for which, base, grad in (("c", vc64, gc), ("n", vn64, gn)):
                p_, q_ = base.copy(), base.copy()
                p_[v, ch] += h
                q_[v, ch] -= h
                if which == "c":
                    fd = (loss64(p_, vn64, L64, views) - loss64(q_, vn64, L64, views)) / (2 * h)
                else:
                    fd = (loss64(vc64, p_, L64, views) - loss64(vc64, q_, L64, views)) / (2 * h)
                assert abs(fd - grad[v, ch]) <= fd_tol(grad[v, ch], views, loss_v), (which, v, ch, fd, grad[v, ch])
    for ch in range(3):
        p_, q_ = L64.copy(), L64.copy()
        p_[ch] += h
        q_[ch] -= h
        fd = (loss64(vc64, vn64, p_, range(V)) - loss64(vc64, vn64, q_, range(V))) / (2 * h)
        say(f"  dlight[{ch}]: FD {fd:.10g} vs adjoint {gl[ch]:.10g}")
        assert abs(fd - gl[ch]) <= fd_tol(gl[ch], range(V), loss_cpu), (ch, fd, gl[ch])
