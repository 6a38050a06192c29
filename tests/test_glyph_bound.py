"""The constant-glyph proof behind the tile raster's resolve (CPU).

For each list entry the raster decides whether the reference's shade
fl(fl(nzu * a) * fl(fl(w0 + w1) + w2)) (rasterizer.rs:76) picks one glyph of
the shader at every pixel of the triangle's candidate box; if so, the
winning slot's cell comes from the entry without re-evaluating the pixel.
The decision bounds every pixel's edge-function sum by the FP32 orient o
plus or minus E = 10u (Sb + Q), Sb each edge's |A| |py - ay| + |B| |px - ax|
at its worst corner of the box, summed, and Q the orient's own terms
(u = 2^-24).  It takes
the shade factor from a hardware reciprocal square root (within 1 ulp), so
it widens E by 2^-19 (|o| + E), counts the step index of the lower end and
compares the upper end with that step's threshold.

Here that decision is restated in float32 numpy (one rounding per
operation, as on the GPU) for triangles whose unit normal z sits within a
few ulps of default_shader's thresholds, and checked by brute force against
the reference's per-pixel shade over every pixel of each box: a box whose
glyph the decision calls constant must show one glyph at every pixel.
Triangles whose glyph does vary across the box must exist in the sample,
so the decision is tested where it matters."""
from __future__ import annotations

import numpy as np
import pytest

f32 = np.float32
THR = np.array([0.20, 0.30, 0.40, 0.50, 0.60, 0.70, 0.80, 0.90, 1.0], np.float32)


def _index(s):
    """Step of default_shader (rasterizer.rs:5-27): first i with s <= thr[i];
    9 (the fallback ' ') when none, NaN included."""
    s = np.asarray(s, np.float32)
    i = (THR[None, :] < s[..., None]).sum(-1)
    return np.where(np.isnan(s), 9, i)


def _triangles(n, rng, W=3840, H=2160):
    """Small front-facing screen-space triangles on planes z = kx x + ky y + c
    whose unit normal z is a threshold times (1 + N(0, 2e-7)) (plus a share
    away from every threshold)."""
    c = np.stack([rng.uniform(60, W / 2 - 60, n), rng.uniform(60, H - 60, n)], 1)
    d = rng.uniform(-3.2, 3.2, (n, 3, 2))
    xy = (c[:, None, :] + d).astype(np.float32)
    e1, e2 = xy[:, 1] - xy[:, 0], xy[:, 2] - xy[:, 0]
    flip = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0] < 0
    xy[flip] = xy[flip][:, [0, 2, 1]]
    near = rng.uniform(0, 1, n) < 0.9
    target = np.where(near, rng.choice(THR, n) * (1 + rng.normal(0, 2e-7, n)), rng.uniform(0.05, 1.0, n))
    k = np.sqrt(np.maximum(1.0 / np.minimum(target, 1.0) ** 2 - 1.0, 0.0))
    ph = rng.uniform(0, 2 * np.pi, n)
    z = (k * np.cos(ph))[:, None] * xy[..., 0] + (k * np.sin(ph))[:, None] * xy[..., 1] + rng.uniform(-50, 50, (n, 1))
    return np.concatenate([xy, z.astype(np.float32)[..., None]], 2).astype(np.float32), W, H, near


def _setup(t, W, H):
    """Per-triangle quantities of setup_tri / unit_normal_z / box_x, float32."""
    x1, y1, z1 = t[:, 0, 0], t[:, 0, 1], t[:, 0, 2]
    x2, y2, z2 = t[:, 1, 0], t[:, 1, 1], t[:, 1, 2]
    x3, y3, z3 = t[:, 2, 0], t[:, 2, 1], t[:, 2, 2]
    A0, B0, A1, B1, A2, B2 = x3 - x2, y3 - y2, x1 - x3, y1 - y3, x2 - x1, y2 - y1
    o = A2 * (y3 - y1) - B2 * (x3 - x1)
    a = f32(1) / o
    e1x, e1y, e1z = x2 - x1, y2 - y1, z2 - z1
    e2x, e2y, e2z = x3 - x1, y3 - y1, z3 - z1
    nx = e1y * e2z - e1z * e2y
    ny = e1z * e2x - e1x * e2z
    nz = e1x * e2y - e1y * e2x
    n2 = f32(0) + ((nx * nx + nz * nz) + (ny * ny + f32(0)))
    na = (nz / np.sqrt(n2)) * a
    minx, maxx = np.minimum(x1, np.minimum(x2, x3)), np.maximum(x1, np.maximum(x2, x3))
    miny, maxy = np.minimum(y1, np.minimum(y2, y3)), np.maximum(y1, np.maximum(y2, y3))
    x0 = np.ceil(np.maximum(minx, f32(1))).astype(np.int64)
    x1d = np.ceil(np.minimum(maxx * f32(2), f32(W - 1))).astype(np.int64)
    y0 = np.ceil(np.maximum(miny, f32(1))).astype(np.int64)
    y1d = np.ceil(np.minimum(maxy, f32(H - 1))).astype(np.int64)
    bx0 = np.maximum(x0, np.ceil(minx - f32(0.03125)).astype(np.int64))
    bx1 = np.minimum(x1d - 1, np.floor(maxx + f32(0.03125)).astype(np.int64))
    return dict(x1=x1, y1=y1, x2=x2, y2=y2, x3=x3, y3=y3, A0=A0, B0=B0, A1=A1, B1=B1, A2=A2, B2=B2, o=o, na=na,
                nz=nz, n2=n2, a=a, bx0=bx0, bx1=bx1, by0=y0, by1=y1d - 1)


def _na_hardware(s, eps):
    """na as the raster takes it for the decision: fl(fl(nz * rsq(n2)) * a),
    with the reciprocal square root off by eps (relative; the hardware's is
    within 1 ulp)."""
    with np.errstate(all="ignore"):
        r = (1.0 / np.sqrt(s["n2"].astype(np.float64)) * (1.0 + eps)).astype(np.float32)
        return (s["nz"] * r) * s["a"]


def _box_terms(s, X0, X1, Y0, Y1):
    """box_glyph's Sb: each edge's |A| |py - ay| + |B| |px - ax| at its own
    worst corner of [X0, X1] x [Y0, Y1], the six products added left to
    right as the device does."""
    p = [np.abs(s["A0"]) * np.maximum(np.abs(Y0 - s["y2"]), np.abs(Y1 - s["y2"])),
         np.abs(s["B0"]) * np.maximum(np.abs(X0 - s["x2"]), np.abs(X1 - s["x2"])),
         np.abs(s["A1"]) * np.maximum(np.abs(Y0 - s["y3"]), np.abs(Y1 - s["y3"])),
         np.abs(s["B1"]) * np.maximum(np.abs(X0 - s["x3"]), np.abs(X1 - s["x3"])),
         np.abs(s["A2"]) * np.maximum(np.abs(Y0 - s["y1"]), np.abs(Y1 - s["y1"])),
         np.abs(s["B2"]) * np.maximum(np.abs(X0 - s["x1"]), np.abs(X1 - s["x1"]))]
    Sb = p[0]
    for q in p[1:]:
        Sb = Sb + q
    return Sb


def _box_glyph_index(s, na_h):
    """The raster's decision (box_glyph in sloth_tiled.hip) from the shade
    factor na_h: the constant step index, or -1."""
    X0, X1 = s["bx0"].astype(np.float32), s["bx1"].astype(np.float32)
    Y0, Y1 = s["by0"].astype(np.float32), s["by1"].astype(np.float32)

    Sb = _box_terms(s, X0, X1, Y0, Y1)
    ey, ex = s["y3"] - s["y1"], s["x3"] - s["x1"]
    o = s["A2"] * ey - s["B2"] * ex
    E = f32(5.9604645e-07) * ((Sb + np.abs(s["A2"]) * np.abs(ey)) + np.abs(s["B2"]) * np.abs(ex))
    E2 = E + f32(2.0 ** -19) * (np.abs(o) + E)
    with np.errstate(all="ignore"):
        s1, s2 = na_h * (o - E2), na_h * (o + E2)
        mag = np.maximum(np.abs(na_h), np.maximum(np.abs(s1), np.abs(s2)))
    lo, hi = np.minimum(s1, s2), np.maximum(s1, s2)
    i = _index(lo)
    ok = np.isfinite(mag) & ((i >= len(THR)) | (hi <= THR[np.minimum(i, len(THR) - 1)]))
    return np.where(ok, i, -1)


@pytest.mark.parametrize("eps", [-2.0 ** -23, 0.0, 2.0 ** -23])
def test_constant_glyph_decision_holds_per_pixel(eps):
    """The decision from a reciprocal square root 1 ulp low, exact, and 1
    ulp high, against the per-pixel shade of the reference's exact na."""
    rng = np.random.default_rng(2024)
    t, W, H, near = _triangles(40000, rng)
    s = _setup(t, W, H)
    claim = _box_glyph_index(s, _na_hardware(s, eps))
    lo_idx = np.full(len(t), 99)
    hi_idx = np.full(len(t), -1)
    bw = (s["bx1"] - s["bx0"] + 1).max()
    bh = (s["by1"] - s["by0"] + 1).max()
    assert bw <= 10 and bh <= 10
    with np.errstate(all="ignore"):
        for dy in range(bh):
            for dx in range(bw):
                px = (s["bx0"] + dx).astype(np.float32)
                py = (s["by0"] + dy).astype(np.float32)
                inbox = (s["bx0"] + dx <= s["bx1"]) & (s["by0"] + dy <= s["by1"])
                # rasterizer.rs:71-76, one rounding per operation
                w0 = s["A0"] * (py - s["y2"]) - s["B0"] * (px - s["x2"])
                w1 = s["A1"] * (py - s["y3"]) - s["B1"] * (px - s["x3"])
                w2 = s["A2"] * (py - s["y1"]) - s["B2"] * (px - s["x1"])
                shade = s["na"] * ((w0 + w1) + w2)
                idx = _index(shade)
                lo_idx = np.where(inbox, np.minimum(lo_idx, idx), lo_idx)
                hi_idx = np.where(inbox, np.maximum(hi_idx, idx), hi_idx)
    claimed = claim >= 0
    nonempty = hi_idx >= 0
    # every claimed box shows exactly its claimed glyph at every pixel
    bad = claimed & nonempty & ((lo_idx != claim) | (hi_idx != claim))
    assert not bad.any(), np.nonzero(bad)[0][:10]
    # the sample holds boxes whose glyph varies (the decision must decline
    # those); away from the thresholds (shades spread over [0.05, 1]) it
    # accepts nearly every box, and close to them some constant ones too
    varies = nonempty & (lo_idx != hi_idx)
    assert varies.sum() >= 100, varies.sum()
    assert not (claimed & varies).any()
    far = nonempty & ~near
    assert (claimed & far).sum() >= 0.99 * far.sum(), ((claimed & far).sum(), far.sum())
    assert (claimed & near & nonempty).sum() >= 1000


def test_box_terms_cover_the_exact_sum():
    """The decision's Sb is at least the exact maximum over the entry's box
    of sum |A| |py - ay| + |B| |px - ax|, which a sum of convex functions
    takes at a corner of the box (a smaller Sb would narrow the shade
    interval below what the rounding allows), and at most three times it
    (each edge's own worst corner is at most the sum's worst)."""
    rng = np.random.default_rng(99)
    t, W, H, _ = _triangles(20000, rng)
    s = _setup(t, W, H)
    ok = (s["bx0"] <= s["bx1"]) & (s["by0"] <= s["by1"])
    X0, X1 = s["bx0"].astype(np.float32), s["bx1"].astype(np.float32)
    Y0, Y1 = s["by0"].astype(np.float32), s["by1"].astype(np.float32)
    Sb = _box_terms(s, X0, X1, Y0, Y1)
    f = lambda k: s[k].astype(np.float64)
    edges = ((f("x2"), f("y2"), f("x3") - f("x2"), f("y3") - f("y2")),
             (f("x3"), f("y3"), f("x1") - f("x3"), f("y1") - f("y3")),
             (f("x1"), f("y1"), f("x2") - f("x1"), f("y2") - f("y1")))
    corner = np.zeros(len(t))
    for cx in (s["bx0"].astype(np.float64), s["bx1"].astype(np.float64)):
        for cy in (s["by0"].astype(np.float64), s["by1"].astype(np.float64)):
            tot = sum(np.abs(A) * np.abs(cy - ay) + np.abs(B) * np.abs(cx - ax) for ax, ay, A, B in edges)
            corner = np.maximum(corner, tot)
    assert ok.sum() > 10000
    ratio = Sb[ok].astype(np.float64) / np.maximum(corner[ok], 1e-300)
    assert (ratio >= 1 - 1e-6).all(), ratio.min()
    assert ratio.max() <= 3 * (1 + 1e-6), ratio.max()
    # some boxes meet the exact sum (every edge's worst corner the same)
    assert (ratio < 1 + 1e-6).any()
