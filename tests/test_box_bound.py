"""The candidate-box proof behind the tiled and global GPU paths (CPU).

The GPU classifies each triangle from an FP32 error bound: no pixel outside
[floor(minx), ceil(maxx)] (class BOX) or outside [ceil(minx - 1/32),
floor(maxx + 1/32)] (class TIGHT) can pass the reference's edge test
(rasterizer.rs:71-75).  Here the classification is restated in float32 numpy
(same operation order, one rounding per operation) and every pixel the
reference's FP32 edge functions cover is found by brute force, for triangles
built to sit near the bound: large coordinates, slivers, sub-pixel sizes.
A covered pixel outside its class's box would be a pixel the GPU never tests.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32
K = f32(8.0) * f32(5.9604645e-08)


def _sat(v):
    v = np.asarray(v, np.float32)
    out = np.where(v > 0, np.minimum(v, f32(2 ** 30)), 0).astype(np.int64)
    return out


def edge_terms_bound(v, X0, X1, Y0, Y1):
    """edge_terms_bound (sloth_device.h) in float32: the edges' summed
    |A| |py - ay| + |B| |px - ax| over [X0, X1] x [Y0, Y1] is at most
    (sum |A|) DY + (sum |B|) DX, DX = max(|X1 - minx|, |X0 - maxx|)."""
    (x1, y1), (x2, y2), (x3, y3) = v
    minx, maxx = min(x1, x2, x3), max(x1, x2, x3)
    miny, maxy = min(y1, y2, y3), max(y1, y2, y3)
    A0, B0, A1, B1, A2, B2 = x3 - x2, y3 - y2, x1 - x3, y1 - y3, x2 - x1, y2 - y1
    DX = max(abs(X1 - minx), abs(X0 - maxx))
    DY = max(abs(Y1 - miny), abs(Y0 - maxy))
    return ((abs(A0) + abs(A1)) + abs(A2)) * DY + ((abs(B0) + abs(B1)) + abs(B2)) * DX


def classify(t, W, H):
    """(class, x0, x1d, y0, y1d) with class in EMPTY/BOX/TIGHT/SPANS/FULL."""
    (x1, y1), (x2, y2), (x3, y3) = [(f32(p[0]), f32(p[1])) for p in t]
    minx, maxx = min(x1, x2, x3), max(x1, x2, x3)
    miny, maxy = min(y1, y2, y3), max(y1, y2, y3)
    x0 = int(_sat(np.ceil(max(minx, f32(1)))))
    y0 = int(_sat(np.ceil(max(miny, f32(1)))))
    x1d = int(_sat(np.ceil(min(maxx * f32(2), f32(W - 1)))))
    y1d = int(_sat(np.ceil(min(maxy, f32(H - 1)))))
    if x0 >= x1d or y0 >= y1d:
        return "EMPTY", x0, x1d, y0, y1d
    lim = f32(4194304.0)
    if not all(abs(c) < lim for c in (x1, y1, x2, y2, x3, y3)):
        return "FULL", x0, x1d, y0, y1d
    A0, B0 = x3 - x2, y3 - y2
    A1, B1 = x1 - x3, y1 - y3
    A2, B2 = x2 - x1, y2 - y1
    X0, X1, Y0, Y1 = f32(x0), f32(x1d - 1), f32(y0), f32(y1d - 1)
    bound = K * edge_terms_bound(((x1, y1), (x2, y2), (x3, y3)), X0, X1, Y0, Y1)
    ey, ex = y3 - y1, x3 - x1
    o = A2 * ey - B2 * ex
    eo = K * (abs(A2) * abs(ey) + abs(B2) * abs(ex))
    if o + eo < -bound:
        return "EMPTY", x0, x1d, y0, y1d
    lo = o - eo
    if lo > 0:
        bw = bound * (maxx - minx)
        if bw < f32(0.0078125) * lo:
            return "TIGHT", x0, x1d, y0, y1d
        if bw < f32(0.5) * lo:
            return "BOX", x0, x1d, y0, y1d
    return "SPANS", x0, x1d, y0, y1d


def covered(t, x0, x1d, y0, y1d, pad=4):
    """Pixels of the domain near the triangle whose FP32 edge functions are
    all >= 0 (rasterizer.rs:71-75), as (xs, ys)."""
    (x1, y1), (x2, y2), (x3, y3) = [(f32(p[0]), f32(p[1])) for p in t]
    minx, maxx = min(x1, x2, x3), max(x1, x2, x3)
    xa = max(x0, int(np.floor(minx)) - pad)
    xb = min(x1d - 1, int(np.ceil(maxx)) + pad)
    if xa > xb:
        return np.zeros(0, int), np.zeros(0, int)
    px, py = np.meshgrid(np.arange(xa, xb + 1), np.arange(y0, y1d))
    px = px.astype(np.float32)
    py = py.astype(np.float32)
    w0 = (x3 - x2) * (py - y2) - (y3 - y2) * (px - x2)
    w1 = (x1 - x3) * (py - y3) - (y1 - y3) * (px - x3)
    w2 = (x2 - x1) * (py - y1) - (y2 - y1) * (px - x1)
    m = (w0 >= 0) & (w1 >= 0) & (w2 >= 0)
    return px[m].astype(np.int64), py[m].astype(np.int64)


def _triangles(rng, n):
    out = []
    for i in range(n):
        kind = i % 4
        base = rng.choice([3.0, 40.0, 1000.0, 70000.0, 1.5e6])
        c = np.array([base + rng.uniform(0, 8), rng.uniform(2, 60)])
        if kind == 0:      # small, arbitrary
            t = c + rng.normal(0, rng.choice([0.3, 1.0, 3.0]), (3, 2))
        elif kind == 1:    # sliver
            d = rng.normal(0, 6, 2)
            perp = np.array([-d[1], d[0]]) * rng.choice([1e-2, 1e-3, 1e-4])
            t = np.stack([c, c + d, c + 0.5 * d + perp])
        elif kind == 2:    # vertex x close to an integer
            t = c + rng.normal(0, 2, (3, 2))
            t[:, 0] = np.round(t[:, 0]) + rng.choice([-1, 1], 3) * rng.choice([1e-3, 1e-5, 0.0], 3)
        else:              # medium
            t = c + rng.normal(0, 8, (3, 2))
        t = t.astype(np.float32)
        o = (t[1, 0] - t[0, 0]) * (t[2, 1] - t[0, 1]) - (t[1, 1] - t[0, 1]) * (t[2, 0] - t[0, 0])
        if o < 0:
            t = t[[0, 2, 1]]
        out.append(t)
    return out


def test_covered_pixels_stay_in_the_class_box():
    rng = np.random.default_rng(5)
    W, H = 1 << 23, 80
    seen = {}
    for t in _triangles(rng, 3000):
        cls, x0, x1d, y0, y1d = classify(t, W, H)
        seen[cls] = seen.get(cls, 0) + 1
        if cls not in ("TIGHT", "BOX", "EMPTY"):
            continue
        xs, ys = covered(t, x0, x1d, y0, y1d)
        if cls == "EMPTY":
            assert len(xs) == 0, f"EMPTY triangle covers pixels: {t.tolist()}"
            continue
        minx, maxx = t[:, 0].min(), t[:, 0].max()
        if cls == "TIGHT":
            lo = int(np.ceil(f32(minx - f32(0.03125))))
            hi = int(np.floor(f32(maxx + f32(0.03125))))
        else:
            lo, hi = int(np.floor(minx)), int(np.ceil(maxx))
        if len(xs):
            assert xs.min() >= lo and xs.max() <= hi, (cls, t.tolist(), xs.min(), xs.max(), lo, hi)
    # the sample reaches every class the proof decides
    assert seen.get("TIGHT", 0) > 500 and seen.get("BOX", 0) > 0 and seen.get("SPANS", 0) > 0, seen


def test_tight_box_is_tighter():
    """The tight box drops the column past maxx that the loose box keeps."""
    t = np.array([[10.2, 5.1], [13.6, 5.3], [11.0, 8.9]], np.float32)
    cls, *_ = classify(t, 100, 40)
    assert cls == "TIGHT"
    assert int(np.floor(f32(t[:, 0].max() + f32(0.03125)))) == 13
    assert int(np.ceil(t[:, 0].max())) == 14


def test_error_model_of_the_bound():
    """The proof's ingredients, checked per pixel against exact values (float64
    holds the products of float32 differences exactly): each FP32 edge function
    is within 4.0001 u (|A| |py - ay| + |B| |px - ax|) of the exact one, the
    sum of those terms over the domain stays below `bound`, and the exact
    orient is at least `lo`."""
    rng = np.random.default_rng(17)
    u = 2.0 ** -24
    worst = 0.0
    tightest = np.inf
    checked = 0
    for t in _triangles(rng, 1500):
        cls, x0, x1d, y0, y1d = classify(t, 1 << 23, 80)
        if cls not in ("TIGHT", "BOX"):
            continue
        xb = min(x1d, x0 + 48)
        px, py = np.meshgrid(np.arange(x0, xb), np.arange(y0, y1d))
        pxf, pyf = px.astype(np.float32), py.astype(np.float32)
        v = [(f32(p[0]), f32(p[1])) for p in t]
        total = np.zeros(px.shape)
        for (ax, ay), (bx, by) in ((v[1], v[2]), (v[2], v[0]), (v[0], v[1])):
            # edge from a to b: w = (bx-ax)(py-ay) - (by-ay)(px-ax); the exact
            # value uses the exact vertex differences
            w_fp = (bx - ax) * (pyf - ay) - (by - ay) * (pxf - ax)
            Ae, Be = float(bx) - float(ax), float(by) - float(ay)
            w_ex = Ae * (py - float(ay)) - Be * (px - float(ax))
            term = abs(Ae) * np.abs(py - float(ay)) + abs(Be) * np.abs(px - float(ax))
            err = np.abs(w_fp.astype(np.float64) - w_ex)
            ok = err <= 4.0001 * u * term + 1e-300
            assert ok.all(), (t.tolist(), float((err - 4.0001 * u * term).max()))
            worst = max(worst, float((err / np.maximum(term, 1e-300)).max()) / u)
            total += term
        # the class's bound covers the worst pixel's summed terms
        (x1, y1), (x2, y2), (x3, y3) = v
        X0, X1, Y0, Y1 = f32(x0), f32(x1d - 1), f32(y0), f32(y1d - 1)
        # the bound as the device computes it (edge_terms_bound), float32
        S = edge_terms_bound(v, X0, X1, Y0, Y1)
        assert 4.0001 * u * total.max() <= float(K * S)
        # ... and against the summed terms' exact maximum over the whole
        # domain, which a sum of convex functions takes at a corner
        corner = max(sum(abs(float(bx) - float(ax)) * abs(cy - float(ay)) + abs(float(by) - float(ay)) * abs(cx - float(ax))
                         for (ax, ay), (bx, by) in ((v[1], v[2]), (v[2], v[0]), (v[0], v[1])))
                     for cx in (float(x0), float(x1d - 1)) for cy in (float(y0), float(y1d - 1)))
        assert float(S) >= corner * (1 - 1e-6), (t.tolist(), float(S), corner)
        tightest = min(tightest, float(S) / corner) if corner > 0 else tightest
        o_ex = (float(x2) - float(x1)) * (float(y3) - float(y1)) - (float(y2) - float(y1)) * (float(x3) - float(x1))
        A2, B2, ey, ex = x2 - x1, y2 - y1, y3 - y1, x3 - x1
        lo = (A2 * ey - B2 * ex) - K * (abs(A2) * abs(ey) + abs(B2) * abs(ex))
        assert o_ex >= float(lo)
        checked += 1
    assert checked > 500
    assert worst <= 4.0001
    # the bound is not loose for nothing: some domains meet it within 1%
    assert tightest < 1.01, tightest
