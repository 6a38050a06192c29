"""Pins of the CPU oracle (CPU only).

The Rust reference cannot run here (no Rust toolchain), so the oracle's
restatement is checked against what fixes it independently:

* the reference's own unit tests, geometry.rs:191-235 (golden vectors);
* default_shader (rasterizer.rs:5-27): the table form against the literal
  if-chain for every one of the 2^32 binary32 bit patterns;
* exact arithmetic: triangles with small integer coordinates, where every
  binary32 operation of orient() is exact, so coverage must equal the
  inclusive half-plane test done with Python integers over the reference's
  pixel domain (rasterizer.rs:58-75), including the doubled x bound, the
  row wrap of id = y*W + 2x and the newline cells;
* closed forms: z at a pixel on vertex v1 is v1.z, identical triangles tie
  to the first drawn, the screen transform of context.rs:104-133;
* matrix products against float64 (a transposed or reordered operand fails).
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle

I4 = np.eye(4, dtype=np.float32).reshape(16)


def col_major(rows):
    return np.asarray(rows, np.float32).T.reshape(16)


# --- geometry.rs:191-235 ----------------------------------------------------

DEFAULT_TRI = np.array([[1.0, -1.0, -1.0, 1.0], [-1.0, -1.0, 1.0, 1.0], [1.0, 1.0, -1.0, 1.0]], np.float32)


def test_reference_test_aabb():
    mn, mx = oracle.tri_aabb(DEFAULT_TRI)
    assert mn.tolist() == [-1.0, -1.0, -1.0, 1.0]
    assert mx.tolist() == [1.0, 1.0, 1.0, 1.0]


def test_reference_test_transform():
    T = col_major([[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1], [0, 0, 0, 1]])   # new_translation(1,1,1)
    t = oracle.tri_mul(T, DEFAULT_TRI)
    mn, mx = oracle.tri_aabb(t)
    assert mn.tolist() == [0.0, 0.0, 0.0, 1.0]
    assert mx.tolist() == [2.0, 2.0, 2.0, 1.0]


def test_reference_test_normal():
    tri = np.array([[-1, 1, 0, 1], [0, 1, 1, 1], [1, 1, 0, 1]], np.float32)
    assert oracle.tri_normal(tri).tolist() == [0.0, 1.0, 0.0, 0.0]


# --- default_shader, exhaustive -------------------------------------------

def test_default_shader_exhaustive():
    """Every binary32 bit pattern: step table == the literal if-chain."""
    chunks = [(i, i + 4096) for i in range(0, 65536, 4096)]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        bad = sum(ex.map(lambda c: oracle.shader_exhaustive_mismatch(*c), chunks))
    assert bad == 0


def test_default_shader_spot_values():
    f = oracle.default_shader
    assert f(-1.0) == "." and f(0.2) == "." and f(np.nextafter(np.float32(0.2), np.float32(1))) == ":"
    assert f(0.5) == "=" and f(1.0) == "@" and f(np.nextafter(np.float32(1.0), np.float32(2))) == " "
    assert f(float("nan")) == " " and f(float("inf")) == " " and f(float("-inf")) == "."


# --- matrix products ------------------------------------------------------

def test_mat4_mul_is_a_product_in_the_right_order():
    rng = np.random.default_rng(0)
    for _ in range(50):
        A = rng.integers(-8, 9, (4, 4)).astype(np.float32)
        B = rng.integers(-8, 9, (4, 4)).astype(np.float32)
        got = oracle.mat4_mul(A.T.reshape(16), B.T.reshape(16)).reshape(4, 4).T
        assert np.array_equal(got, (A.astype(np.float64) @ B.astype(np.float64)).astype(np.float32))
    A = rng.normal(size=(4, 4)).astype(np.float32)
    B = rng.normal(size=(4, 4)).astype(np.float32)
    got = oracle.mat4_mul(A.T.reshape(16), B.T.reshape(16)).reshape(4, 4).T
    assert np.allclose(got, A.astype(np.float64) @ B.astype(np.float64), rtol=1e-6, atol=1e-6)


def test_tri_mul_is_matrix_times_vector():
    rng = np.random.default_rng(1)
    M = rng.integers(-5, 6, (4, 4)).astype(np.float32)
    tri = np.concatenate([rng.integers(-9, 10, (3, 3)), np.ones((3, 1))], 1).astype(np.float32)
    got = oracle.tri_mul(M.T.reshape(16), tri)
    assert np.array_equal(got, (M.astype(np.float64) @ tri.T.astype(np.float64)).T.astype(np.float32))


def test_rotation_euler():
    for roll, pitch, yaw in [(0.0, 0.3, 0.0), (0.7, -1.2, 2.5), (0.0, math.pi, 0.0)]:
        R = oracle.rotation_euler(roll, pitch, yaw).reshape(4, 4).T.astype(np.float64)
        cr, sr, cp, sp, cy, sy = (math.cos(roll), math.sin(roll), math.cos(pitch), math.sin(pitch),
                                  math.cos(yaw), math.sin(yaw))
        Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
        Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
        Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
        assert np.allclose(R[:3, :3], Rz @ Ry @ Rx, atol=2e-6)
        assert R[3].tolist() == [0, 0, 0, 1] and R[:3, 3].tolist() == [0, 0, 0]


def test_update_closed_form():
    U = oracle.update(160, 48, [[0.5, 2.0, 1.0], [3.0, 0.25, 0.0]]).reshape(4, 4).T
    s = np.float32(np.float32(min(48.0, 80.0)) / np.float32(3.0)) / np.float32(2.0)
    assert U[0, 0] == s and U[1, 1] == -s and U[2, 2] == s
    assert U[0, 3] == 40.0 and U[1, 3] == 24.0 and U[3, 3] == 1.0


# --- exact-arithmetic brute force -----------------------------------------

def exact_reference_frame(W, H, tris, image=True):
    """Sequential draw with Python integers for small integer coordinates:
    orient() is exact there, so coverage is the inclusive half-plane test.
    Returns the winning triangle per z slot (z all equal -> first drawn wins)
    and the newline cells."""
    win = np.full(W * H, -1, np.int64)
    zb = np.full(W * H, np.inf)
    nl_rows = {}
    for ti, (v1, v2, v3) in enumerate(tris):
        xs = [v1[0], v2[0], v3[0]]
        ys = [v1[1], v2[1], v3[1]]
        x0 = max(math.ceil(max(min(xs), 1)), 0)
        y0 = max(math.ceil(max(min(ys), 1)), 0)
        x1 = max(math.ceil(min(2 * max(xs), W - 1)), 0)
        y1 = max(math.ceil(min(max(ys), H - 1)), 0)

        def orient(a, b, c):
            return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])

        for y in range(y0, y1):
            for x in range(x0, x1):
                p = (x, y)
                if orient(v2, v3, p) >= 0 and orient(v3, v1, p) >= 0 and orient(v1, v2, p) >= 0:
                    z = v1[2]   # all vertices share z in these cases
                    i = y * W + 2 * x
                    if z < zb[i]:
                        zb[i] = z
                        win[i] = ti
            nl_rows[y] = ti
    return win, nl_rows


@pytest.mark.parametrize("seed", range(6))
def test_coverage_matches_exact_arithmetic(seed):
    rng = np.random.default_rng(seed)
    W, H = (24, 9) if seed % 2 else (25, 10)
    tris = []
    for _ in range(60):
        v = rng.integers(-3, max(W, H) + 3, (3, 2))
        tris.append([(int(v[k, 0]), int(v[k, 1]), 7) for k in range(3)])
    xyz = np.array(tris, np.float32)
    rgb = np.zeros((len(tris), 3), np.uint8)
    fr = oracle.render(W, H, I4, xyz, rgb, I4)
    want, nl_rows = exact_reference_frame(W, H, tris)
    assert np.array_equal(fr.win, want)
    for y in range(H):
        c = fr.cells[y * W + 1]
        # cell y*W+1 is written by the newline of row y (after every pixel of
        # its triangle's row y) and by the winners of z slots y*W and y*W+1,
        # which can only be wrapped pixels of row y-1: the newline is the last
        # write iff its triangle is not earlier than those winners
        ws = [int(want[i]) for i in (y * W, y * W + 1) if i < W * H and want[i] >= 0]
        expect_nl = y in nl_rows and (not ws or nl_rows[y] >= max(ws))
        assert (c == oracle.CELL_NEWLINE) == expect_nl, (y, ws, nl_rows.get(y))


def test_depth_at_vertex_is_v1z():
    W, H = 40, 20
    xyz = np.array([[[10, 5, 0.3125], [30, 5, -2.75], [10, 15, 4.5]]], np.float32)
    fr = oracle.render(W, H, I4, xyz, np.zeros((1, 3), np.uint8), I4)
    assert fr.z[5 * W + 20] == np.float32(0.3125)


def test_ties_go_to_the_first_drawn():
    W, H = 40, 20
    t = [[2, 2, 1], [15, 2, 1], [2, 15, 1]]
    xyz = np.array([t, t], np.float32)
    rgb = np.array([[10, 20, 30], [40, 50, 60]], np.uint8)
    fr = oracle.render(W, H, I4, xyz, rgb, I4)
    covered = fr.win >= 0
    assert covered.any() and (fr.win[covered] == 0).all()
    cells = fr.cells[fr.cells != oracle.CELL_BLANK]
    assert ((cells >> 8) & 0xFF).tolist().count(40) == 0


def test_empty_x_range_still_writes_newlines():
    """rasterizer.rs:89-91 runs per row of the y domain, covered or not."""
    W, H = 20, 10
    xyz = np.array([[[-9, 2, 0], [-8, 2, 0], [-9, 7, 0]]], np.float32)   # x domain empty
    fr = oracle.render(W, H, I4, xyz, np.zeros((1, 3), np.uint8), I4)
    nl = [y for y in range(H) if fr.cells[y * W + 1] == oracle.CELL_NEWLINE]
    assert nl == [2, 3, 4, 5, 6]
    assert (fr.win == -1).all()


def test_clear_state():
    f = oracle.Frame(7, 3, True)
    assert f.cells.size == 7 * 3 + 3 and (f.cells == oracle.CELL_BLANK).all()
    assert (f.z == oracle.F32_MAX).all()
    g = oracle.Frame(7, 3, False)
    assert g.cells.size == 21


# --- data.js (structure only) ---------------------------------------------

def test_datajs_structure():
    """src-webify/data.js is an export of skull.obj at 100x100 made by an
    older build of the renderer (its <span> format predates context.rs:79-87
    and its frame 0 covers pixels of row 25 that the current rasterizer
    cannot reach: the mesh's topmost transformed vertex is at y = 25.0 with
    x = 26.35).  It still fixes the frame layout, which the oracle must share:
    newline cells at y*W + 1, pixel cells doubled, the material colours, and
    nearly all coverage."""
    from sloth_amd import meshgen

    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "skull_datajs.npz"))
    xyz, rgb, counts = d["xyz"], d["rgb"], d["mesh_counts"]
    W, H = int(d["width"]), int(d["height"])
    offs = np.concatenate([[0], np.cumsum(counts)])
    U = oracle.update(W, H, [oracle.mesh_bbox(xyz[offs[i]:offs[i + 1]])[1] for i in range(len(counts))])
    pitches = meshgen.turntable_pitches(100)
    palette = d["palette"]
    agree = total = 0
    for k in range(0, 100, 9):
        R = oracle.rotation_euler(0.0, pitches[k], 0.0)
        f = oracle.Frame(W, H, True)
        for i in range(len(counts)):
            f.draw_mesh(U, xyz[offs[i]:offs[i + 1]], rgb[offs[i]:offs[i + 1]], R)
        g, e = f.glyphs(), d["glyphs"][k]
        assert np.array_equal(np.nonzero(g == 10)[0], np.nonzero(e == 10)[0])
        cm = (g != 32) & (g != 10)
        ce = (e != 32) & (e != 10)
        agree += int((cm == ce).sum())
        total += cm.size
        cols = {tuple(int(v) for v in palette[i]) for i in np.unique(d["color_index"][k])}
        mine = {((c >> 8) & 255, (c >> 16) & 255, (c >> 24) & 255) for c in np.unique(f.cells)}
        assert mine <= cols | {(0, 0, 0)}
    assert agree / total > 0.98


# --- FP64 shading model (the added differentiable form) --------------------

def _shading_case(seed=2):
    from sloth_amd import meshgen

    m = meshgen.uv_sphere(16, 8, 1.0, seed=seed)
    W, H = 60, 30
    U = oracle.update(W, H, [m.bbox_max()])
    R = oracle.rotation_euler(0.2, 2.2, 0.1)
    fr = oracle.render(W, H, U, m.xyz, m.rgb, R)
    rng = np.random.default_rng(seed)
    vc = rng.uniform(0, 1, (m.verts.shape[0], 3))
    vn = m.verts / np.linalg.norm(m.verts, axis=1, keepdims=True)
    L = np.array([0.3, 0.5, -0.8])
    target = rng.uniform(0, 1, (W * H, 3)).astype(np.float32)
    return m, oracle.mat4_mul(U, R), fr, vc.astype(np.float64), vn.astype(np.float64), L, target


def test_shade_adjoint_matches_central_differences():
    """The loss is quadratic in each parameter group, so central differences
    are exact up to float64 rounding."""
    m, M, fr, vc, vn, L, target = _shading_case()
    loss, _, (gc, gn, gl) = oracle.shade_f64(fr, M, m.xyz, m.faces, vc, vn, L, target, want_grad=True)
    assert loss > 0
    h = 1e-3
    rng = np.random.default_rng(0)
    used = np.unique(m.faces[fr.win[fr.win >= 0]])
    for v in rng.choice(used, 6, replace=False):
        for ch in range(3):
            for base, grad, slot in ((vc, gc, 0), (vn, gn, 1)):
                p, q = base.copy(), base.copy()
                p[v, ch] += h
                q[v, ch] -= h
                args_p = [vc, vn, L]
                args_q = [vc, vn, L]
                args_p[slot], args_q[slot] = p, q
                fd = (oracle.shade_f64(fr, M, m.xyz, m.faces, *args_p, target)[0]
                      - oracle.shade_f64(fr, M, m.xyz, m.faces, *args_q, target)[0]) / (2 * h)
                assert abs(fd - grad[v, ch]) <= 1e-7 * max(1.0, abs(fd))
    for ch in range(3):
        p, q = L.copy(), L.copy()
        p[ch] += h
        q[ch] -= h
        fd = (oracle.shade_f64(fr, M, m.xyz, m.faces, vc, vn, p, target)[0]
              - oracle.shade_f64(fr, M, m.xyz, m.faces, vc, vn, q, target)[0]) / (2 * h)
        assert abs(fd - gl[ch]) <= 1e-7 * max(1.0, abs(fd))


def test_shade_reduces_to_flat_shade():
    """Per-corner normals = the face's unit normal, L = (0,0,1), colours = 1:
    the intensity equals the reference's nz_unit * a * (w0 + w1 + w2)."""
    from sloth_amd import meshgen

    base = meshgen.uv_sphere(16, 8, seed=3)
    n = base.n
    faces = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    e1 = base.xyz[:, 1].astype(np.float64) - base.xyz[:, 0]
    e2 = base.xyz[:, 2].astype(np.float64) - base.xyz[:, 0]
    fn = np.cross(e1, e2)
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    W, H = 60, 30
    U = oracle.update(W, H, [base.bbox_max()])
    R = oracle.rotation_euler(0.1, 1.9, 0.0)
    fr = oracle.render(W, H, U, base.xyz, base.rgb, R)
    M = oracle.mat4_mul(U, R)
    _, out, _ = oracle.shade_f64(fr, M, base.xyz, faces, np.ones((3 * n, 3)), np.repeat(fn, 3, 0),
                                 np.array([0.0, 0.0, 1.0]), want_out=True)
    for idx in np.nonzero(fr.win >= 0)[0]:
        T = fr.win[idx]
        t = oracle.tri_mul(M, np.concatenate([base.xyz[T], np.ones((3, 1), np.float32)], 1)).astype(np.float64)
        x, y = fr.winxy[2 * idx], fr.winxy[2 * idx + 1]
        o = (t[1, 0] - t[0, 0]) * (t[2, 1] - t[0, 1]) - (t[1, 1] - t[0, 1]) * (t[2, 0] - t[0, 0])
        w = [(t[2, 0] - t[1, 0]) * (y - t[1, 1]) - (t[2, 1] - t[1, 1]) * (x - t[1, 0]),
             (t[0, 0] - t[2, 0]) * (y - t[2, 1]) - (t[0, 1] - t[2, 1]) * (x - t[2, 0]),
             (t[1, 0] - t[0, 0]) * (y - t[0, 1]) - (t[1, 1] - t[0, 1]) * (x - t[0, 0])]
        n_s = np.cross(t[1, :3] - t[0, :3], t[2, :3] - t[0, :3])
        flat = n_s[2] / np.linalg.norm(n_s) * sum(w) / o
        assert abs(out[idx, 0] - flat) < 1e-6


def test_encode_known_answers():
    """Context::flush bytes for three hand-made cells, written out from the
    format strings of context.rs:58-61 (println! of the chars), :62-78
    (crossterm 0.18.0 PrintStyledContent: background, foreground, char,
    reset) and :79-87 (the -j spans), and the terminal-mode MoveTo(0, 0)
    (:53-55).  Glyph bytes are chars below U+0100, so 0xE9 is 'é' (2 UTF-8
    bytes)."""
    cells = np.array([ord("@") | (255 << 8) | (0 << 16) | (7 << 24),
                      0x0A,
                      0xE9 | (1 << 8) | (22 << 16) | (133 << 24)], np.uint32)
    assert oracle.encode(cells, 0) == b"@\n\xc3\xa9\n"
    assert oracle.encode(cells, 0, image=False) == b"\x1b[1;1H@\n\xc3\xa9\n"
    assert oracle.encode(cells, 2) == (b'<span style="color:rgb(255,0,7)">@'
                                       b'<span style="color:rgb(0,0,0)">\n'
                                       b'<span style="color:rgb(1,22,133)">\xc3\xa9')
    bg = b"\x1b[48;2;25;25;25m"
    assert oracle.encode(cells, 1) == (bg + b"\x1b[38;2;255;0;7m@\x1b[0m" +
                                       bg + b"\x1b[38;2;0;0;0m\n\x1b[0m" +
                                       bg + b"\x1b[38;2;1;22;133m\xc3\xa9\x1b[0m")
    assert oracle.encode(cells[:0], 1, image=False) == b"\x1b[1;1H"
