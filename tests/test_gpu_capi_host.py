"""The C-ABI from a C++ host (GPU).

`tests/capi_host` (built by `build()` from `tests/capi_host.cpp`) links
libsloth_hip.so and uses only `include/sloth_hip.h`, as a foreign binding of
the reference would: it loads a model file or takes a triangle soup, runs
Context::update from the meshes' bounding boxes, clear, draw_mesh per mesh,
reads the frame and encodes it.  Its output must equal the oracle's frame and
flush bytes, and its utransform the restatement of Context::update.
"""
from __future__ import annotations

import math
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle
from _helpers import assert_frame_equal, oracle_render, say

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
HOST = os.path.join(HERE, "capi_host")


def _run(tmp_path, spec: bytes, W, H, image):
    sp, op = tmp_path / "spec.bin", tmp_path / "out.bin"
    sp.write_bytes(spec)
    if not os.path.exists(HOST):
        raise AssertionError("tests/capi_host is not built (run __graft_entry__.build())")
    r = subprocess.run([HOST, str(sp), str(op)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    b = op.read_bytes()
    (st,) = struct.unpack_from("<I", b, 0)
    assert st == 0, st
    (nc,) = struct.unpack_from("<Q", b, 4)
    o = 12
    cells = np.frombuffer(b, np.uint32, nc, o)
    o += 4 * nc
    z = np.frombuffer(b, np.float32, W * H, o)
    o += 4 * W * H
    win = np.frombuffer(b, np.int32, W * H, o)
    o += 4 * W * H
    U = np.frombuffer(b, np.float32, 16, o)
    o += 64
    (nb,) = struct.unpack_from("<Q", b, o)
    o += 8
    enc = b[o:o + nb]
    assert nc == W * H + (H if image else 0)
    return cells, z, win, U, enc


def _header(kind, W, H, image, mode, T):
    return struct.pack("<5I", kind, W, H, int(image), mode) + np.asarray(T, np.float32).reshape(16).tobytes()


def test_soup_frame_and_flush(tmp_path):
    from sloth_amd import meshgen

    m, W, H, U, R = meshgen.workload_sphere_160x48()
    say(f"case C++ host: soup {m.n} tris {W}x{H}, ANSI flush")
    spec = (_header(0, W, H, True, 1, R)
            + np.asarray(U, np.float32).tobytes() + struct.pack("<I", m.n)
            + np.ascontiguousarray(m.xyz, np.float32).tobytes() + np.ascontiguousarray(m.rgb, np.uint8).tobytes())
    cells, z, win, U_out, enc = _run(tmp_path, spec, W, H, True)
    ref = oracle.render(W, H, U, m.xyz, m.rgb, R)
    assert_frame_equal(cells, z, win, ref, W)
    assert np.array_equal(U_out.view(np.uint32), np.asarray(U, np.float32).view(np.uint32))
    assert enc == oracle.encode(ref.cells, 1, image=True)


def _sphere_obj(path, n_lon=20, n_lat=10):
    from sloth_amd import meshgen

    m = meshgen.uv_sphere(n_lon, n_lat, 1.0, seed=4)
    with open(path, "w") as f:
        for v in m.verts:
            f.write(f"v {v[0]:.6f} {v[1]:.6f} {v[2]:.6f}\n")
        for a, b, c in m.faces:
            f.write(f"f {a + 1} {b + 1} {c + 1}\n")


@pytest.mark.parametrize("name,W,H,image,mode", [("cube.obj", 100, 40, True, 2), ("cube.stl", 81, 25, False, 1),
                                                  ("sphere.obj", 121, 50, True, 0)])
def test_model_file_frame_and_flush(tmp_path, sloth, name, W, H, image, mode):
    path = os.path.join(HERE, "golden", "models", name)
    if name == "sphere.obj":
        path = str(tmp_path / name)
        _sphere_obj(path)
    from sloth_amd import meshgen

    T = meshgen.rotation_euler(0.4, math.pi + 0.7, 0.1)
    say(f"case C++ host: model {name} {W}x{H} image={image} flush mode {mode}")
    pb = path.encode()
    spec = _header(1, W, H, image, mode, T) + struct.pack("<I", len(pb)) + pb
    cells, z, win, U_out, enc = _run(tmp_path, spec, W, H, image)

    mf = sloth.ModelFile(path)
    arrays = [mf.mesh_arrays(i) for i in range(mf.n_meshes)]
    U = meshgen.screen_transform(W, H, [a[3] for a in arrays])
    assert np.array_equal(U_out.view(np.uint32), U.view(np.uint32)), (U_out, U)
    ref = oracle_render(W, H, U, [(pos, rgb, T, None) for pos, rgb, _, _ in arrays], image=image)
    assert_frame_equal(cells, z, win, ref, W)
    assert enc == oracle.encode(ref.cells, mode, image=image)


def test_shading_pair_from_host(tmp_path):
    """The differentiable-shading forward/backward pair and the one-pass step
    driven through the bare header (a binding's calls): 3 views of a
    48 x 24 UV sphere at 128x96, target from colours + N(0, 0.1^2) and the
    light rotated 10 degrees.  Visibility bit for bit, loss and gradients
    against the FP64 CPU adjoint to 1e-5 norm-relative."""
    from sloth_amd import meshgen

    wl = meshgen.workload_shading_views(n_views=3, size=128, n_lon=48, n_lat=24, seed=13)
    m, W, H, U, Rs = wl["mesh"], wl["W"], 96, wl["U"], wl["Rs"]
    U = meshgen.screen_transform(W, H, [m.bbox_max()])
    V, nv = len(Rs), m.verts.shape[0]
    say(f"case C++ host: shading pair, {V} views {W}x{H}, {m.n} tris")
    f32 = lambda a: np.ascontiguousarray(a, np.float32).tobytes()   # noqa: E731
    spec = (_header(2, W, H, True, 0, np.eye(4, dtype=np.float32)) + f32(U) + struct.pack("<I", V)
            + b"".join(f32(R) for R in Rs) + struct.pack("<I", m.n) + f32(m.xyz)
            + np.ascontiguousarray(m.rgb, np.uint8).tobytes() + struct.pack("<I", nv)
            + np.ascontiguousarray(m.faces, np.uint32).tobytes() + f32(wl["vcolor"]) + f32(wl["vnormal"])
            + f32(wl["light"]) + f32(wl["vcolor_target"]) + f32(wl["light_target"]))
    sp, op = tmp_path / "spec.bin", tmp_path / "out.bin"
    sp.write_bytes(spec)
    r = subprocess.run([HOST, str(sp), str(op)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    b = op.read_bytes()
    (st,) = struct.unpack_from("<I", b, 0)
    assert st == 0, st
    (ns,) = struct.unpack_from("<Q", b, 4)
    assert ns == V * W * H
    o = 12
    win = np.frombuffer(b, np.int32, ns, o)
    o += 4 * ns
    target = np.frombuffer(b, np.float32, 3 * ns, o).reshape(ns, 3)
    o += 12 * ns
    got = []
    for _ in range(2):
        (loss,) = struct.unpack_from("<d", b, o)
        o += 8
        dc = np.frombuffer(b, np.float32, 3 * nv, o).reshape(nv, 3)
        o += 12 * nv
        dn = np.frombuffer(b, np.float32, 3 * nv, o).reshape(nv, 3)
        o += 12 * nv
        dl = np.frombuffer(b, np.float32, 3, o)
        o += 12
        got.append((loss, dc, dn, dl))
    assert o == len(b)

    nz = W * H
    loss_cpu, gc, gn, gl = 0.0, 0.0, 0.0, 0.0
    for f, R in enumerate(Rs):
        fr = oracle.Frame(W, H, True)
        fr.draw_mesh(U, m.xyz, m.rgb, R)
        assert np.array_equal(win[f * nz:(f + 1) * nz], fr.win), f
        lo, _, (a, c, d) = oracle.shade_f64(fr, oracle.mat4_mul(U, R), m.xyz, m.faces, wl["vcolor"],
                                           wl["vnormal"], wl["light"], target[f * nz:(f + 1) * nz],
                                           want_grad=True)
        loss_cpu, gc, gn, gl = loss_cpu + lo, gc + a, gn + c, gl + d
    for loss, dc, dn, dl in got:
        assert abs(loss - loss_cpu) <= 1e-5 * abs(loss_cpu)
        for g, ref in ((dc, gc), (dn, gn), (dl, gl)):
            assert np.linalg.norm(g.astype(np.float64) - ref) <= 1e-5 * np.linalg.norm(ref)
