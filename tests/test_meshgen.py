"""Synthetic workloads (CPU): sizes named by the benchmark configs, the
turntable schedule of main.rs:55-58, 92-99, determinism."""
from __future__ import annotations

import numpy as np

from sloth_amd import meshgen


def test_counts():
    assert meshgen.uv_sphere(32, 16).n == 960
    assert meshgen.uv_sphere(224, 224).n == 99_904
    v, f = meshgen.icosphere_indexed(3)
    assert f.shape[0] == 20 * 4 ** 3 and v.shape[0] == 10 * 4 ** 3 + 2
    assert meshgen.subdivided_box(5, seed=0).n == 12 * 25


def test_outward_winding():
    """Triangles wind counter-clockwise seen from outside (normals point away
    from the centre), so front faces are the ones facing the viewer."""
    for m in (meshgen.uv_sphere(16, 8), meshgen.displaced_icosphere(2, seed=0), meshgen.subdivided_box(2, seed=0)):
        t = m.xyz.astype(np.float64)
        n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
        c = t.mean(axis=1)
        assert ((n * c).sum(axis=1) > 0).all()


def test_deterministic():
    a = meshgen.displaced_icosphere(4, seed=7)
    b = meshgen.displaced_icosphere(4, seed=7)
    assert np.array_equal(a.xyz.view(np.uint32), b.xyz.view(np.uint32)) and np.array_equal(a.rgb, b.rgb)
    c = meshgen.displaced_icosphere(4, seed=8)
    assert not np.array_equal(a.xyz, c.xyz)


def test_turntable_schedule():
    """pitch_0 = fl(pi), step = fl(fl(2 pi) * fl(1/N)), accumulated in f32."""
    p = meshgen.turntable_pitches(100)
    pi = np.float32(np.pi)
    step = np.float32(np.float32(2 * pi) * np.float32(np.float32(1.0) / np.float32(100)))
    acc = pi
    for k in range(100):
        assert p[k] == acc
        acc = np.float32(acc + step)
    assert meshgen.turntable_pitches(1, 0.3)[0] == np.float32(np.float32(0.3) + pi)


def test_rotation_matches_libm_sinf_cosf():
    """The workloads' model matrices are Rotation3::from_euler_angles in
    binary32 with the C library's sinf/cosf (main.rs:76-77, Rust's
    f32::sin_cos): bit for bit the oracle's, over every pitch of the 256-frame
    turntable and 2000 random Euler triples.  float64 sin/cos rounded to f32
    differ from sinf/cosf in the last bit on 7 of those pitches, so this test
    fails for that construction."""
    import oracle

    for p in meshgen.turntable_pitches(256):
        a = meshgen.rotation_euler(0.0, float(p), 0.0)
        b = oracle.rotation_euler(0.0, float(p), 0.0)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), p
    rng = np.random.default_rng(11)
    for roll, pitch, yaw in rng.uniform(-7.0, 7.0, size=(2000, 3)).astype(np.float32):
        a = meshgen.rotation_euler(float(roll), float(pitch), float(yaw))
        b = oracle.rotation_euler(float(roll), float(pitch), float(yaw))
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (roll, pitch, yaw)


def test_rotation_f64_rounding_differs():
    """Guard for the test above: the float64-rounded construction is a
    different function on the turntable (so the pin has teeth)."""
    import math

    diff = 0
    for p in meshgen.turntable_pitches(256):
        s, c = meshgen._sin_cos_f32(p)
        diff += int(s != np.float32(math.sin(float(p))) or c != np.float32(math.cos(float(p))))
    assert diff > 0
