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
