"""CPU-side checks of the C-ABI library: it loads, exports exactly the
functions include/sloth_hip.h declares, the Python binding and the Rust
binding of INTEGRATION.md declare the same set, and argument errors come back as status codes without touching a GPU."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sloth_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(sloth_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    assert "sloth_draw_mesh" in names and "sloth_context_clear" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    import sloth_amd

    L = sloth_amd.load_library()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_covers_header():
    import sloth_amd

    assert sorted(sloth_amd.SIGNATURES) == declared_functions()


def test_rust_binding_covers_header():
    """The Rust extern block in INTEGRATION.md declares every entry point of
    the header, and nothing else."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = text[text.index('extern "C" {'):]
    block = block[:block.index("\n}")]
    names = sorted(set(re.findall(r"\bpub fn (sloth_[a-z0-9_]+)\s*\(", block)))
    assert names == declared_functions()


def test_struct_layouts():
    import sloth_amd

    assert C.sizeof(sloth_amd.sloth_ramp) == 4 + 64 + 16 + 4
    assert C.sizeof(sloth_amd.sloth_frame_desc) == 12 + 64
    assert C.sizeof(sloth_amd.sloth_draw_desc) == 8 + 8 + 64


def test_argument_errors_without_device():
    import sloth_amd

    L = sloth_amd.load_library()
    out = C.c_void_p()
    assert L.sloth_context_create(0, 10, 1, C.byref(out)) == sloth_amd.SLOTH_EINVAL
    assert L.sloth_context_create(10, 70000, 1, C.byref(out)) == sloth_amd.SLOTH_EINVAL
    assert L.sloth_context_create(10, 10, 1, None) == sloth_amd.SLOTH_EINVAL
    assert L.sloth_draw_mesh(None, None, None, None) == sloth_amd.SLOTH_EINVAL
    assert b"null" in L.sloth_last_error()
    assert L.sloth_batch_create(None, 0, None, 0, None, C.byref(out)) == sloth_amd.SLOTH_EINVAL
    r = sloth_amd.sloth_ramp()
    r.n = 17
    assert L.sloth_mesh_destroy(None) == 0


def test_host_helpers_match_oracle():
    """sloth_mat4_mul / sloth_screen_transform are host arithmetic; they must
    agree bit for bit with the oracle's restatement."""
    import oracle
    import sloth_amd

    rng = np.random.default_rng(0)
    for _ in range(200):
        a = rng.normal(0, 10, 16).astype(np.float32)
        b = rng.normal(0, 10, 16).astype(np.float32)
        assert np.array_equal(sloth_amd.mat4_mul(a, b).view(np.uint32), oracle.mat4_mul(a, b).view(np.uint32))
    for W, H in ((160, 48), (1920, 1080), (3840, 2160), (7, 3), (65535, 1)):
        bm = rng.uniform(0.1, 3, (3, 3)).astype(np.float32)
        assert np.array_equal(sloth_amd.screen_transform(W, H, bm).view(np.uint32),
                              oracle.update(W, H, bm).view(np.uint32))


def test_default_shader_table():
    import sloth_amd

    r = sloth_amd.sloth_ramp()
    assert sloth_amd.load_library().sloth_default_ramp(C.byref(r)) == 0
    assert r.n == 9 and bytes(r.glyphs[:9]) == b".:-=+*#%@" and r.fallback == ord(" ")
    assert [np.float32(r.thresholds[i]) for i in range(9)] == [
        np.float32(x) for x in (0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0)]


def test_rotation_helper_matches_oracle():
    """sloth_rotation_euler (host code of the library, C library sinf/cosf)
    is bit for bit the oracle's Rotation3::from_euler_angles."""
    import numpy as np

    import oracle
    import sloth_amd
    from sloth_amd import meshgen

    rng = np.random.default_rng(5)
    cases = [(0.0, float(p), 0.0) for p in meshgen.turntable_pitches(256)]
    cases += [tuple(float(v) for v in r) for r in rng.uniform(-7, 7, (2000, 3)).astype(np.float32)]
    for r, p, y in cases:
        a = sloth_amd.rotation_euler(r, p, y)
        b = oracle.rotation_euler(r, p, y)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (r, p, y)


def test_export_schedule():
    """The -j export stops after the requested count or once the pitch after
    a frame exceeds 9.42477 (main.rs:88-106)."""
    import numpy as np

    from sloth_amd import meshgen

    p = meshgen.export_pitches(100)
    assert len(p) == 100 and np.array_equal(p, meshgen.turntable_pitches(100))
    q = meshgen.export_pitches(20, 2.5)
    step = np.float32(np.float32(np.float32(2.0) * np.float32(np.pi)) * np.float32(np.float32(1.0) / np.float32(20)))
    assert len(q) < 20 and np.float32(q[-1] + step) > np.float32(9.42477) and q[-1] <= np.float32(9.42477)
