"""The no-allocation entry point (sloth_render_frame_noalloc) against the
context path it varies (sloth_context_create + draw_mesh + read), bit for bit,
on two calls in a row -- the second with a smaller frame and mesh -- and both
against the oracle.  Device buffers are the caller's (hipMalloc'd here
through the library's own HIP runtime)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from _helpers import assert_frame_equal, oracle_render, say

pytestmark = pytest.mark.gpu


class _Dev:
    """Device buffers owned by the test (the caller)."""

    def __init__(self, sloth):
        sloth.load_library()
        self.hip = C.CDLL("libamdhip64.so.7")   # the runtime the library already loaded
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(1, nbytes)) == 0
        self.bufs.append(p)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        self.hip.hipDeviceSynchronize()
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _context_frame(sloth, m, W, H, U, R, image):
    ctx = sloth.Context(W, H, image=image)
    ctx.utransform = U
    ctx.clear()
    sloth.draw_mesh(ctx, sloth.SimpleMesh.from_mesh(m), R)
    return ctx.read()


def test_noalloc_matches_context_path(sloth):
    from sloth_amd import meshgen

    dev = _Dev(sloth)
    try:
        cases = [
            ("640x480 displaced icosphere s4", meshgen.displaced_icosphere(4, seed=12), 640, 480, True,
             meshgen.rotation_euler(0.2, 2.4, 0.1)),
            ("161x48 uv sphere 32x16, terminal mode", meshgen.uv_sphere(32, 16, seed=0), 161, 48, False,
             meshgen.rotation_euler(0.0, 3.4, 0.0)),
            ("160x48 uv sphere 16x8", meshgen.uv_sphere(16, 8, seed=3), 160, 48, True,
             meshgen.rotation_euler(0.3, 1.2, 0.0)),
        ]
        for name, m, W, H, image, R in cases:
            say(f"case noalloc {name}: {m.n} tris")
            U = meshgen.screen_transform(W, H, [m.bbox_max()])
            nc = W * H + (H if image else 0)
            d_xyz, d_rgb = dev.upload(m.xyz), dev.upload(m.rgb)
            d_cells, d_z = dev.alloc(4 * nc), dev.alloc(4 * W * H)
            sloth.render_frame_noalloc(d_xyz, d_rgb, m.n, W, H, image, U, R, d_cells, d_z)
            cells = dev.download(d_cells, nc, np.uint32)
            z = dev.download(d_z, W * H, np.float32)
            c_cells, c_z, _ = _context_frame(sloth, m, W, H, U, R, image)
            assert np.array_equal(cells, c_cells)
            assert np.array_equal(z.view(np.uint32), c_z.view(np.uint32))
            ref = oracle_render(W, H, U, [(m.xyz, m.rgb, R, None)], image=image)
            assert_frame_equal(cells, z, None, ref, W)

        # frames past the fixed scratch are refused, nothing is launched
        with pytest.raises(sloth.SlothError) as e:
            sloth.render_frame_noalloc(d_xyz, d_rgb, 0, 3840, 2160, True, U, R, d_cells, d_z)
        assert e.value.code == sloth.SLOTH_EINVAL
    finally:
        dev.free()


def test_noalloc_random_sequence(sloth):
    """Twenty calls in a row on one device, each frame of a random size (up to
    2560 x 1440, the largest that fits the fixed scratch, and down to one
    cell), parity and mode, with random soups or closed meshes: each call
    starts from what the previous one left in the shared scratch, and each
    frame must equal the oracle's."""
    from sloth_amd import meshgen

    rng = np.random.default_rng(77)
    dev = _Dev(sloth)
    try:
        for k in range(20):
            if k == 7:
                W, H = 2560, 1440
            else:
                W, H = int(rng.integers(1, 400)), int(rng.integers(1, 300))
            image = bool(rng.integers(0, 2))
            if rng.integers(0, 2):
                m = meshgen.displaced_icosphere(int(rng.integers(1, 5)), seed=int(rng.integers(0, 100)))
                xyz, rgb = m.xyz, m.rgb
            else:
                n = int(rng.integers(1, 3000))
                xyz = (rng.uniform(-1, 1, (n, 1, 3)) + rng.normal(0, 0.1, (n, 3, 3))).astype(np.float32)
                rgb = rng.integers(0, 256, (n, 3), dtype=np.uint8)
            U = meshgen.screen_transform(W, H, [np.array([1.0, 1.0, 1.0], np.float32)])
            R = meshgen.rotation_euler(*rng.uniform(0, 6.283, 3))
            say(f"case noalloc sequence {k}: {W}x{H} image={image} {len(xyz)} tris")
            nc = W * H + (H if image else 0)
            d_xyz, d_rgb = dev.upload(xyz), dev.upload(rgb)
            d_cells, d_z = dev.alloc(4 * nc), dev.alloc(4 * W * H)
            sloth.render_frame_noalloc(d_xyz, d_rgb, len(xyz), W, H, image, U, R, d_cells, d_z)
            cells = dev.download(d_cells, nc, np.uint32)
            z = dev.download(d_z, W * H, np.float32)
            assert_frame_equal(cells, z, None, oracle_render(W, H, U, [(xyz, rgb, R, None)], image=image), W)
            dev.free()
    finally:
        dev.free()
