"""ctypes binding of the CPU oracle (liboracle.so) -- test infrastructure.

Only the tests, the smoke check and the benchmark's CPU-baseline leg import
this module; the product never does.  The C sources restate the reference's
rasterizer.rs / geometry.rs / context.rs in binary32 with its operation order.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

CELL_BLANK = 0x20          # (' ', (0,0,0))
CELL_NEWLINE = 0x0A        # ('\n', (0,0,0))
F32_MAX = np.float32(3.4028234663852886e38)


class Ramp(C.Structure):
    _fields_ = [("n", C.c_uint32), ("thresholds", C.c_float * 16), ("glyphs", C.c_uint8 * 16),
                ("fallback", C.c_uint8), ("pad", C.c_uint8 * 3)]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        u32p = C.POINTER(C.c_uint32)
        u8p = C.POINTER(C.c_uint8)
        i32p = C.POINTER(C.c_int32)
        u64p = C.POINTER(C.c_uint64)
        L.oracle_mat4_mul.argtypes = [fp, fp, fp]
        L.oracle_rotation_euler.argtypes = [C.c_float, C.c_float, C.c_float, fp]
        L.oracle_tri_mul.argtypes = [fp, fp]
        L.oracle_tri_aabb.argtypes = [fp, fp]
        L.oracle_tri_normal.argtypes = [fp, fp]
        L.oracle_default_shader.argtypes = [C.c_float]
        L.oracle_default_shader.restype = C.c_uint8
        L.oracle_ramp_shade.argtypes = [C.POINTER(Ramp), C.c_float]
        L.oracle_ramp_shade.restype = C.c_uint8
        L.oracle_default_ramp.argtypes = [C.POINTER(Ramp)]
        L.oracle_draw_mesh.argtypes = [C.c_uint32, C.c_uint32, C.c_int, fp, u32p, fp, i32p, i32p, fp, u8p,
                                       C.c_uint32, fp, C.POINTER(Ramp), C.c_uint32, u64p]
        dp = C.POINTER(C.c_double)
        L.oracle_shade_f64.argtypes = [C.c_uint32, C.c_uint32, fp, fp, i32p, C.c_uint32, dp, dp, dp, i32p, i32p,
                                       fp, dp, dp, dp, dp]
        L.oracle_shade_f64.restype = C.c_double
        L.oracle_update.argtypes = [C.c_uint32, C.c_uint32, fp, C.c_uint32, fp]
        L.oracle_mesh_bbox.argtypes = [fp, C.c_uint32, fp]
        L.oracle_clear.argtypes = [C.c_uint32, C.c_uint32, C.c_int, u32p, fp]
        L.oracle_render_frames.argtypes = [C.c_uint32, C.c_uint32, C.c_int, fp, fp, C.c_uint32, fp,
                                           u8p, C.c_uint32, u32p, fp, C.POINTER(Ramp), C.c_int]
        L.oracle_render_frames.restype = C.c_int
        L.oracle_encode.argtypes = [u32p, C.c_uint64, C.c_int, C.c_int, C.c_char_p, C.c_size_t]
        L.oracle_encode.restype = C.c_size_t
        L.oracle_shader_exhaustive_mismatch.argtypes = [C.c_uint32, C.c_uint32]
        L.oracle_shader_exhaustive_mismatch.restype = C.c_uint32
        L.oracle_set_arith_variant.argtypes = [C.c_int]
        _lib = L
    return _lib


NORM_SEQUENTIAL = 1    # norm_squared as ((x*x + y*y) + z*z) + w*w
MATVEC_PAIRWISE = 2    # Matrix4 * Vector4 rows as (m0 x + m2 z) + (m1 y + m3 w)


def set_arith_variant(v: int) -> None:
    """Alternative operation orders for the oracle's calls on the calling
    thread (0 = the restatement of nalgebra 0.22's order).  For bounding what
    the unpinned order could change; never used to check the GPU."""
    lib().oracle_set_arith_variant(int(v))


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(C.POINTER(t))


def default_ramp() -> Ramp:
    r = Ramp()
    lib().oracle_default_ramp(C.byref(r))
    return r


def make_ramp(thresholds, glyphs, fallback) -> Ramp:
    r = Ramp()
    r.n = len(thresholds)
    for i, (t, g) in enumerate(zip(thresholds, glyphs)):
        r.thresholds[i] = float(np.float32(t))
        r.glyphs[i] = ord(g) if isinstance(g, str) else int(g)
    r.fallback = ord(fallback) if isinstance(fallback, str) else int(fallback)
    return r


def mat4_mul(A, B) -> np.ndarray:
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    out = np.empty(16, np.float32)
    lib().oracle_mat4_mul(_p(A, C.c_float), _p(B, C.c_float), _p(out, C.c_float))
    return out


def rotation_euler(roll, pitch, yaw) -> np.ndarray:
    out = np.empty(16, np.float32)
    lib().oracle_rotation_euler(float(np.float32(roll)), float(np.float32(pitch)),
                                float(np.float32(yaw)), _p(out, C.c_float))
    return out


def tri_mul(M, tri) -> np.ndarray:
    M = np.ascontiguousarray(M, np.float32)
    t = np.ascontiguousarray(tri, np.float32).reshape(12).copy()
    lib().oracle_tri_mul(_p(M, C.c_float), _p(t, C.c_float))
    return t.reshape(3, 4)


def tri_aabb(tri) -> tuple[np.ndarray, np.ndarray]:
    t = np.ascontiguousarray(tri, np.float32).reshape(12)
    out = np.empty(8, np.float32)
    lib().oracle_tri_aabb(_p(t, C.c_float), _p(out, C.c_float))
    return out[:4], out[4:]


def tri_normal(tri) -> np.ndarray:
    t = np.ascontiguousarray(tri, np.float32).reshape(12)
    out = np.empty(4, np.float32)
    lib().oracle_tri_normal(_p(t, C.c_float), _p(out, C.c_float))
    return out


def default_shader(s: float) -> str:
    return chr(lib().oracle_default_shader(float(np.float32(s))))


def update(width, height, bbox_maxes) -> np.ndarray:
    bm = np.ascontiguousarray(np.asarray(bbox_maxes, np.float32).reshape(-1, 3))
    out = np.empty(16, np.float32)
    lib().oracle_update(width, height, _p(bm, C.c_float), bm.shape[0], _p(out, C.c_float))
    return out


def mesh_bbox(xyz) -> tuple[np.ndarray, np.ndarray]:
    x = np.ascontiguousarray(xyz, np.float32)
    out = np.empty(6, np.float32)
    lib().oracle_mesh_bbox(_p(x, C.c_float), x.shape[0], _p(out, C.c_float))
    return out[:3], out[3:]


class Frame:
    """Context state of one frame: cells (packed glyph|R|G|B uint32), z, and
    the winning global triangle index per z slot (-1 = none)."""

    def __init__(self, width: int, height: int, image: bool = True):
        self.W, self.H, self.image = width, height, bool(image)
        self.cells = np.empty(width * height + (height if image else 0), np.uint32)
        self.z = np.empty(width * height, np.float32)
        self.win = np.full(width * height, -1, np.int32)
        self.winxy = np.full(2 * width * height, -1, np.int32)   # winning pixel (x, y) per z slot
        self.tri_base = 0
        self.stats = np.zeros(3, np.uint64)   # candidates, inside, z-passes
        self.clear()

    def clear(self):
        lib().oracle_clear(self.W, self.H, int(self.image), _p(self.cells, C.c_uint32),
                           _p(self.z, C.c_float))
        self.win[:] = -1
        self.winxy[:] = -1
        self.tri_base = 0

    def draw_mesh(self, U, xyz, rgb, R, ramp: Ramp | None = None):
        U = np.ascontiguousarray(U, np.float32)
        R = np.ascontiguousarray(R, np.float32)
        xyz = np.ascontiguousarray(xyz, np.float32)
        rgb = np.ascontiguousarray(rgb, np.uint8)
        ramp = ramp if ramp is not None else default_ramp()
        n = xyz.shape[0]
        lib().oracle_draw_mesh(self.W, self.H, int(self.image), _p(U, C.c_float),
                               _p(self.cells, C.c_uint32), _p(self.z, C.c_float),
                               _p(self.win, C.c_int32), _p(self.winxy, C.c_int32), _p(xyz, C.c_float),
                               _p(rgb, C.c_uint8),
                               n, _p(R, C.c_float), C.byref(ramp), self.tri_base,
                               _p(self.stats, C.c_uint64))
        self.tri_base += n

    def glyphs(self) -> np.ndarray:
        return (self.cells & 0xFF).astype(np.uint8)

    def text(self) -> str:
        return self.glyphs().tobytes().decode("latin-1")


def render(width, height, U, xyz, rgb, R, image=True, ramp=None) -> Frame:
    f = Frame(width, height, image)
    f.draw_mesh(U, xyz, rgb, R, ramp)
    return f


def render_frames(width, height, U, Rs, xyz, rgb, image=True, nthreads=1, keep=True, ramp=None):
    """Frames with per-frame utransform U[f] and transform Rs[f] (clear +
    draw_mesh each).  Returns (cells [F,Ccells], z [F,W*H]) or (None, None)."""
    Rs = np.ascontiguousarray(np.asarray(Rs, np.float32).reshape(-1, 16))
    F = Rs.shape[0]
    U = np.asarray(U, np.float32).reshape(-1, 16)
    if U.shape[0] == 1:
        U = np.repeat(U, F, axis=0)
    U = np.ascontiguousarray(U)
    xyz = np.ascontiguousarray(xyz, np.float32)
    rgb = np.ascontiguousarray(rgb, np.uint8)
    nc = width * height + (height if image else 0)
    cells = np.empty((F, nc), np.uint32) if keep else None
    z = np.empty((F, width * height), np.float32) if keep else None
    ramp = ramp if ramp is not None else default_ramp()
    rc = lib().oracle_render_frames(
        width, height, int(image), _p(U, C.c_float), _p(Rs, C.c_float), F, _p(xyz, C.c_float),
        _p(rgb, C.c_uint8), xyz.shape[0], _p(cells, C.c_uint32) if keep else None,
        _p(z, C.c_float) if keep else None, C.byref(ramp), int(nthreads))
    if rc != 0:
        raise RuntimeError("oracle_render_frames failed")
    return cells, z


def shader_exhaustive_mismatch(lo: int = 0, hi: int = 65536) -> int:
    return int(lib().oracle_shader_exhaustive_mismatch(lo, hi))


def shade_f64(frame: Frame, M, xyz, faces, vcolor, vnormal, light, target=None, want_out=False,
              want_grad=False):
    """Differentiable-shading model in float64 over the visible fragments of
    `frame` (one draw; M = utransform * transform).  Returns
    (loss, out [W*H,3] or None, (dcolor, dnormal, dlight) or None)."""
    M = np.ascontiguousarray(M, np.float32).reshape(16)
    xyz = np.ascontiguousarray(xyz, np.float32)
    faces = np.ascontiguousarray(faces, np.int32)
    vc = np.ascontiguousarray(vcolor, np.float64)
    vn = np.ascontiguousarray(vnormal, np.float64)
    L = np.ascontiguousarray(light, np.float64)
    nv = vc.shape[0]
    dp = C.POINTER(C.c_double)
    tgt = None if target is None else np.ascontiguousarray(target, np.float32)
    out = np.zeros((frame.W * frame.H, 3)) if want_out else None
    grads = (np.zeros((nv, 3)), np.zeros((nv, 3)), np.zeros(3)) if want_grad else None
    loss = lib().oracle_shade_f64(
        frame.W, frame.H, _p(M, C.c_float), _p(xyz, C.c_float), _p(faces, C.c_int32), nv, _p(vc, C.c_double),
        _p(vn, C.c_double), _p(L, C.c_double), _p(frame.win, C.c_int32), _p(frame.winxy, C.c_int32),
        _p(tgt, C.c_float) if tgt is not None else None, _p(out, C.c_double) if out is not None else None,
        _p(grads[0], C.c_double) if grads else None, _p(grads[1], C.c_double) if grads else None,
        _p(grads[2], C.c_double) if grads else None)
    return float(loss), out, grads


def encode(cells, mode: int, image: bool = True) -> bytes:
    """Context::flush's bytes: mode 0 plain text (context.rs:58-61), mode 1
    the colour terminal stream (context.rs:62-78), mode 2 the -j HTML spans
    (context.rs:79-87); a terminal-mode frame (image False) starts with the
    cursor move (context.rs:53-55)."""
    c = np.ascontiguousarray(cells, np.uint32)
    n = lib().oracle_encode(_p(c, C.c_uint32), c.size, mode, int(bool(image)), None, 0)
    buf = C.create_string_buffer(max(1, n))
    lib().oracle_encode(_p(c, C.c_uint32), c.size, mode, int(bool(image)), buf, n)
    return buf.raw[:n]
