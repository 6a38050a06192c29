"""Seeded synthetic meshes and frame schedules for the raster workloads.

The reference renders OBJ/STL files (inputs.rs:95-129); the workloads here are
synthetic meshes of the same form: a triangle soup of f32 positions with w = 1
(geometry.rs:17-23, 78-81) and one u8 RGB colour per triangle, taken from the
triangle's first vertex as the reference does for vertex-coloured OBJ files
(geometry.rs:117-125).

Everything is computed in float64 and rounded to float32 once at the end, and
only uses +,-,*,/,sqrt and integer hashing (plus sin/cos for the UV sphere),
so a mesh is a pure function of its arguments.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Mesh:
    """A triangle soup: xyz [N,3,3] float32, rgb [N,3] uint8.

    Optionally carries the indexed form it came from (vertex positions,
    per-vertex colours and triangle vertex ids), used by the shading adjoint.
    """

    xyz: np.ndarray
    rgb: np.ndarray
    verts: np.ndarray | None = None      # [V,3] float32
    vcolor: np.ndarray | None = None     # [V,3] uint8
    faces: np.ndarray | None = None      # [N,3] int32

    @property
    def n(self) -> int:
        return int(self.xyz.shape[0])

    def bbox_max(self) -> np.ndarray:
        """Bounding-box maximum of an OBJ-loaded SimpleMesh (geometry.rs:85-88,
        129-135): the box starts at the origin, so each component is >= 0."""
        m = self.xyz.reshape(-1, 3).max(axis=0) if self.n else np.zeros(3, np.float32)
        return np.maximum(m, np.float32(0.0)).astype(np.float32)


def _soup(verts64: np.ndarray, faces: np.ndarray, vcolor: np.ndarray) -> Mesh:
    verts = verts64.astype(np.float32)
    faces = faces.astype(np.int32)
    xyz = np.ascontiguousarray(verts[faces])                 # [N,3,3]
    rgb = np.ascontiguousarray(vcolor[faces[:, 0]])          # first vertex's colour
    return Mesh(xyz=xyz, rgb=rgb, verts=verts, vcolor=vcolor, faces=faces)


def _vertex_colors(nv: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, size=(nv, 3), dtype=np.uint8)


def uv_sphere(n_lon: int = 32, n_lat: int = 16, radius: float = 1.0, seed: int = 0) -> Mesh:
    """UV sphere: pole fans (2*n_lon triangles) plus (n_lat-2)*n_lon quads split
    in two; outward counter-clockwise winding.  32 x 16 gives 960 triangles."""
    verts = [(0.0, radius, 0.0)]
    for i in range(1, n_lat):
        th = math.pi * i / n_lat
        st, ct = math.sin(th), math.cos(th)
        for j in range(n_lon):
            ph = 2.0 * math.pi * j / n_lon
            verts.append((radius * st * math.cos(ph), radius * ct, radius * st * math.sin(ph)))
    verts.append((0.0, -radius, 0.0))
    v = np.array(verts, dtype=np.float64)
    south = len(verts) - 1

    def ring(i, j):
        return 1 + (i - 1) * n_lon + (j % n_lon)

    faces = []
    for j in range(n_lon):
        faces.append((0, ring(1, j + 1), ring(1, j)))
    for i in range(1, n_lat - 1):
        for j in range(n_lon):
            a, b = ring(i, j), ring(i, j + 1)
            c, d = ring(i + 1, j), ring(i + 1, j + 1)
            faces.append((a, b, d))
            faces.append((a, d, c))
    for j in range(n_lon):
        faces.append((south, ring(n_lat - 1, j), ring(n_lat - 1, j + 1)))
    f = np.array(faces, dtype=np.int64)
    return _soup(v, f, _vertex_colors(v.shape[0], seed))


_ICO_FACES = np.array(
    [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
     (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
     (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)


def _icosahedron() -> np.ndarray:
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
                  (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)],
                 dtype=np.float64)
    return v / np.sqrt((v * v).sum(axis=1, keepdims=True))


def icosphere_indexed(level: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere subdivided `level` times (20*4^level triangles).  Each
    face splits into 4 consecutive children, so triangle order has spatial
    locality.  Returns (verts float64 [V,3], faces int64 [N,3])."""
    v = _icosahedron()
    f = _ICO_FACES.copy()
    for _ in range(level):
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        e = np.stack([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)], 1)
        e = e.reshape(-1, 2)
        key = np.minimum(e[:, 0], e[:, 1]) * (1 << 32) + np.maximum(e[:, 0], e[:, 1])
        uniq, inv = np.unique(key, return_inverse=True)
        lo, hi = uniq >> 32, uniq & 0xFFFFFFFF
        mid = v[lo] + v[hi]
        mid = mid / np.sqrt((mid * mid).sum(axis=1, keepdims=True))
        mid_id = (inv + v.shape[0]).reshape(-1, 3)
        v = np.concatenate([v, mid], 0)
        ab, bc, ca = mid_id[:, 0], mid_id[:, 1], mid_id[:, 2]
        f = np.stack([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                      np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)], 1).reshape(-1, 3)
    return v, f


def _hash01(ix: np.ndarray, iy: np.ndarray, iz: np.ndarray, seed: int) -> np.ndarray:
    """Integer lattice hash -> float64 in [-1, 1)."""
    m = np.uint64(0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        h = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) & m
        h ^= (iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)) & m
        h ^= (iz.astype(np.uint64) * np.uint64(0x165667B19E3779F9)) & m
        h ^= np.uint64(seed) * np.uint64(0x27D4EB2F165667C5)
        h ^= h >> np.uint64(33)
        h = (h * np.uint64(0xFF51AFD7ED558CCD)) & m
        h ^= h >> np.uint64(33)
        h = (h * np.uint64(0xC4CEB9FE1A85EC53)) & m
        h ^= h >> np.uint64(33)
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 / float(1 << 53)) - 1.0


def value_noise(p: np.ndarray, seed: int, freq: float = 4.0) -> np.ndarray:
    """3-D value noise: hashed lattice values, trilinear blend with smoothstep
    weights.  p: [V,3] float64 -> [V] float64 in [-1, 1]."""
    q = p * freq
    i0 = np.floor(q)
    t = q - i0
    s = t * t * (3.0 - 2.0 * t)
    i0 = i0.astype(np.int64)
    out = np.zeros(p.shape[0])
    for dx in (0, 1):
        wx = s[:, 0] if dx else 1.0 - s[:, 0]
        for dy in (0, 1):
            wy = s[:, 1] if dy else 1.0 - s[:, 1]
            for dz in (0, 1):
                wz = s[:, 2] if dz else 1.0 - s[:, 2]
                out += wx * wy * wz * _hash01(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed)
    return out


def displaced_icosphere(level: int, seed: int, amplitude: float = 0.05,
                        color_seed: int | None = None) -> Mesh:
    """Icosphere with radius 1 + amplitude * value_noise(seed); triangle areas
    then vary from sub-pixel to tens of pixels at 1080p."""
    v, f = icosphere_indexed(level)
    r = 1.0 + amplitude * value_noise(v, seed)
    v = v * r[:, None]
    return _soup(v, f, _vertex_colors(v.shape[0], seed if color_seed is None else color_seed))


def subdivided_box(n: int, seed: int) -> Mesh:
    """Axis-aligned cube [-1,1]^3, each face split into n x n quads, two
    triangles per quad (12 n^2 triangles), outward counter-clockwise."""
    g = np.linspace(-1.0, 1.0, n + 1)
    faces_v = []
    faces_f = []
    base = 0
    # (normal axis, sign, u axis, v axis) with u x v = sign * normal
    specs = [(0, 1, 1, 2), (0, -1, 2, 1), (1, 1, 2, 0), (1, -1, 0, 2), (2, 1, 0, 1), (2, -1, 1, 0)]
    for ax, sg, ua, va in specs:
        uu, vv = np.meshgrid(g, g, indexing="ij")
        pts = np.zeros((n + 1, n + 1, 3))
        pts[..., ax] = float(sg)
        pts[..., ua] = uu
        pts[..., va] = vv
        faces_v.append(pts.reshape(-1, 3))
        idx = np.arange((n + 1) * (n + 1)).reshape(n + 1, n + 1) + base
        a = idx[:-1, :-1].ravel()
        b = idx[1:, :-1].ravel()
        c = idx[1:, 1:].ravel()
        d = idx[:-1, 1:].ravel()
        faces_f.append(np.stack([a, b, c], 1))
        faces_f.append(np.stack([a, c, d], 1))
        base += (n + 1) * (n + 1)
    v = np.concatenate(faces_v, 0)
    f = np.concatenate(faces_f, 0)
    return _soup(v, f, _vertex_colors(v.shape[0], seed))


# ---------------------------------------------------------------------------
# transforms (host side; the device receives matrices and never evaluates trig)
# ---------------------------------------------------------------------------

def _f32(x) -> np.float32:
    return np.float32(x)


_libm = None


def _sin_cos_f32(x) -> tuple:
    """Rust's f32::sin_cos on Linux: the C library's binary32 sinf and cosf
    (not float64 sin/cos rounded to f32, which differs in the last bit for
    some angles)."""
    global _libm
    if _libm is None:
        lib = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
        for name in ("sinf", "cosf"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_float
            fn.argtypes = [ctypes.c_float]
        _libm = lib
    a = float(np.float32(x))
    return np.float32(_libm.sinf(a)), np.float32(_libm.cosf(a))


def rotation_euler(roll: float, pitch: float, yaw: float) -> np.ndarray:
    """Rotation3::from_euler_angles(roll, pitch, yaw).to_homogeneous()
    (main.rs:76-77) in binary32, as a column-major float32[16], with the
    angles' sines and cosines from the C library's sinf/cosf as Rust's
    f32::sin_cos gives them."""
    sr, cr = _sin_cos_f32(roll)
    sp, cp = _sin_cos_f32(pitch)
    sy, cy = _sin_cos_f32(yaw)
    f = _f32
    with np.errstate(all="ignore"):
        m00 = cy * cp
        m01 = cy * sp * sr - sy * cr
        m02 = cy * sp * cr + sy * sr
        m10 = sy * cp
        m11 = sy * sp * sr + cy * cr
        m12 = sy * sp * cr - cy * sr
        m20 = -sp
        m21 = cp * sr
        m22 = cp * cr
    z = f(0.0)
    return np.array([m00, m10, m20, z, m01, m11, m21, z, m02, m12, m22, z, z, z, z, f(1.0)],
                    dtype=np.float32)


def turntable_pitches(n_frames: int, start_pitch: float = 0.0) -> np.ndarray:
    """The `-j N` export schedule (main.rs:55-58, 92-99; inputs.rs:147): pitch
    starts at fl(y + pi), the step is fl(fl(2 pi) * fl(1/N)) and each frame adds
    it in f32."""
    pi = _f32(math.pi)
    pitch = _f32(_f32(start_pitch) + pi)
    step = _f32(_f32(_f32(2.0) * pi) * _f32(_f32(1.0) / _f32(n_frames)))
    out = np.empty(n_frames, np.float32)
    for k in range(n_frames):
        out[k] = pitch
        pitch = _f32(pitch + step)
    return out


def export_pitches(n_frames: int, start_pitch: float = 0.0) -> np.ndarray:
    """The pitches the `-j n_frames` export renders (main.rs:55-58, 88-106):
    the turntable schedule, stopped after n_frames frames or as soon as the
    pitch after a frame exceeds 9.42477 (then that frame is the last)."""
    pi = _f32(math.pi)
    pitch = _f32(_f32(start_pitch) + pi)
    step = _f32(_f32(_f32(2.0) * pi) * _f32(_f32(1.0) / _f32(n_frames)))
    out = []
    for k in range(n_frames):
        out.append(pitch)
        pitch = _f32(pitch + step)
        if pitch > _f32(9.42477) or k == n_frames - 1:
            break
    return np.array(out, np.float32)


def screen_transform(width: int, height: int, bbox_maxes) -> np.ndarray:
    """Context::update (context.rs:93-141): utransform from the size and the
    largest bounding-box maximum over the meshes, column-major float32[16]."""
    scale = _f32(0.0)
    for bm in bbox_maxes:
        for c in range(3):
            scale = max(scale, _f32(bm[c])) if not np.isnan(bm[c]) else scale
    fw, fh = _f32(width & 0xFFFF), _f32(height & 0xFFFF)
    with np.errstate(all="ignore"):
        s = _f32(_f32(_f32(min(fh, _f32(fw / _f32(2.0)))) / scale) / _f32(2.0))
    z = _f32(0.0)
    return np.array([s, z, z, z, z, -s, z, z, z, z, s, z,
                     _f32(fw / _f32(4.0)), _f32(fh / _f32(2.0)), z, _f32(1.0)], dtype=np.float32)


# ---------------------------------------------------------------------------
# named workloads
# ---------------------------------------------------------------------------

def workload_sphere_160x48():
    """One UV sphere (32 x 16, 960 triangles, seed 0), pitch fl(0.3 + pi) as
    `-y 0.3` gives, 160x48 image mode."""
    m = uv_sphere(32, 16, 1.0, seed=0)
    pitch = turntable_pitches(1, 0.3)[0]
    R = rotation_euler(0.0, float(pitch), 0.0)
    U = screen_transform(160, 48, [m.bbox_max()])
    return m, 160, 48, U, R


def workload_icosphere_4k():
    """Displaced icosphere, 8 subdivisions (1,310,720 triangles, seed 1), one
    frame at 3840x2160 image mode, default pitch (pi)."""
    m = displaced_icosphere(8, seed=1)
    R = rotation_euler(0.0, float(turntable_pitches(1)[0]), 0.0)
    U = screen_transform(3840, 2160, [m.bbox_max()])
    return m, 3840, 2160, U, R


def workload_sweep_1080p(n_frames: int = 256, level: int = 7):
    """Displaced icosphere, 7 subdivisions (327,680 triangles, seed 2), the
    `-j n_frames` turntable schedule at 1920x1080 image mode."""
    m = displaced_icosphere(level, seed=2)
    pitches = turntable_pitches(n_frames)
    Rs = np.stack([rotation_euler(0.0, float(p), 0.0) for p in pitches])
    U = screen_transform(1920, 1080, [m.bbox_max()])
    return m, 1920, 1080, U, Rs


RESOLUTIONS = [(80, 24), (640, 480), (1280, 720), (2560, 1440)]


def workload_uneven_meshes(n_meshes: int = 64, seed: int = 3, lo: float = 1e3, hi: float = 1e6,
                           scale_counts: float = 1.0):
    """64 meshes with triangle counts log-uniform in [lo, hi]: an icosphere at
    the level whose 20*4^k is nearest, or a subdivided box (12 n^2 triangles,
    n = round(sqrt(N/12))); each with its own resolution and seeded rotation.
    Returns a list of (mesh, W, H, U, R)."""
    rng = np.random.default_rng(seed)
    items = []
    for i in range(n_meshes):
        target = math.exp(rng.uniform(math.log(lo), math.log(hi))) * scale_counts
        kind = rng.integers(0, 2)
        W, H = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
        roll, pitch, yaw = (float(x) for x in rng.uniform(0.0, 2.0 * math.pi, size=3))
        if kind == 0:
            k = int(np.argmin([abs(20 * 4 ** j - target) for j in range(0, 9)]))
            m = displaced_icosphere(k, seed=seed * 1000 + i)
        else:
            nn = max(1, int(round(math.sqrt(target / 12.0))))
            m = subdivided_box(nn, seed=seed * 1000 + i)
        R = rotation_euler(roll, pitch, yaw)
        U = screen_transform(W, H, [m.bbox_max()])
        items.append((m, W, H, U, R))
    return items


def workload_shading_views(n_views: int = 32, size: int = 1024, n_lon: int = 224, n_lat: int = 224,
                           seed: int = 5):
    """The differentiable-shading workload: a UV sphere (224 x 224 = 99,904
    triangles), n_views seeded rotations at size x size image mode, per-vertex
    colours uniform in [0, 1), unit radial normals, light (0.577, 0.577,
    -0.577); the target's parameters are the colours + N(0, 0.1^2) and the
    light rotated 10 degrees about y.  Returns a dict of those arrays."""
    m = uv_sphere(n_lon, n_lat, 1.0, seed=seed)
    rng = np.random.default_rng(seed)
    U = screen_transform(size, size, [m.bbox_max()])
    Rs = [rotation_euler(*rng.uniform(0, 2 * math.pi, 3)) for _ in range(n_views)]
    vc = rng.uniform(0, 1, (m.verts.shape[0], 3)).astype(np.float32)
    vn = (m.verts / np.linalg.norm(m.verts, axis=1, keepdims=True)).astype(np.float32)
    light = np.array([0.577, 0.577, -0.577], np.float32)
    a = math.radians(10.0)
    Ry = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    vc_target = (vc + rng.normal(0, 0.1, vc.shape)).astype(np.float32)
    light_target = (Ry @ light.astype(np.float64)).astype(np.float32)
    return {"mesh": m, "W": size, "H": size, "U": U, "Rs": Rs, "vcolor": vc, "vnormal": vn, "light": light,
            "vcolor_target": vc_target, "light_target": light_target}
