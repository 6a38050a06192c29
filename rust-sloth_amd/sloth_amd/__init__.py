"""Python mirror of sloth's raster API over libsloth_hip.so (ctypes).

Names follow the reference (ecumene/rust-sloth):

    SimpleMesh      geometry.rs:78-81   triangle soup + bounding_box
    Context         context.rs:12-142   blank / clear / update / utransform,
                                        frame_buffer / z_buffer
    draw_mesh       rasterizer.rs:39-46
    default_shader  rasterizer.rs:5-27  (a step table, see Shader)

The rasterization runs on the GPU through the C-ABI in include/sloth_hip.h.
There is no CPU fallback: importing works without a GPU, but creating a
Context or SimpleMesh raises SlothError if the library or the device is
missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import meshgen  # noqa: F401  (synthetic workloads)

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsloth_hip.so")

SLOTH_OK, SLOTH_EINVAL, SLOTH_ENOMEM, SLOTH_EHIP, SLOTH_ENODEV, SLOTH_ESTATE = range(6)
ENCODE_TEXT, ENCODE_ANSI, ENCODE_HTML = 0, 1, 2


def _encode(fn, *args) -> bytes:
    n = C.c_uint64()
    _check(fn(*args, None, 0, C.byref(n)))
    buf = C.create_string_buffer(max(1, n.value))
    _check(fn(*args, buf, n.value, C.byref(n)))
    return buf.raw[:n.value]
_ERRNAMES = {1: "EINVAL", 2: "ENOMEM", 3: "EHIP", 4: "ENODEV", 5: "ESTATE"}


class SlothError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sloth error {_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class sloth_ramp(C.Structure):
    _fields_ = [("n", C.c_uint32), ("thresholds", C.c_float * 16), ("glyphs", C.c_uint8 * 16),
                ("fallback", C.c_uint8), ("reserved", C.c_uint8 * 3)]


class sloth_frame_desc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("image", C.c_int32),
                ("utransform", C.c_float * 16)]


class sloth_draw_desc(C.Structure):
    _fields_ = [("frame", C.c_uint32), ("mesh", C.c_void_p), ("transform", C.c_float * 16)]


# exported symbol -> (restype, argtypes); the CPU tests check this list against
# include/sloth_hip.h
_fp = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_vp = C.c_void_p
_vpp = C.POINTER(C.c_void_p)
SIGNATURES = {
    "sloth_init": (C.c_int, [C.c_int]),
    "sloth_last_error": (C.c_char_p, []),
    "sloth_default_ramp": (C.c_int, [C.POINTER(sloth_ramp)]),
    "sloth_mesh_create": (C.c_int, [_fp, _u8p, C.c_uint32, _vpp]),
    "sloth_mesh_create_device": (C.c_int, [_vp, _vp, C.c_uint32, _vpp]),
    "sloth_mesh_create_indexed": (C.c_int, [_fp, C.c_uint32, _u32p, _u8p, C.c_uint32, _vpp]),
    "sloth_mesh_is_indexed": (C.c_int, [_vp, _u32p]),
    "sloth_mesh_bbox": (C.c_int, [_vp, _fp, _fp]),
    "sloth_mesh_triangles": (C.c_uint32, [_vp]),
    "sloth_mesh_destroy": (C.c_int, [_vp]),
    "sloth_model_load": (C.c_int, [C.c_char_p, _vpp]),
    "sloth_model_mesh_count": (C.c_int, [_vp, C.POINTER(C.c_uint32)]),
    "sloth_model_mesh": (C.c_int, [_vp, C.c_uint32, C.POINTER(_fp), C.POINTER(_u8p), C.POINTER(C.c_uint32),
                                   _fp, _fp]),
    "sloth_model_upload": (C.c_int, [_vp, C.c_uint32, _vpp]),
    "sloth_model_destroy": (C.c_int, [_vp]),
    "sloth_context_create": (C.c_int, [C.c_uint32, C.c_uint32, C.c_int, _vpp]),
    "sloth_context_destroy": (C.c_int, [_vp]),
    "sloth_context_update": (C.c_int, [_vp, _fp, C.c_uint32]),
    "sloth_context_set_utransform": (C.c_int, [_vp, _fp]),
    "sloth_context_get_utransform": (C.c_int, [_vp, _fp]),
    "sloth_context_clear": (C.c_int, [_vp]),
    "sloth_draw_mesh": (C.c_int, [_vp, _vp, _fp, C.POINTER(sloth_ramp)]),
    "sloth_context_render": (C.c_int, [_vp, _vp]),
    "sloth_context_read": (C.c_int, [_vp, _u32p, _fp, _i32p]),
    "sloth_context_encode": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "sloth_batch_encode_frame": (C.c_int, [_vp, C.c_uint32, C.c_int, C.c_char_p, C.c_uint64,
                                           C.POINTER(C.c_uint64)]),
    "sloth_context_device_buffers": (C.c_int, [_vp, _vpp, _vpp, C.POINTER(C.c_uint64),
                                               C.POINTER(C.c_uint64)]),
    "sloth_batch_create": (C.c_int, [C.POINTER(sloth_frame_desc), C.c_uint32,
                                     C.POINTER(sloth_draw_desc), C.c_uint32,
                                     C.POINTER(sloth_ramp), _vpp]),
    "sloth_batch_set_transforms": (C.c_int, [_vp, _fp]),
    "sloth_batch_run": (C.c_int, [_vp, _vp]),
    "sloth_batch_run_per_frame": (C.c_int, [_vp, _vp]),
    "sloth_batch_sync": (C.c_int, [_vp]),
    "sloth_batch_read_frame": (C.c_int, [_vp, C.c_uint32, _u32p, _fp, _i32p]),
    "sloth_batch_device_frame": (C.c_int, [_vp, C.c_uint32, _vpp, _vpp]),
    "sloth_batch_encode_export": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_char_p, C.c_uint64,
                                            C.POINTER(C.c_uint64)]),
    "sloth_batch_profile": (C.c_int, [_vp, _fp, C.POINTER(C.c_char_p), _u32p]),
    "sloth_batch_time": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_int, _fp]),
    "sloth_batch_set_tile": (C.c_int, [_vp, C.c_uint32, C.c_uint32]),
    "sloth_batch_set_walk": (C.c_int, [_vp, C.c_int]),
    "sloth_batch_autotune": (C.c_int, [_vp, _u32p, _u32p]),
    "sloth_batch_plan": (C.c_int, [_vp, _u32p, _u32p, _i32p, _u32p]),
    "sloth_batch_set_bin": (C.c_int, [_vp, C.c_uint32]),
    "sloth_batch_destroy": (C.c_int, [_vp]),
    "sloth_shading_create": (C.c_int, [_vp, _u32p, C.c_uint32, C.c_uint32, _vpp]),
    "sloth_shading_visibility": (C.c_int, [_vp]),
    "sloth_shading_slots": (C.c_int, [_vp, C.POINTER(C.c_uint64)]),
    "sloth_shading_set_target": (C.c_int, [_vp, _fp]),
    "sloth_shading_forward": (C.c_int, [_vp, _fp, _fp, _fp, _fp, C.POINTER(C.c_double)]),
    "sloth_shading_backward": (C.c_int, [_vp, _fp, _fp, _fp]),
    "sloth_shading_step": (C.c_int, [_vp, _fp, _fp, _fp, C.POINTER(C.c_double), _fp, _fp, _fp]),
    "sloth_shading_time": (C.c_int, [_vp, C.c_uint32, C.c_uint32, _fp]),
    "sloth_shading_destroy": (C.c_int, [_vp]),
    "sloth_render_frame_noalloc": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, _fp, _fp,
                                             C.POINTER(sloth_ramp), _vp, _vp, _vp]),
    "sloth_screen_transform": (C.c_int, [C.c_uint32, C.c_uint32, _fp, C.c_uint32, _fp]),
    "sloth_mat4_mul": (C.c_int, [_fp, _fp, _fp]),
    "sloth_rotation_euler": (C.c_int, [C.c_float, C.c_float, C.c_float, _fp]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libsloth_hip.so and declares every exported signature.  Raises
    SlothError if the library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SlothError(SLOTH_ENODEV, f"{path} is missing: build it first (__graft_entry__.build())")
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int):
    if rc != SLOTH_OK:
        raise SlothError(rc, load_library().sloth_last_error().decode(errors="replace"))


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)


_initialised = False


def init(device: int = 0):
    """Selects and checks the GPU (gfx950)."""
    global _initialised
    _check(load_library().sloth_init(device))
    _initialised = True


def _ensure_init():
    if not _initialised:
        init(0)


# ---------------------------------------------------------------------------
# shaders
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class Shader:
    """A step shader: first i with shade <= thresholds[i] -> glyphs[i], else
    fallback (NaN included).  The reference's `Fn(f32) -> char` closures
    (rasterizer.rs:39-42) of that form."""

    thresholds: tuple
    glyphs: str
    fallback: str = " "

    def to_c(self) -> sloth_ramp:
        if len(self.thresholds) != len(self.glyphs) or len(self.glyphs) > 16:
            raise ValueError("a shader needs equally many thresholds and glyphs, at most 16")
        r = sloth_ramp()
        r.n = len(self.thresholds)
        for i, (t, g) in enumerate(zip(self.thresholds, self.glyphs)):
            r.thresholds[i] = float(np.float32(t))
            r.glyphs[i] = ord(g)
        r.fallback = ord(self.fallback)
        return r

    def __call__(self, shade: float) -> str:
        s = np.float32(shade)
        for t, g in zip(self.thresholds, self.glyphs):
            if s <= np.float32(t):
                return g
        return self.fallback


default_shader = Shader((0.20, 0.30, 0.40, 0.50, 0.60, 0.70, 0.80, 0.90, 1.0), ".:-=+*#%@", " ")


# ---------------------------------------------------------------------------
# SimpleMesh
# ---------------------------------------------------------------------------

@dataclass
class AABB:
    """geometry.rs:5-15 (xyz; w is always 1)."""

    min: np.ndarray
    max: np.ndarray


class SimpleMesh:
    """geometry.rs:78-81: a device-resident triangle soup.

    positions: [n,3,3] float32 (v1, v2, v3 as x,y,z; w = 1), colors: [n,3]
    uint8.  `bounding_box` follows the OBJ loader (geometry.rs:85-88,
    129-135)."""

    def __init__(self, positions, colors):
        _ensure_init()
        L = load_library()
        self.positions = np.ascontiguousarray(positions, np.float32).reshape(-1, 3, 3)
        self.colors = np.ascontiguousarray(colors, np.uint8).reshape(-1, 3)
        if self.positions.shape[0] != self.colors.shape[0]:
            raise ValueError("positions and colors disagree on the triangle count")
        h = C.c_void_p()
        _check(L.sloth_mesh_create(_fptr(self.positions), self.colors.ctypes.data_as(_u8p),
                                   self.positions.shape[0], C.byref(h)))
        self._h = h
        mn = np.zeros(3, np.float32)
        mx = np.zeros(3, np.float32)
        _check(L.sloth_mesh_bbox(h, _fptr(mn), _fptr(mx)))
        self.bounding_box = AABB(mn, mx)

    @classmethod
    def from_mesh(cls, m, indexed: bool = False) -> "SimpleMesh":
        """A device copy of a meshgen Mesh: its soup, or (indexed=True) its
        vertex rows and faces."""
        if indexed:
            return cls.indexed(m.verts, m.faces, m.rgb)
        return cls(m.xyz, m.rgb)

    @classmethod
    def indexed(cls, vertices, faces, colors) -> "SimpleMesh":
        """An indexed device mesh (sloth_mesh_create_indexed): vertex rows
        [V,3] float32, faces [n,3] vertex ids, colours [n,3] uint8; renders as
        the soup vertices[faces]."""
        _ensure_init()
        L = load_library()
        self = cls.__new__(cls)
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        f = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
        c = np.ascontiguousarray(colors, np.uint8).reshape(-1, 3)
        if f.shape[0] != c.shape[0]:
            raise ValueError("faces and colors disagree on the triangle count")
        h = C.c_void_p()
        _check(L.sloth_mesh_create_indexed(_fptr(v), v.shape[0], f.ctypes.data_as(_u32p), c.ctypes.data_as(_u8p),
                                           f.shape[0], C.byref(h)))
        self._h = h
        self.positions = v[f.astype(np.int64)] if f.size else np.zeros((0, 3, 3), np.float32)
        self.colors = c
        mn = np.zeros(3, np.float32)
        mx = np.zeros(3, np.float32)
        _check(L.sloth_mesh_bbox(h, _fptr(mn), _fptr(mx)))
        self.bounding_box = AABB(mn, mx)
        return self

    @property
    def is_indexed(self) -> bool:
        return bool(load_library().sloth_mesh_is_indexed(self._h, None))

    @classmethod
    def _from_model(cls, model, i: int) -> "SimpleMesh":
        L = load_library()
        pos, rgb, bmin, bmax = model.mesh_arrays(i)
        self = cls.__new__(cls)
        _ensure_init()
        self.positions, self.colors = pos, rgb
        h = C.c_void_p()
        _check(L.sloth_model_upload(model._h, i, C.byref(h)))
        self._h = h
        self.bounding_box = AABB(bmin, bmax)
        return self

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.sloth_mesh_destroy(h)
            self._h = None


class ModelFile:
    """One model file parsed on the host (inputs.rs:93-129): `.obj` through
    the tobj 3.2.2 rules, `.stl` through stl_io 0.4.2's; see
    sloth_model_load.  No GPU is needed to parse."""

    def __init__(self, path: str):
        L = load_library()
        h = C.c_void_p()
        _check(L.sloth_model_load(os.fsencode(path), C.byref(h)))
        self._h = h
        n = C.c_uint32()
        _check(L.sloth_model_mesh_count(h, C.byref(n)))
        self.n_meshes = int(n.value)

    def mesh_arrays(self, i: int):
        """(positions [n,3,3] f32, colors [n,3] u8, bbox min, bbox max), copies."""
        L = load_library()
        p, c, n = _fp(), _u8p(), C.c_uint32()
        mn = np.zeros(3, np.float32)
        mx = np.zeros(3, np.float32)
        _check(L.sloth_model_mesh(self._h, i, C.byref(p), C.byref(c), C.byref(n), _fptr(mn), _fptr(mx)))
        k = int(n.value)
        pos = np.ctypeslib.as_array(p, shape=(k * 9,)).reshape(k, 3, 3).copy() if k else np.zeros((0, 3, 3), np.float32)
        rgb = np.ctypeslib.as_array(c, shape=(k * 3,)).reshape(k, 3).copy() if k else np.zeros((0, 3), np.uint8)
        return pos, rgb, mn, mx

    def meshes(self) -> list:
        """The file's meshes as device SimpleMesh objects, in queue order."""
        return [SimpleMesh._from_model(self, i) for i in range(self.n_meshes)]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.sloth_model_destroy(h)
            self._h = None


def load_meshes(paths) -> list:
    """match_meshes (inputs.rs:93-129): the mesh queue of one or more files,
    in order (a string splits on spaces, as the reference's argument does)."""
    if isinstance(paths, str):
        paths = paths.split(" ")
    out = []
    for p in paths:
        out.extend(ModelFile(p).meshes())
    return out


# ---------------------------------------------------------------------------
# Context
# ---------------------------------------------------------------------------

class Context:
    """context.rs:12-142 on the GPU.

    The frame buffer is exposed as packed uint32 cells (byte0 glyph, bytes
    1..3 R,G,B); `frame_buffer` decodes it to the reference's
    (char, (r,g,b)) list."""

    def __init__(self, width: int, height: int, image: bool = True):
        _ensure_init()
        L = load_library()
        h = C.c_void_p()
        _check(L.sloth_context_create(width, height, int(bool(image)), C.byref(h)))
        self._h = h
        self.width, self.height, self.image = int(width), int(height), bool(image)
        self._meshes = []          # keep drawn meshes alive until clear

    @classmethod
    def blank(cls, image: bool, width: int, height: int) -> "Context":
        """Context::blank (context.rs:22-34) with image mode's fixed size."""
        return cls(width, height, image)

    @property
    def utransform(self) -> np.ndarray:
        m = np.empty(16, np.float32)
        _check(load_library().sloth_context_get_utransform(self._h, _fptr(m)))
        return m

    @utransform.setter
    def utransform(self, m):
        m = np.ascontiguousarray(m, np.float32).reshape(16)
        _check(load_library().sloth_context_set_utransform(self._h, _fptr(m)))

    def update(self, meshes) -> None:
        """Context::update (context.rs:93-141)."""
        bm = np.ascontiguousarray(np.stack([m.bounding_box.max for m in meshes]) if meshes
                                  else np.zeros((0, 3), np.float32), np.float32)
        _check(load_library().sloth_context_update(self._h, _fptr(bm), bm.shape[0]))

    def clear(self) -> None:
        """Context::clear (context.rs:35-45)."""
        _check(load_library().sloth_context_clear(self._h))
        self._meshes = []

    def draw(self, mesh: SimpleMesh, transform, shader: Shader = default_shader) -> None:
        t = np.ascontiguousarray(transform, np.float32).reshape(16)
        r = shader.to_c()
        _check(load_library().sloth_draw_mesh(self._h, mesh.handle, _fptr(t), C.byref(r)))
        self._meshes.append(mesh)

    def render(self, stream: int | None = None) -> None:
        _check(load_library().sloth_context_render(self._h, C.c_void_p(stream or 0)))

    def read(self, want_winners: bool = False):
        """(cells uint32, z float32, winners int32 or None)."""
        nc = self.width * self.height + (self.height if self.image else 0)
        cells = np.empty(nc, np.uint32)
        z = np.empty(self.width * self.height, np.float32)
        win = np.empty(self.width * self.height, np.int32) if want_winners else None
        _check(load_library().sloth_context_read(
            self._h, cells.ctypes.data_as(_u32p), _fptr(z),
            win.ctypes.data_as(_i32p) if want_winners else None))
        return cells, z, win

    @property
    def z_buffer(self) -> np.ndarray:
        return self.read()[1]

    @property
    def frame_buffer(self) -> list:
        cells = self.read()[0]
        return [(chr(c & 0xFF), ((c >> 8) & 0xFF, (c >> 16) & 0xFF, (c >> 24) & 0xFF)) for c in cells]

    def text(self) -> str:
        """The no-colour flush (context.rs:58-61) without the trailing newline."""
        return (self.read()[0] & 0xFF).astype(np.uint8).tobytes().decode("latin-1")

    def encode(self, mode: int = ENCODE_TEXT) -> bytes:
        """Context::flush's bytes (context.rs:50-92), encoded on the GPU."""
        return _encode(load_library().sloth_context_encode, self._h, mode)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.sloth_context_destroy(h)
            self._h = None


def draw_mesh(context: Context, mesh: SimpleMesh, transform, shader: Shader = default_shader) -> None:
    """rasterizer.rs:39-46."""
    context.draw(mesh, transform, shader)


def render_frame_noalloc(d_positions: int, d_colors: int, n: int, width: int, height: int, image: bool,
                         utransform, transform, d_cells: int, d_z: int, shader: Shader = default_shader,
                         stream: int | None = None) -> None:
    """One frame of one draw from device arrays into device buffers, with
    no device allocation (sloth_render_frame_noalloc); returns without
    waiting.  Pointers are device addresses (ints)."""
    _ensure_init()
    u = np.ascontiguousarray(utransform, np.float32).reshape(16)
    t = np.ascontiguousarray(transform, np.float32).reshape(16)
    r = shader.to_c()
    _check(load_library().sloth_render_frame_noalloc(
        C.c_void_p(d_positions), C.c_void_p(d_colors), n, width, height, int(bool(image)), _fptr(u), _fptr(t),
        C.byref(r), C.c_void_p(d_cells), C.c_void_p(d_z), C.c_void_p(stream or 0)))


def screen_transform(width: int, height: int, bbox_maxes) -> np.ndarray:
    bm = np.ascontiguousarray(np.asarray(bbox_maxes, np.float32).reshape(-1, 3))
    out = np.empty(16, np.float32)
    _check(load_library().sloth_screen_transform(width, height, _fptr(bm), bm.shape[0], _fptr(out)))
    return out


def rotation_euler(roll: float, pitch: float, yaw: float) -> np.ndarray:
    """Rotation3::from_euler_angles(...).to_homogeneous() (main.rs:76-77),
    column-major float32[16], from the library's host helper."""
    out = np.empty(16, np.float32)
    _check(load_library().sloth_rotation_euler(float(np.float32(roll)), float(np.float32(pitch)),
                                               float(np.float32(yaw)), _fptr(out)))
    return out


def mat4_mul(a, b) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32).reshape(16)
    b = np.ascontiguousarray(b, np.float32).reshape(16)
    out = np.empty(16, np.float32)
    _check(load_library().sloth_mat4_mul(_fptr(a), _fptr(b), _fptr(out)))
    return out


# ---------------------------------------------------------------------------
# batches
# ---------------------------------------------------------------------------

@dataclass
class FrameSpec:
    width: int
    height: int
    utransform: np.ndarray
    image: bool = True


@dataclass
class DrawSpec:
    frame: int
    mesh: SimpleMesh
    transform: np.ndarray


class Batch:
    """Many frames rendered by one launch sequence (captured in a HIP graph)."""

    def __init__(self, frames: list, draws: list, shader: Shader = default_shader):
        _ensure_init()
        L = load_library()
        fa = (sloth_frame_desc * len(frames))()
        for i, f in enumerate(frames):
            fa[i].width, fa[i].height, fa[i].image = f.width, f.height, int(bool(f.image))
            u = np.ascontiguousarray(f.utransform, np.float32).reshape(16)
            for k in range(16):
                fa[i].utransform[k] = float(u[k])
        da = (sloth_draw_desc * max(1, len(draws)))()
        for i, d in enumerate(draws):
            da[i].frame = d.frame
            da[i].mesh = d.mesh.handle.value
            t = np.ascontiguousarray(d.transform, np.float32).reshape(16)
            for k in range(16):
                da[i].transform[k] = float(t[k])
        r = shader.to_c()
        h = C.c_void_p()
        _check(L.sloth_batch_create(fa, len(frames), da, len(draws), C.byref(r), C.byref(h)))
        self._h = h
        self.frames = frames
        self.draws = draws
        self._keep = [d.mesh for d in draws]

    def set_transforms(self, transforms) -> None:
        """One column-major 4x4 per draw, in the order the draws were given;
        the next run renders with them."""
        t = np.ascontiguousarray(transforms, np.float32)
        if t.size != 16 * len(self.draws):
            raise ValueError(f"set_transforms needs {len(self.draws)} matrices of 16 floats, got {t.size} floats")
        t = t.reshape(-1, 16)
        _check(load_library().sloth_batch_set_transforms(self._h, _fptr(t)))

    def run(self, stream: int | None = None) -> None:
        _check(load_library().sloth_batch_run(self._h, C.c_void_p(stream or 0)))

    def run_per_frame(self, stream: int | None = None) -> None:
        _check(load_library().sloth_batch_run_per_frame(self._h, C.c_void_p(stream or 0)))

    def sync(self) -> None:
        _check(load_library().sloth_batch_sync(self._h))

    def read_frame(self, f: int, want_winners: bool = False):
        fr = self.frames[f]
        nc = fr.width * fr.height + (fr.height if fr.image else 0)
        cells = np.empty(nc, np.uint32)
        z = np.empty(fr.width * fr.height, np.float32)
        win = np.empty(fr.width * fr.height, np.int32) if want_winners else None
        _check(load_library().sloth_batch_read_frame(
            self._h, f, cells.ctypes.data_as(_u32p), _fptr(z),
            win.ctypes.data_as(_i32p) if want_winners else None))
        return cells, z, win

    def encode_frame(self, f: int, mode: int = ENCODE_TEXT) -> bytes:
        return _encode(load_library().sloth_batch_encode_frame, self._h, f, mode)

    def encode_export(self, first: int = 0, count: int | None = None, mode: int = ENCODE_HTML) -> bytes:
        """The `-j` export of frames [first, first + count) (main.rs:55-58,
        85-106), encoded on the GPU."""
        count = len(self.frames) - first if count is None else count
        return _encode(load_library().sloth_batch_encode_export, self._h, first, count, mode)

    def device_frame(self, f: int):
        c = C.c_void_p()
        z = C.c_void_p()
        _check(load_library().sloth_batch_device_frame(self._h, f, C.byref(c), C.byref(z)))
        return c.value, z.value

    def time(self, steps: int, warmup: int, per_frame: bool = False) -> float:
        """Milliseconds for `steps` renders after `warmup` untimed ones (HIP
        events on the batch's stream)."""
        ms = C.c_float()
        _check(load_library().sloth_batch_time(self._h, steps, warmup, int(per_frame), C.byref(ms)))
        return float(ms.value)

    def set_tile(self, tile_w: int, tile_h: int) -> None:
        _check(load_library().sloth_batch_set_tile(self._h, tile_w, tile_h))

    def set_walk(self, mode: int) -> None:
        """Tile raster variant: 1 big boxes walked by wavefronts, 0 per lane,
        -1 the sizing run's choice."""
        _check(load_library().sloth_batch_set_walk(self._h, mode))

    def autotune(self) -> tuple:
        """Times the candidate tile sizes and keeps the fastest; returns it."""
        w, h = C.c_uint32(), C.c_uint32()
        _check(load_library().sloth_batch_autotune(self._h, C.byref(w), C.byref(h)))
        return int(w.value), int(h.value)

    def plan(self) -> dict:
        """The raster plan in use: tile size in cells, the big-box walk and
        the binning kernel's triangles per thread."""
        w, h, bw, tpt = C.c_uint32(), C.c_uint32(), C.c_int32(), C.c_uint32()
        _check(load_library().sloth_batch_plan(self._h, C.byref(w), C.byref(h), C.byref(bw), C.byref(tpt)))
        return {"tile_w": int(w.value), "tile_h": int(h.value), "big_walk": int(bw.value),
                "bin_tris_per_thread": int(tpt.value)}

    def set_bin(self, tris_per_thread: int) -> None:
        """Triangles per thread of the binning kernel: 1, 2 or 4."""
        _check(load_library().sloth_batch_set_bin(self._h, tris_per_thread))

    def profile(self) -> dict:
        cap = 16
        ms = (C.c_float * cap)()
        names = (C.c_char_p * cap)()
        n = C.c_uint32(cap)
        _check(load_library().sloth_batch_profile(self._h, ms, names, C.byref(n)))
        return {names[i].decode(): float(ms[i]) for i in range(min(n.value, cap))}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.sloth_batch_destroy(h)
            self._h = None


class Shading:
    """Differentiable shading of a batch's visible fragments (per-vertex
    colours and normals, light direction; visibility fixed).  The forward and
    backward pair of the C-ABI (sloth_shading_forward / _backward)."""

    def __init__(self, batch: Batch, faces, n_verts: int):
        L = load_library()
        self.batch = batch
        self.faces = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
        self.n_verts = int(n_verts)
        h = C.c_void_p()
        _check(L.sloth_shading_create(batch._h, self.faces.ctypes.data_as(_u32p), self.faces.shape[0],
                                      self.n_verts, C.byref(h)))
        self._h = h
        n = C.c_uint64()
        _check(L.sloth_shading_slots(h, C.byref(n)))
        self.n_slots = int(n.value)

    def visibility(self) -> None:
        """Re-renders the visibility after Batch.set_transforms."""
        _check(load_library().sloth_shading_visibility(self._h))

    def set_target(self, target) -> None:
        t = None if target is None else np.ascontiguousarray(target, np.float32).reshape(self.n_slots, 3)
        _check(load_library().sloth_shading_set_target(self._h, _fptr(t) if t is not None else None))
        self._target = t

    def forward(self, vcolor, vnormal, light, want_out: bool = False):
        vc = np.ascontiguousarray(vcolor, np.float32).reshape(self.n_verts, 3)
        vn = np.ascontiguousarray(vnormal, np.float32).reshape(self.n_verts, 3)
        li = np.ascontiguousarray(light, np.float32).reshape(3)
        out = np.empty((self.n_slots, 3), np.float32) if want_out else None
        loss = C.c_double()
        _check(load_library().sloth_shading_forward(self._h, _fptr(vc), _fptr(vn), _fptr(li),
                                                    _fptr(out) if want_out else None, C.byref(loss)))
        return float(loss.value), out

    def backward(self):
        dc = np.empty((self.n_verts, 3), np.float32)
        dn = np.empty((self.n_verts, 3), np.float32)
        dl = np.empty(3, np.float32)
        _check(load_library().sloth_shading_backward(self._h, _fptr(dc), _fptr(dn), _fptr(dl)))
        return dc, dn, dl

    def step(self, vcolor, vnormal, light):
        """Loss and gradients in one pass: (loss, d_vcolor, d_vnormal, d_light)."""
        vc = np.ascontiguousarray(vcolor, np.float32).reshape(self.n_verts, 3)
        vn = np.ascontiguousarray(vnormal, np.float32).reshape(self.n_verts, 3)
        li = np.ascontiguousarray(light, np.float32).reshape(3)
        loss = C.c_double()
        dc = np.empty((self.n_verts, 3), np.float32)
        dn = np.empty((self.n_verts, 3), np.float32)
        dl = np.empty(3, np.float32)
        _check(load_library().sloth_shading_step(self._h, _fptr(vc), _fptr(vn), _fptr(li), C.byref(loss),
                                                 _fptr(dc), _fptr(dn), _fptr(dl)))
        return float(loss.value), dc, dn, dl

    def time(self, steps: int, warmup: int) -> float:
        """Milliseconds for `steps` training steps (loss and gradients)."""
        ms = C.c_float()
        _check(load_library().sloth_shading_time(self._h, steps, warmup, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.sloth_shading_destroy(h)
            self._h = None
