"""Model-file ingestion (CPU: parsing only): the mesh queue of inputs.rs:93-129
built by sloth_model_load from OBJ files (tobj 3.2.2 rules, geometry.rs:83-138)
and STL files (stl_io 0.4.2 rules, geometry.rs:146-181).  Expected values are
written out by hand from the files in each test; cube.obj / cube.mtl /
cube.stl are copies of the reference's models/ files."""
from __future__ import annotations

import os
import struct

import numpy as np
import pytest

import sloth_amd

HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(HERE, "golden", "models")


def _load(path):
    mf = sloth_amd.ModelFile(path)
    return [mf.mesh_arrays(i) for i in range(mf.n_meshes)]


def test_obj_models_faces_and_colours(tmp_path):
    (tmp_path / "m.mtl").write_text(
        "newmtl half\nKd 0.5 1.0 0.0\n\nnewmtl odd\nKd 1.5 -0.2 nan\nnewmtl nokd\nNs 10\n")
    (tmp_path / "m.obj").write_text(
        "# comment\nmtllib m.mtl\n"
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 1.5 -2\nvt 0 0\nvn 0 0 1\n"
        "o first\nusemtl half\n"
        "f 1/1/1 2/1/1 3/1/1 4/1/1\n"            # quad -> (1,2,3), (1,3,4)
        "l 1 2\nf 1 2\n"                            # lines are ignored
        "usemtl half\nf -5 -4 -1\n"                 # same material: same model; negative ids
        "usemtl odd\nf 2//1 3//1 5//1\n"            # material change: new model
        "g grp\nusemtl nokd\nf 1 2 3 4 5\n")        # pentagon fan: (1,2,3) (1,3,4) (1,4,5)
    ms = _load(str(tmp_path / "m.obj"))
    assert len(ms) == 3
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 1.5, -2]], np.float32)
    pos, rgb, mn, mx = ms[0]
    assert np.array_equal(pos, v[[[0, 1, 2], [0, 2, 3], [0, 1, 4]]])
    assert (rgb == [127, 255, 0]).all()                     # 0.5 * 255 = 127.5 -> 127
    assert np.array_equal(mn, [0, 0, -2]) and np.array_equal(mx, [1, 1.5, 0])
    pos, rgb, mn, mx = ms[1]
    assert np.array_equal(pos, v[[[1, 2, 4]]]) and (rgb == [255, 0, 0]).all()   # saturating, NaN -> 0
    assert np.array_equal(mn, [0, 0, -2])                    # the box starts at the origin
    pos, rgb, _, _ = ms[2]
    assert np.array_equal(pos, v[[[0, 1, 2], [0, 2, 3], [0, 3, 4]]]) and (rgb == 0).all()   # no Kd: 0


def test_obj_without_materials_and_vertex_colours(tmp_path):
    (tmp_path / "a.obj").write_text("v 0 0 0 1 0 0\nv 2 0 0 0 1 0\nv 0 3 0 0 0 1\nf 1 2 3\nf 3 1 2\n")
    (pos, rgb, mn, mx), = _load(str(tmp_path / "a.obj"))
    assert pos.shape == (2, 3, 3) and (rgb == 1).all()           # no materials: (1,1,1), colours unused
    assert np.array_equal(mx, [2, 3, 0]) and np.array_equal(mn, [0, 0, 0])
    (tmp_path / "c.mtl").write_text("newmtl k\nKd 0.2 0.2 0.2\n")
    (tmp_path / "c.obj").write_text("mtllib c.mtl\nv 0 0 0 1 0 0\nv 2 0 0 0 1 0\nv 0 3 0 0 0.5 1\n"
                                    "usemtl k\nf 1 2 3\nf 3 1 2\n")
    (pos, rgb, _, _), = _load(str(tmp_path / "c.obj"))
    assert rgb.tolist() == [[255, 0, 0], [0, 127, 255]]           # the first corner's colour


def test_obj_errors(tmp_path):
    (tmp_path / "nolib.obj").write_text("mtllib missing.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    (tmp_path / "k.mtl").write_text("newmtl k\nKd 1 1 1\n")
    (tmp_path / "nomat.obj").write_text("mtllib k.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    (tmp_path / "range.obj").write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")
    (tmp_path / "late.obj").write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\no next\nv 0 1 0\nf 1 2 3\n")
    (tmp_path / "x.ply").write_text("")
    for name in ("nolib.obj", "nomat.obj", "range.obj", "late.obj", "x.ply", "absent.obj"):
        with pytest.raises(sloth_amd.SlothError):
            sloth_amd.ModelFile(str(tmp_path / name))


def test_obj_face_before_its_vertex(tmp_path):
    """Faces resolve when their model ends (tobj), so a face may name a vertex
    defined after it inside the same model."""
    (tmp_path / "f.obj").write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\nv 0 1 0\n")
    (pos, _, _, _), = _load(str(tmp_path / "f.obj"))
    assert np.array_equal(pos[0], np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32))


def test_cube_obj():
    ms = _load(os.path.join(MODELS, "cube.obj"))
    # six `o` objects of one quad each, in file order, Kd * 255 truncated
    expect = [[26, 30, 163], [36, 163, 36], [163, 21, 155], [163, 26, 23], [163, 160, 30], [31, 157, 163]]
    assert len(ms) == 6
    for (pos, rgb, mn, mx), col in zip(ms, expect):
        assert pos.shape == (2, 3, 3) and (rgb == col).all()
        assert np.array_equal(pos[0, 0], pos[1, 0])           # fan from the first corner
    assert np.array_equal(ms[0][0][0], np.array([[-0.511545, 0.570342, -1.023090],
                                                  [0.511545, -0.452748, -1.023090],
                                                  [-0.211889, -1.176181, -0.0]], np.float32))


def test_stl_ascii_and_binary(tmp_path):
    (pos, rgb, mn, mx), = _load(os.path.join(MODELS, "cube.stl"))
    assert pos.shape[0] == 12 and (rgb == [255, 255, 0]).all()
    assert np.array_equal(pos[0], np.array([[-9.62079, -2.56574, 34.9837], [32.6492, -15.3124, -4.75013],
                                            [24.5587, 17.1678, 20.6188]], np.float32))
    assert np.array_equal(mn, pos.reshape(-1, 3).min(0)) and np.array_equal(mx, pos.reshape(-1, 3).max(0))
    tris = np.array([[[-1, -2, -3], [-4, -5, -6], [-7, -8, -9.5]], [[-1, 0, -3], [-2, -2, -2], [-3, -1, -1]]],
                    np.float32)
    blob = b"binary header".ljust(80, b"\0") + struct.pack("<I", 2)
    for t in tris:
        blob += struct.pack("<3f", 0, 0, 1) + t.tobytes() + b"\0\0"
    (tmp_path / "b.STL").write_bytes(blob + b"trailing")
    (pos, rgb, mn, mx), = _load(str(tmp_path / "b.STL"))
    assert np.array_equal(pos, tris) and (rgb == [255, 255, 0]).all()
    assert np.array_equal(mx, [-1, 0, -1])                    # no origin in an STL box
    (tmp_path / "short.stl").write_bytes(blob[:120])
    with pytest.raises(sloth_amd.SlothError):
        sloth_amd.ModelFile(str(tmp_path / "short.stl"))


@pytest.mark.skipif(not os.path.exists("/root/reference/models/skull.obj"), reason="reference models not mounted")
def test_skull_matches_the_fixture_parser():
    """The skull fixture was parsed by an independent restatement of the same
    rules (tests/golden/make_fixtures.py); the loader must agree bit for bit."""
    ms = _load("/root/reference/models/skull.obj")
    d = np.load(os.path.join(HERE, "golden", "skull_datajs.npz"))
    assert [m[0].shape[0] for m in ms] == d["mesh_counts"].tolist()
    pos = np.concatenate([m[0] for m in ms])
    rgb = np.concatenate([m[1] for m in ms])
    assert np.array_equal(pos.view(np.uint32), d["xyz"].view(np.uint32)) and np.array_equal(rgb, d["rgb"])
