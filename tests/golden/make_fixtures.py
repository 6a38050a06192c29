"""Writes the golden fixtures under tests/golden/ from the reference's own data
files.  Run in a container where the reference tree is mounted; the GPU tests
read only the .npz files this script writes.

skull_datajs.npz
    Input: models/skull.obj + skull.mtl parsed with tobj 3.2.2's
    GPU_LOAD_OPTIONS semantics (inputs.rs:108): one model per `o`/`g`/`usemtl`
    group, polygons fan-triangulated from their first vertex, triangle colour
    = material Kd * 255 truncated to u8 (geometry.rs:109-116).
    Expected output: src-webify/data.js, the `-j 100 -w 100` export of that
    file (100 frames of 100x100 image mode, 10,100 cells each).  Its <span>
    colour tags are emitted only on a colour change, so a cell's colour is the
    last tag before it.  Stored as glyph bytes plus a colour-palette index.
"""
from __future__ import annotations

import os
import re
import sys

import numpy as np

REF = os.environ.get("SLOTH_REFERENCE", "/root/reference")
OUT = os.path.dirname(os.path.abspath(__file__))


def parse_mtl(path):
    mats = {}
    cur = None
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "newmtl":
                cur = p[1]
                mats[cur] = None
            elif p[0] == "Kd" and cur is not None:
                mats[cur] = np.array([float(x) for x in p[1:4]], np.float32)
    return mats


def parse_obj(path):
    """Returns a list of models: (name, material name, [ (i,j,k) vertex ids ])
    and the position array.  Fan triangulation (a, b, c), (a, c, d), ..."""
    pos = []
    models = []
    name, mtl, faces = "unnamed", None, []
    mtllib = None

    def push():
        nonlocal faces
        if faces:
            models.append((name, mtl, faces))
            faces = []

    with open(path) as f:
        for line in f:
            p = line.split()
            if not p or p[0].startswith("#"):
                continue
            if p[0] == "v":
                pos.append([float(x) for x in p[1:4]])
            elif p[0] == "f":
                ids = []
                for tok in p[1:]:
                    i = int(tok.split("/")[0])
                    ids.append(i - 1 if i > 0 else len(pos) + i)
                for k in range(1, len(ids) - 1):
                    faces.append((ids[0], ids[k], ids[k + 1]))
            elif p[0] in ("o", "g"):
                push()
                name = " ".join(p[1:])
            elif p[0] == "usemtl":
                push()
                mtl = p[1]
            elif p[0] == "mtllib":
                mtllib = p[1]
    push()
    return np.array(pos, np.float32), models, mtllib


def skull_meshes():
    obj = os.path.join(REF, "models", "skull.obj")
    pos, models, mtllib = parse_obj(obj)
    mats = parse_mtl(os.path.join(REF, "models", mtllib))
    xyz, rgb, counts = [], [], []
    for _, mtl, faces in models:
        f = np.array(faces, np.int64)
        kd = mats[mtl]
        col = (kd * np.float32(255.0)).astype(np.float32)
        col = np.array([int(c) for c in col], np.uint8)  # `as u8` truncates
        xyz.append(pos[f])
        rgb.append(np.repeat(col[None, :], f.shape[0], 0))
        counts.append(f.shape[0])
    return (np.ascontiguousarray(np.concatenate(xyz), np.float32),
            np.ascontiguousarray(np.concatenate(rgb), np.uint8), np.array(counts, np.int32))


SPAN = re.compile(r'<span style="color:rgb\((\d+),(\d+),(\d+)\)">')


def decode_datajs():
    with open(os.path.join(REF, "src-webify", "data.js")) as f:
        s = f.read()
    assert s.startswith("let frames = [\n")
    body = s[len("let frames = [\n"):]
    chunks = body.split("`")
    # chunks: '', frame0, ',\n', frame1, ..., frameN, '];\n'
    frames = chunks[1::2]
    glyphs, colors = [], []
    for fr in frames:
        assert fr.startswith("\n")
        fr = fr[1:]           # println!("`") newline before the cells
        g, c = [], []
        cur = (0, 0, 0)
        i = 0
        while i < len(fr):
            m = SPAN.match(fr, i)
            if m:
                cur = tuple(int(x) for x in m.groups())
                i = m.end()
                continue
            g.append(ord(fr[i]))
            c.append(cur)
            i += 1
        glyphs.append(g)
        colors.append(c)
    return glyphs, colors


def main():
    xyz, rgb, counts = skull_meshes()
    glyphs, colors = decode_datajs()
    nf = len(glyphs)
    ncell = len(glyphs[0])
    assert all(len(g) == ncell for g in glyphs), "ragged frames"
    palette = sorted({c for fr in colors for c in fr})
    pidx = {c: i for i, c in enumerate(palette)}
    G = np.array(glyphs, np.uint8)
    P = np.array([[pidx[c] for c in fr] for fr in colors], np.uint8)
    np.savez_compressed(os.path.join(OUT, "skull_datajs.npz"), xyz=xyz, rgb=rgb,
                        mesh_counts=counts, glyphs=G, color_index=P,
                        palette=np.array(palette, np.uint8), width=np.int32(100),
                        height=np.int32(100))
    print(f"skull: {xyz.shape[0]} triangles in {len(counts)} meshes {counts.tolist()}; "
          f"data.js: {nf} frames x {ncell} cells; palette {palette}")


if __name__ == "__main__":
    sys.exit(main())
