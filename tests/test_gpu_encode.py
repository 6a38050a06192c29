"""Context::flush byte streams (context.rs:50-92) encoded on the GPU, against
the oracle's restatement, byte for byte."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from _helpers import say

pytestmark = pytest.mark.gpu


def test_encode_sphere_frame(sloth):
    from sloth_amd import meshgen

    m, W, H, U, R = meshgen.workload_sphere_160x48()
    ctx = sloth.Context(W, H, image=True)
    ctx.utransform = U
    ctx.clear()
    sloth.draw_mesh(ctx, sloth.SimpleMesh.from_mesh(m), R)
    ref = oracle.render(W, H, U, m.xyz, m.rgb, R)
    say("case encode sphere 160x48")
    assert ctx.encode(sloth.ENCODE_TEXT) == oracle.encode(ref.cells, 0)
    assert ctx.encode(sloth.ENCODE_ANSI) == oracle.encode(ref.cells, 1)
    assert ctx.encode(sloth.ENCODE_HTML) == oracle.encode(ref.cells, 2)
    text = ctx.encode(sloth.ENCODE_TEXT).decode()
    # newline cells (rows inside some triangle's domain) plus println!'s newline
    assert text.count("\n") == int((ref.cells == oracle.CELL_NEWLINE).sum()) + 1 and "@" in text


def test_encode_non_ascii_glyphs(sloth):
    """Glyph bytes >= 0x80 are chars U+0080..U+00FF, two UTF-8 bytes each."""
    from sloth_amd import meshgen, Shader

    m = meshgen.uv_sphere(20, 10, seed=2)
    W, H = 90, 30
    U = meshgen.screen_transform(W, H, [m.bbox_max()])
    R = meshgen.rotation_euler(0.0, 2.5, 0.0)
    sh = Shader((0.25, 0.5, 0.75, 1.0), "·°¤é", "ÿ")
    ctx = sloth.Context(W, H, image=True)
    ctx.utransform = U
    ctx.clear()
    sloth.draw_mesh(ctx, sloth.SimpleMesh.from_mesh(m), R, sh)
    cells, _, _ = ctx.read()
    say("case encode non-ascii")
    for mode in (sloth.ENCODE_TEXT, sloth.ENCODE_ANSI, sloth.ENCODE_HTML):
        got = ctx.encode(mode)
        assert got == oracle.encode(cells, mode)
        got.decode("utf-8")


def test_encode_1080p_batch_frame(sloth):
    from sloth_amd import meshgen

    m, W, H, U, Rs = meshgen.workload_sweep_1080p(4, level=5)
    sm = sloth.SimpleMesh.from_mesh(m)
    b = sloth.Batch([sloth.FrameSpec(W, H, U, True) for _ in Rs], [sloth.DrawSpec(k, sm, R) for k, R in enumerate(Rs)])
    b.run()
    b.sync()
    say("case encode 1080p batch frames")
    for f in (0, 3):
        cells, _, _ = b.read_frame(f)
        assert b.encode_frame(f, sloth.ENCODE_TEXT) == oracle.encode(cells, 0)
        assert b.encode_frame(f, sloth.ENCODE_ANSI) == oracle.encode(cells, 1)
        assert b.encode_frame(f, sloth.ENCODE_HTML) == oracle.encode(cells, 2)


def test_encode_terminal_mode_and_odd_cell_counts(sloth):
    """A terminal-mode context starts with the cursor move; frames whose cell
    count is odd (even W, odd H in image mode) encode exactly their own cells
    (the device buffer is padded to an even count)."""
    from sloth_amd import meshgen

    m = meshgen.uv_sphere(16, 8, seed=4)
    R = meshgen.rotation_euler(0.0, 0.7, 0.0)
    for W, H, image in ((80, 25, True), (80, 25, False), (64, 21, True), (33, 17, False)):
        U = meshgen.screen_transform(W, H, [m.bbox_max()])
        ctx = sloth.Context(W, H, image=image)
        ctx.utransform = U
        ctx.clear()
        sloth.draw_mesh(ctx, sloth.SimpleMesh.from_mesh(m), R)
        cells, _, _ = ctx.read()
        say(f"case encode {W}x{H} image={image}")
        assert cells.size == W * H + (H if image else 0)
        for mode in (sloth.ENCODE_TEXT, sloth.ENCODE_ANSI, sloth.ENCODE_HTML):
            got = ctx.encode(mode)
            assert got == oracle.encode(cells, mode, image=image)
            assert got.startswith(b"\x1b[1;1H") == (not image)


@pytest.mark.parametrize("n_frames,start,mode", [(20, 0.0, 2), (20, 2.5, 2), (7, 0.4, 0)])
def test_export_framing(sloth, n_frames, start, mode):
    """The `-j N` export of the reference's skull.obj at 100x100 (main.rs:55-58,
    85-106): the frames of the export schedule in one batch, encoded with the
    framing around each frame's flush, against the oracle's frames framed the
    same way.  A start pitch of 2.5 ends the export early (pitch > 9.42477);
    mode 0 is the --no-color export (plain flush, newline included)."""
    import os

    from sloth_amd import meshgen

    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "skull_datajs.npz"))
    xyz, rgb, counts = d["xyz"], d["rgb"], d["mesh_counts"]
    offs = np.concatenate([[0], np.cumsum(counts)])
    W, H = 100, 100
    U = meshgen.screen_transform(W, H, [oracle.mesh_bbox(xyz[offs[i]:offs[i + 1]])[1] for i in range(len(counts))])
    meshes = [sloth.SimpleMesh(xyz[offs[i]:offs[i + 1]], rgb[offs[i]:offs[i + 1]]) for i in range(len(counts))]
    pitches = meshgen.export_pitches(n_frames, start)
    frames = [sloth.FrameSpec(W, H, U, True) for _ in pitches]
    draws = [sloth.DrawSpec(f, m, sloth.rotation_euler(0.0, float(p), 0.0))
             for f, p in enumerate(pitches) for m in meshes]
    b = sloth.Batch(frames, draws)
    b.run()
    b.sync()
    say(f"case export framing: {len(pitches)} of {n_frames} frames, start {start}, mode {mode}")
    got = b.encode_export(mode=mode)
    want = [b"let frames = [\n"]
    for k, p in enumerate(pitches):
        f = oracle.Frame(W, H, True)
        R = oracle.rotation_euler(0.0, float(p), 0.0)
        for i in range(len(counts)):
            f.draw_mesh(U, xyz[offs[i]:offs[i + 1]], rgb[offs[i]:offs[i + 1]], R)
        want += [b"`\n", oracle.encode(f.cells, mode, image=True), b"`,\n" if k + 1 < len(pitches) else b"`];\n"]
    assert got == b"".join(want)
    if start == 2.5:
        assert len(pitches) < n_frames
