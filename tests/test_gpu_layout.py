"""Output-buffer layout through the C-ABI (GPU).

* sloth_context_read writes exactly the frame's W*H (+H) cells and W*H
  depths: sizes whose cell count is odd (image mode with H odd, terminal mode
  with W and H odd) leave a sentinel after the caller's buffer untouched;
* sloth_context_device_buffers reports the frame's own cell count;
* every batch frame's cells start on a 256-byte boundary
  (sloth_batch_device_frame) and frames of sizes whose cell counts leave
  every residue modulo 64 still match the oracle.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from _helpers import assert_frame_equal, oracle_render, random_soup, say

pytestmark = pytest.mark.gpu

SENTINEL = 0xDEADBEEF


def _scene(sloth, W, H, seed):
    from sloth_amd import meshgen

    rng = np.random.default_rng(seed)
    U = meshgen.screen_transform(W, H, [np.array([1.0, 1.0, 1.0], np.float32)])
    xyz, rgb = random_soup(rng, 300, W, H, U, spread=1.3, size=0.1)
    R = meshgen.rotation_euler(*rng.uniform(0, 2 * math.pi, 3))
    return U, xyz, rgb, R


@pytest.mark.parametrize("W,H,image", [(40, 21, True), (39, 21, False), (38, 7, True), (2, 1, True), (1, 1, False)])
def test_context_read_writes_only_the_frame(sloth, W, H, image):
    U, xyz, rgb, R = _scene(sloth, W, H, W * 100 + H)
    say(f"case context read {W}x{H} image={image}")
    ctx = sloth.Context(W, H, image)
    ctx.utransform = U
    ctx.clear()
    mesh = sloth.SimpleMesh(xyz, rgb)
    ctx.draw(mesh, R)
    nc = W * H + (H if image else 0)
    cells = np.full(nc + 8, SENTINEL, np.uint32)
    z = np.full(W * H + 8, np.float32(-7.0), np.float32)
    win = np.full(W * H + 8, -7, np.int32)
    L = sloth.load_library()
    rc = L.sloth_context_read(ctx._h, cells.ctypes.data_as(C.POINTER(C.c_uint32)),
                              z.ctypes.data_as(C.POINTER(C.c_float)), win.ctypes.data_as(C.POINTER(C.c_int32)))
    assert rc == 0
    assert (cells[nc:] == SENTINEL).all(), cells[nc:]
    assert (z[W * H:] == np.float32(-7.0)).all()
    assert (win[W * H:] == -7).all()
    ref = oracle_render(W, H, U, [(xyz, rgb, R, None)], image=image)
    assert_frame_equal(cells[:nc], z[:W * H], win[:W * H], ref, W)

    d_cells, d_z = C.c_void_p(), C.c_void_p()
    n_cells, n_z = C.c_uint64(), C.c_uint64()
    assert L.sloth_context_device_buffers(ctx._h, C.byref(d_cells), C.byref(d_z), C.byref(n_cells),
                                          C.byref(n_z)) == 0
    assert (n_cells.value, n_z.value) == (nc, W * H)


def test_batch_frames_aligned_and_exact(sloth):
    """Frames whose cell counts (W*H + H) take many residues modulo 64; each
    frame's cells start 256-byte aligned and match the oracle."""
    sizes = [(64, 33), (66, 17), (30, 9), (126, 5), (2, 3), (100, 37), (64, 32), (18, 61)]
    frames, draws, plan, keep = [], [], [], []
    for f, (W, H) in enumerate(sizes):
        U, xyz, rgb, R = _scene(sloth, W, H, 7000 + f)
        m = sloth.SimpleMesh(xyz, rgb)
        keep.append(m)
        frames.append(sloth.FrameSpec(W, H, U, True))
        draws.append(sloth.DrawSpec(f, m, R))
        plan.append((W, H, U, xyz, rgb, R))
    assert len({(W * H + H) % 64 for W, H in sizes}) >= 6
    b = sloth.Batch(frames, draws)
    say("case batch of 8 frames with unaligned cell counts")
    b.run()
    b.sync()
    L = sloth.load_library()
    for f, (W, H, U, xyz, rgb, R) in enumerate(plan):
        d_cells, d_z = C.c_void_p(), C.c_void_p()
        assert L.sloth_batch_device_frame(b._h, f, C.byref(d_cells), C.byref(d_z)) == 0
        assert d_cells.value % 256 == 0, (f, hex(d_cells.value))
        cells, z, win = b.read_frame(f, want_winners=True)
        ref = oracle_render(W, H, U, [(xyz, rgb, R, None)])
        assert_frame_equal(cells, z, win, ref, W)
