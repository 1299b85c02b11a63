"""Host transfers of the C-ABI (needs a GPU): the reference layout with its
three ghost layers per side <-> the device's active-cell layout.

upload_q0 / upload_faces take the active cells of the caller's array;
q0() / faces() return every cell, the ghost layers filled with the active
cells that the boundary conditions map them to (src/bound.cpp:19-95:
periodic wraps, open copies the edge cell; an axis of one cell maps every
ghost to it).  The expected arrays are built here with numpy index maps and
compared bit for bit.  Shapes cover ragged tiles (sizes that are not
multiples of the 32-cell transpose tile), 2-D and 1-D grids, and a grid with
more than 65535 cells along x (ny and nz are limited to 65535 by create: they
go into the launch grids' y and z dimensions).
"""
import ctypes as C

import numpy as np
import pytest

import howmhd as H

pytestmark = pytest.mark.gpu

PER, OPEN = 0, 1


def _params(nx, ny, nz, bc):
    p = H.Params()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.bc = (C.c_int32 * 3)(*bc)
    p.dx = p.dy = p.dz = 1.0
    p.gam, p.cour, p.rhomin, p.pgmin = 5.0 / 3.0, 0.4, 1e-10, 1e-10
    p.precision = H.PREC["fp64"]
    p.indicator = 0
    return p


def _ghost_map(n, bc):
    a = np.arange(-3, n + 3)
    if n == 1:
        return np.zeros_like(a)
    if bc == PER:
        return np.mod(a, n)
    return np.clip(a, 0, n - 1)


def _expected(arr, shape, bc):
    act = arr[:, 3:-3, 3:-3, 3:-3]
    mx, my, mz = (_ghost_map(shape[d], bc[d]) for d in range(3))
    return act[:, mx][:, :, my][:, :, :, mz]


SHAPES = [
    ((37, 5, 70), (PER, OPEN, PER)),
    ((40, 33, 1), (OPEN, PER, PER)),
    ((64, 1, 1), (OPEN, OPEN, OPEN)),
    ((65540, 3, 1), (PER, OPEN, PER)),
]


@pytest.mark.parametrize("shape,bc", SHAPES, ids=[f"{s[0]}x{s[1]}x{s[2]}" for s, _ in SHAPES])
def test_transfer_round_trip(shape, bc, require_gpu):
    print(f"[transfer] shape={shape} bc={bc}", flush=True)
    rng = np.random.default_rng(sum(shape))
    g = H.GpuSolver(_params(*shape, bc))
    try:
        full = tuple(n + 6 for n in shape)
        q = rng.standard_normal((8,) + full)
        b = rng.standard_normal((3,) + full)
        g.upload_q0(q)
        g.upload_faces(b)
        qd, bd = g.q0(), g.faces()
    finally:
        g.close()
    for name, got, src in (("q", qd, q), ("faces", bd, b)):
        want = _expected(src, shape, bc)
        bad = np.count_nonzero(got.view(np.int64) != want.view(np.int64))
        assert bad == 0, f"{name}: {bad} values differ"


def test_upload_right_after_create_256(require_gpu):
    """A 256^3 context (12 GB of buffers cleared at create) takes an upload at
    once and returns it bit for bit.  The clears run on the context's stream;
    when they ran on the null stream, which the context's non-blocking stream
    does not wait for, the last clears (the transfer staging buffer among
    them) could still run during the first upload and zero part of q(0)."""
    shape, bc = (256, 256, 256), (PER, PER, PER)
    print(f"[transfer] shape={shape} bc={bc}, upload at once after create", flush=True)
    rng = np.random.default_rng(7)
    full = tuple(n + 6 for n in shape)
    q = 1.0 + rng.random((8,) + full)
    for trial in range(2):
        g = H.GpuSolver(_params(*shape, bc))
        try:
            g.upload_q0(q)
            qd = g.q0()
        finally:
            g.close()
        bad = np.count_nonzero(H.active(qd).view(np.int64) != H.active(q).view(np.int64))
        assert bad == 0, f"trial {trial}: {bad} values differ"
