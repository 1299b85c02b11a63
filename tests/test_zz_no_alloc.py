"""howmhd_gpu_tstep_noalloc, the CFL entry point that allocates no device
memory, against howmhd_gpu_tstep (needs a GPU).

Two calls in a row, the second on a smaller grid with smaller signal speeds:
each dt must equal, bit for bit, the dt that howmhd_gpu_tstep gives for the
same state (the shared scratch keeps the first call's per-block maxima, which
the second call must neither read nor need cleared); a third call repeats the
second on a stream of the caller's.  Bad arguments return HOWMHD_EINVAL before
any HIP call.
"""
import ctypes as C

import numpy as np
import pytest

import howmhd as H

pytestmark = pytest.mark.gpu


def _state(problem, steps, **kw):
    """(params, q(0) in the reference layout, dt of howmhd_gpu_tstep) after
    `steps` steps of the plain FP64 path."""
    run = H.parse(H.make_input(problem=problem, nt=1, tend=1.0, **kw))
    q0 = H.init_problem(run)
    p = H.gpu_params(run)
    g = H.GpuSolver(p)
    try:
        g.upload_q0(q0)
        g.init_faces()
        for _ in range(steps):
            g.ssprk(g.tstep())
        dt = g.tstep()      # from the maxima the last stage-5 combination left
        q = g.q0()
        g2 = H.GpuSolver(p)
        try:
            g2.upload_q0(q)
            assert g2.tstep() == dt   # the same state through the CFL kernel
        finally:
            g2.close()
    finally:
        g.close()
    return p, q, dt


def _max_speed(q, gam):
    """Largest |v_d| + c_fast,d over active cells and axes (NumPy)."""
    a = q[:, 3:-3, 3:-3, 3:-3]
    rho = a[0]
    v = a[1:4] / rho
    B2 = (a[4:7] ** 2).sum(0)
    pg = (gam - 1) * (a[7] - 0.5 * rho * (v ** 2).sum(0) - 0.5 * B2)
    cs2 = gam * pg / rho
    best = 0.0
    for d in range(3):
        bb, bn = B2 / rho, a[4 + d] ** 2 / rho
        cf = np.sqrt(0.5 * (bb + cs2 + np.sqrt(np.maximum((bb + cs2) ** 2 - 4 * bn * cs2, 0))))
        best = max(best, float((np.abs(v[d]) + cf).max()))
    return best


def test_tstep_noalloc_matches_tstep(require_gpu):
    import torch
    torch.cuda.set_device(0)
    # first call: 3-D blast (270 blocks of the reduction), second: 2-D
    # Orszag-Tang (4 blocks) whose speeds are below the blast's
    cases = [("mhd_blast3d", 2, dict(nx=48, ny=40, nz=36, cour=0.4)),
             ("orszag_tang", 1, dict(nx=32, ny=32, nz=1, cour=0.8))]
    speeds = []
    for problem, steps, kw in cases:
        p, q, dt = _state(problem, steps, **kw)
        speeds.append(_max_speed(q, p.gam))
        print(f"[noalloc] {problem} {p.nx}x{p.ny}x{p.nz}: tstep dt = {dt!r}", flush=True)
        qd = torch.from_numpy(q).cuda()
        got = H.tstep_noalloc(p, qd)
        torch.cuda.synchronize()
        assert got == dt, (problem, got, dt)
        del qd
    # the second grid's speeds are lower, so a stale maximum of the first
    # call would have shortened its dt
    assert speeds[1] < speeds[0], speeds
    # a third call on a stream of the caller's (the upload and the reduction
    # both queued there)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        qd = torch.from_numpy(q).cuda()
        got = H.tstep_noalloc(p, qd)
    side.synchronize()
    assert got == dt, ("side stream", got, dt)


def test_tstep_noalloc_rejects_bad_arguments(require_gpu):
    import torch
    lib = H.gpu_lib()
    p = H.Params()
    p.nx, p.ny, p.nz = 2, 8, 8        # n == 2 is not a grid howmhd_gpu_create accepts
    p.dx = p.dy = p.dz = 0.1
    p.gam, p.cour = 5.0 / 3.0, 0.4
    q = torch.zeros((8, 8, 14, 14), dtype=torch.float64, device="cuda")
    dt = C.c_double(-1.0)
    rc = lib.howmhd_gpu_tstep_noalloc(C.byref(p), C.c_void_p(q.data_ptr()), None, C.byref(dt))
    assert rc == 1 and dt.value == -1.0
    assert b"n == 1 or n >= 3" in lib.howmhd_gpu_last_error(None)
    p.nx = 4
    rc = lib.howmhd_gpu_tstep_noalloc(C.byref(p), None, None, C.byref(dt))
    assert rc == 1 and b"null argument" in lib.howmhd_gpu_last_error(None)
    p.nx, p.ny, p.nz = 1 << 14, 1 << 14, 1 << 13   # 2^41 cells
    rc = lib.howmhd_gpu_tstep_noalloc(C.byref(p), C.c_void_p(q.data_ptr()), None, C.byref(dt))
    assert rc == 1 and b"2^40" in lib.howmhd_gpu_last_error(None)
