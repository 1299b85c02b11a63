"""Error returns of the C-ABI on a GPU: calls out of order, arguments out of
range and a context that cannot fit are refused with a code and a message,
before any kernel runs, and leave the device usable."""
import numpy as np
import pytest

import howmhd as H

pytestmark = pytest.mark.gpu


def _params(**kw):
    kw.setdefault("nt", 1)
    kw.setdefault("tend", 1.0)
    run = H.parse(H.make_input(**kw))
    return run, H.gpu_params(run)


def test_ssprk_out_of_order_and_bad_dt(require_gpu):
    run, p = _params(problem="orszag_tang", nx=16, ny=16, nz=1)
    q0 = H.init_problem(run)
    g = H.GpuSolver(p)
    try:
        g.upload_q0(q0)
        with pytest.raises(H.HowmhdError, match="init_faces or upload_faces first"):
            g.ssprk(1e-3)
        g.init_faces()
        for bad in (-1e-3, float("nan"), float("inf")):
            with pytest.raises(H.HowmhdError, match="finite and >= 0"):
                g.ssprk(bad)
        g.upload_q0(q0)  # a new q(0) needs its faces again
        with pytest.raises(H.HowmhdError, match="init_faces"):
            g.ssprk(1e-3)
        g.init_faces()
        g.ssprk(g.tstep())   # the context still steps
        assert np.all(np.isfinite(H.active(g.q0())))
    finally:
        g.close()


def test_indicator_mlp_needs_weights(require_gpu):
    run, p = _params(problem="orszag_tang", nx=16, ny=16, nz=1)
    p.indicator = H.INDICATOR["mlp"]
    g = H.GpuSolver(p)
    try:
        g.upload_q0(H.init_problem(run))
        g.init_faces()
        with pytest.raises(H.HowmhdError, match="needs weights"):
            g.ssprk(g.tstep())
    finally:
        g.close()


def test_context_too_large_for_the_device(require_gpu):
    """2048^3 needs several TB: create fails with a message, frees what it
    took, and a normal context works afterwards."""
    _, p = _params(problem="uniform", nx=8, ny=8, nz=8)
    p.nx = p.ny = p.nz = 2048
    with pytest.raises(H.HowmhdError, match="allocation failed"):
        H.GpuSolver(p)
    run, p2 = _params(problem="mhd_blast3d", nx=8, ny=8, nz=8)
    g = H.GpuSolver(p2)
    try:
        g.upload_q0(H.init_problem(run))
        g.init_faces()
        g.ssprk(g.tstep())
        assert np.all(np.isfinite(H.active(g.q0())))
    finally:
        g.close()
