"""Parity of the HIP time step with the CPU restatement (needs a GPU).

Both libraries are driven through the C-ABI with the same q(0), the same dt
sequence (the restatement's), and compared on active cells:
  * the FP-contraction-off build must agree bit for bit (q, faces and dt);
  * the product build within relative L1 1e-12 per conserved variable.
Cases cover 1-D, 2-D and 3-D problems, periodic and open boundaries, the
stage-5 combination and the floors.
"""
import numpy as np
import pytest

import howmhd as H
import oracle_ctypes as O

pytestmark = pytest.mark.gpu

CASES = {
    "ot32": ("orszag_tang", dict(nx=32, ny=32, nz=1, cour=0.8), 6),
    "blast16": ("blast3d", dict(nx=16, ny=16, nz=16), 3),
    "smooth16": ("mpi_smooth3d", dict(nx=16, ny=12, nz=10), 2),
    "briowu": ("brio_wu", dict(nx=64, ny=1, nz=1, x1bc="open", x2bc="open", x3bc="open"), 8),
    "rotor": ("mhd_rotor", dict(nx=32, ny=32, nz=1, gam=1.4), 4),
    "mhdblast_open": ("mhd_blast3d", dict(nx=12, ny=14, nz=16, x1bc="open", x3bc="open"), 3),
    "cpaw": ("cpaw3d", dict(nx=24, ny=12, nz=12, xsize=3.0, ysize=1.5, zsize=1.5), 2),
    "xz2d": ("mpi_smooth3d", dict(nx=20, ny=1, nz=18), 3),
    "floors": ("blast3d", dict(nx=16, ny=16, nz=8, rhomin=0.9999, pgmin=5.0), 2),
}


def _setup(problem, kw):
    kw = dict(kw)
    kw.setdefault("nt", 1)
    kw.setdefault("tend", 1.0)
    run = H.parse(H.make_input(problem=problem, **kw))
    return run, H.init_problem(run), H.gpu_params(run)


def _bitwise_mismatch(a, b):
    a, b = H.active(a), H.active(b)
    return int(np.count_nonzero(a.view(np.int64) != b.view(np.int64)))


@pytest.mark.parametrize("case", list(CASES))
def test_exact_build_bitwise(case, require_gpu):
    problem, kw, nsteps = CASES[case]
    run, q0, p = _setup(problem, kw)
    print(f"[parity exact] case={case} shape=({p.nx},{p.ny},{p.nz}) steps={nsteps}", flush=True)
    ref = O.RefSolver(p)
    gpu = H.GpuSolver(p, exact=True)
    assert H.build_kind(exact=True) == "fp-exact"
    for s in (ref, gpu):
        s.upload_q0(q0)
        s.init_faces()
    assert _bitwise_mismatch(gpu.faces(), ref.faces()) == 0, "initial faces"
    for step in range(nsteps):
        dt = ref.tstep()
        assert gpu.tstep() == dt, f"dt differs at step {step}"
        ref.ssprk(dt)
        gpu.ssprk(dt)
    qr, qg = ref.q0(), gpu.q0()
    br, bg = ref.faces(), gpu.faces()
    nq, nb = _bitwise_mismatch(qg, qr), _bitwise_mismatch(bg, br)
    if nq or nb:
        l1 = H.rel_l1(qg, qr)
        print("rel L1 per variable:", l1)
    assert nq == 0, f"{nq} q values differ"
    assert nb == 0, f"{nb} face values differ"


@pytest.mark.parametrize("case", list(CASES))
def test_product_build_l1(case, require_gpu):
    problem, kw, nsteps = CASES[case]
    run, q0, p = _setup(problem, kw)
    print(f"[parity fast] case={case} shape=({p.nx},{p.ny},{p.nz}) steps={nsteps}", flush=True)
    ref = O.RefSolver(p)
    gpu = H.GpuSolver(p, exact=False)
    for s in (ref, gpu):
        s.upload_q0(q0)
        s.init_faces()
    for step in range(nsteps):
        dt = ref.tstep()
        dtg = gpu.tstep()
        assert abs(dtg - dt) <= 1e-13 * dt
        ref.ssprk(dt)
        gpu.ssprk(dt)
    l1 = H.rel_l1(gpu.q0(), ref.q0())
    print("rel L1 per variable:", l1)
    assert np.all(l1 <= 1e-12), l1


def test_orszag_tang_128_ten_steps(require_gpu):
    """Orszag-Tang 128x128x1, gam 5/3, cour 0.8: product build vs restatement
    after 10 steps, relative L1 <= 1e-12 for all 8 conserved variables."""
    run, q0, p = _setup("orszag_tang", dict(nx=128, ny=128, nz=1, cour=0.8, tend=0.5,
                                            gam="1.6666666666666667"))
    ref = O.RefSolver(p)
    gpu = H.GpuSolver(p)
    for s in (ref, gpu):
        s.upload_q0(q0)
        s.init_faces()
    for _ in range(10):
        dt = ref.tstep()
        ref.ssprk(dt)
        gpu.ssprk(dt)
    l1 = H.rel_l1(gpu.q0(), ref.q0())
    print("OT 128^2 rel L1 per variable after 10 steps:", l1)
    assert np.all(l1 <= 1e-12), l1
    d = gpu.diag()
    print("max |div B| (CT faces):", d.max_divb_ct)


def test_builds_stay_distinct_in_one_process(require_gpu):
    """Both GPU builds loaded into one process keep their own kernels: the
    product build (loaded first here) runs with FP contraction and differs
    from the restatement in the last bits, the FP-contraction-off build stays
    bit-for-bit equal to it."""
    run, q0, p = _setup("orszag_tang", dict(nx=32, ny=32, nz=1, cour=0.8))
    fast = H.GpuSolver(p)
    exact = H.GpuSolver(p, exact=True)
    ref = O.RefSolver(p)
    assert H.build_kind() == "fp-fast" and H.build_kind(exact=True) == "fp-exact"
    for s in (fast, exact, ref):
        s.upload_q0(q0)
        s.init_faces()
    for _ in range(2):
        dt = ref.tstep()
        for s in (fast, exact, ref):
            s.ssprk(dt)
    qr = ref.q0()
    assert _bitwise_mismatch(exact.q0(), qr) == 0
    assert _bitwise_mismatch(fast.q0(), qr) > 0
    assert np.all(H.rel_l1(fast.q0(), qr) <= 1e-12)
