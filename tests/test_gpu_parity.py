"""Parity of the HIP time step with the CPU restatement (needs a GPU).

The GPU library and the restatement are driven through their C-ABIs with the
same q(0) and the same dt sequence (the restatement's), and compared on
active cells:
  * precision fp64 (the plain method with the reference's rounding) must
    agree bit for bit: q, faces and dt;
  * precision fp64fast (FMA contraction, reciprocal forms) rounds
    differently.  On smooth flows it stays within relative L1 1e-12; where
    the flow crosses points at which the characteristic decomposition is
    discontinuous (B_t -> 0, sign of B_n at B_n = 0, af/as at degenerate
    wave speeds) one rounding difference changes the WENO flux there, and
    the form differs by up to ~1e-4 after 10 steps (blast problems).
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
    # every boundary open (the 3-D EMF kernel's open edges in x, y and z)
    "blast_open3": ("mhd_blast3d", dict(nx=14, ny=13, nz=12, x1bc="open", x2bc="open", x3bc="open"), 2),
    # two z chunks of the 3-D EMF kernel, the second one plane deep; ragged
    # x/y tiles
    "tall33": ("mpi_smooth3d", dict(nx=10, ny=7, nz=33), 2),
    # 3-cell-wide column: tiles of the 3-D EMF and face kernels mostly idle,
    # every x stencil wraps
    "thin3": ("mpi_smooth3d", dict(nx=3, ny=4, nz=40), 3),
    # three z planes: the z windows of the 3-D EMF and face kernels wrap
    # (periodic) or clamp (open) at both ends at once
    "flat3_per": ("mpi_smooth3d", dict(nx=12, ny=10, nz=3), 2),
    "flat3_open": ("mhd_blast3d", dict(nx=10, ny=9, nz=3, x3bc="open"), 2),
    # three x-tiles per row and 9 rows: the x sweep's XCD-aware tile order
    # on its first 8 rows, launch order on the ninth; open x: 131 interfaces
    "wide130": ("mpi_smooth3d", dict(nx=130, ny=9, nz=3), 2),
    "wide130_open": ("mhd_blast3d", dict(nx=130, ny=9, nz=3, x1bc="open"), 2),
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
def test_fp64_bitwise(case, require_gpu):
    problem, kw, nsteps = CASES[case]
    run, q0, p = _setup(problem, kw)
    print(f"[parity fp64] case={case} shape=({p.nx},{p.ny},{p.nz}) steps={nsteps}", flush=True)
    ref = O.RefSolver(p)
    gpu = H.GpuSolver(p)
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


@pytest.mark.parametrize("problem,kw", [
    ("mpi_smooth3d", dict(nx=1, ny=12, nz=10)),
    ("mhd_blast3d", dict(nx=1, ny=10, nz=12, x3bc="open")),
])
def test_fp64_bitwise_degenerate_x(problem, kw, require_gpu):
    """A grid of one cell along x.  In the reference the x sweep's store also
    forms the stage combination k1*q + k3*q_prev (src/ssprk.cpp:405-431, called
    at 1024), which undoes the stage-5 preparation's division by k1, and with
    nx = 1 the x sweep is not called (src/ssprk.cpp:1306-1308).  Those steps are
    skipped and the state blows up within one step (density at the floor,
    E ~ 1e12).  The GPU path reproduces that bit for bit over two steps."""
    run, q0, p = _setup(problem, kw)
    print(f"[parity fp64 nx=1] {problem} shape=({p.nx},{p.ny},{p.nz})", flush=True)
    ref = O.RefSolver(p)
    gpu = H.GpuSolver(p)
    for s in (ref, gpu):
        s.upload_q0(q0)
        s.init_faces()
    for step in range(2):
        dt = ref.tstep()
        assert gpu.tstep() == dt, f"dt differs at step {step}"
        ref.ssprk(dt)
        gpu.ssprk(dt)
    qr = ref.q0()
    assert np.all(np.isfinite(H.active(qr)))
    assert _bitwise_mismatch(gpu.q0(), qr) == 0
    assert _bitwise_mismatch(gpu.faces(), ref.faces()) == 0


@pytest.mark.parametrize("problem,kw", [
    ("mhd_blast3d", dict(nx=16, ny=16, nz=16)),
    ("orszag_tang", dict(nx=32, ny=32, nz=1, cour=0.8)),
])
def test_fp64_bitwise_large_weno_operands(problem, kw, require_gpu):
    """The same flow scaled by 1e50 (rho, momenta, E x 1e50, B x 1e25; the
    velocities and the dt sequence stay the same): some WENO weight operands
    t_k = (eps + IS_k)^2 exceed 1e200, where the reference-rounding sweep
    hands its block to the IEEE-division launch.  Still bit for bit."""
    run, q0, p = _setup(problem, kw)
    q0 = q0.copy()
    q0[[0, 1, 2, 3, 7]] *= 1e50
    q0[[4, 5, 6]] *= 1e25
    print(f"[parity fp64 large operands] {problem} {kw}", flush=True)
    ref = O.RefSolver(p)
    gpu = H.GpuSolver(p)
    for s in (ref, gpu):
        s.upload_q0(q0)
        s.init_faces()
    O.large_t_count(reset=True)
    for step in range(2):
        dt = ref.tstep()
        assert gpu.tstep() == dt, f"dt differs at step {step}"
        ref.ssprk(dt)
        gpu.ssprk(dt)
    nlarge = O.large_t_count()
    print("weight evaluations with t_k > 1e200:", nlarge)
    assert nlarge > 0
    qr, qg = ref.q0(), gpu.q0()
    assert np.all(np.isfinite(H.active(qr)))
    assert _bitwise_mismatch(qg, qr) == 0
    assert _bitwise_mismatch(gpu.faces(), ref.faces()) == 0


@pytest.mark.parametrize("problem,kw", [
    ("mhd_blast3d", dict(nx=16, ny=16, nz=16)),
    ("orszag_tang", dict(nx=32, ny=32, nz=1, cour=0.8)),
])
def test_fp64_bitwise_small_sqrt_operands(problem, kw, require_gpu):
    """The same flow with B x 1e-160: B^2 is subnormal, so square-root
    operands of the eigensystem (c_a^2 = B_n^2/rho and the like) fall below
    2^-767, where the correctly rounded square root scales its operand, and
    the Alfven and slow speeds and B_t normalisations are subnormal-derived.
    Still bit for bit."""
    run, q0, p = _setup(problem, kw)
    q0 = q0.copy()
    q0[[4, 5, 6]] *= 1e-160
    print(f"[parity fp64 small sqrt operands] {problem} {kw}", flush=True)
    ref = O.RefSolver(p)
    gpu = H.GpuSolver(p)
    for s in (ref, gpu):
        s.upload_q0(q0)
        s.init_faces()
    O.odd_sqrt_count(reset=True)
    for step in range(2):
        dt = ref.tstep()
        assert gpu.tstep() == dt, f"dt differs at step {step}"
        ref.ssprk(dt)
        gpu.ssprk(dt)
    nodd = O.odd_sqrt_count()
    print("square-root operands outside [2^-767, inf):", nodd)
    assert nodd > 0
    qr, qg = ref.q0(), gpu.q0()
    assert np.all(np.isfinite(H.active(qr)))
    assert _bitwise_mismatch(qg, qr) == 0
    assert _bitwise_mismatch(gpu.faces(), ref.faces()) == 0


def test_fp64_matches_regression_fixtures(require_gpu):
    """The plain FP64 GPU path against the committed regression fixtures of the
    restatement (tests/golden/oracle_regression.json): the same dt at every
    step and the same SHA-256 of q(0) and the faces, without running the
    restatement."""
    import hashlib
    import json
    from pathlib import Path
    gdir = Path(__file__).resolve().parent / "golden"
    want = json.loads((gdir / "oracle_regression.json").read_text())

    def digest(a):
        return hashlib.sha256(np.ascontiguousarray(H.active(a), dtype="<f8").tobytes()).hexdigest()

    for case in want["cases"]:
        print(f"[fixture] {case['problem']} {case['input']}", flush=True)
        run = H.parse(H.make_input(problem=case["problem"], **case["input"]))
        g = H.GpuSolver(H.gpu_params(run))
        try:
            g.upload_q0(H.init_problem(run))
            g.init_faces()
            for k in range(case["steps"]):
                dt = g.tstep()
                assert float(dt).hex() == case["dt"][k], (case["problem"], k)
                g.ssprk(dt)
            assert digest(g.q0()) == case["q_sha256"], case["problem"]
            assert digest(g.faces()) == case["faces_sha256"], case["problem"]
        finally:
            g.close()


@pytest.mark.parametrize("precision", ["fp64", "fp64fast", "fp32"])
def test_dt_from_fused_cfl_maxima(precision, require_gpu):
    """After a step, tstep takes the CFL maxima that the stage-5 combination
    computed while writing q(0); in every form dt is the reference's dt of
    that q(0) bit for bit (the restatement's tstep on the downloaded state)."""
    run, q0, p = _setup("mhd_blast3d", dict(nx=16, ny=12, nz=20, cour=0.4))
    p.precision = H.PREC[precision]
    gpu = H.GpuSolver(p)
    gpu.upload_q0(q0)
    gpu.init_faces()
    for step in range(3):
        dt = gpu.tstep()
        gpu.ssprk(dt)
        ref = O.RefSolver(p)
        ref.upload_q0(gpu.q0())
        assert gpu.tstep() == ref.tstep(), f"{precision}: dt differs after step {step}"


@pytest.mark.parametrize("rho_cell", [-0.25, 1e-300])
def test_fp64_bitwise_nonpositive_density_cell(rho_cell, require_gpu):
    """One cell of the MHD blast given density -0.25 (the sweep's shared
    reciprocal 1/rho is then negative, and its tile goes to the IEEE-division
    launch) or 1e-300, with floors rhomin = pgmin = 1e-3: still bit for bit."""
    run, q0, p = _setup("mhd_blast3d", dict(nx=16, ny=16, nz=16, rhomin=1e-3, pgmin=1e-3))
    q0 = q0.copy()
    H.active(q0)[0, 5, 6, 7] = rho_cell
    ref = O.RefSolver(p)
    gpu = H.GpuSolver(p)
    for s in (ref, gpu):
        s.upload_q0(q0)
        s.init_faces()
    for step in range(2):
        dt = ref.tstep()
        assert gpu.tstep() == dt, f"dt differs at step {step}"
        ref.ssprk(dt)
        gpu.ssprk(dt)
    assert np.all(np.isfinite(H.active(ref.q0())))
    assert _bitwise_mismatch(gpu.q0(), ref.q0()) == 0
    assert _bitwise_mismatch(gpu.faces(), ref.faces()) == 0


def test_fp64_bitwise_bench_workload(require_gpu):
    """The benchmark workload itself, MHD blast at 256^3: one full step
    (tstep + SSPRK(5,4)) of the GPU path equals the restatement bit for bit
    (q, faces and dt)."""
    run, q0, p = _setup("mhd_blast3d", dict(nx=256, ny=256, nz=256, cour=0.4))
    print("[parity fp64] mhd_blast3d 256^3, 1 step", flush=True)
    ref = O.RefSolver(p)
    gpu = H.GpuSolver(p)
    for s in (ref, gpu):
        s.upload_q0(q0)
        s.init_faces()
    del q0
    dt = ref.tstep()
    assert gpu.tstep() == dt
    gpu.ssprk(dt)
    ref.ssprk(dt)
    assert _bitwise_mismatch(gpu.q0(), ref.q0()) == 0
    assert _bitwise_mismatch(gpu.faces(), ref.faces()) == 0
    assert gpu.tstep() == ref.tstep()


SMOOTH = {"smooth16", "cpaw", "mhdblast_open"}


@pytest.mark.parametrize("case", list(CASES))
def test_fp64fast_l1(case, require_gpu):
    problem, kw, nsteps = CASES[case]
    run, q0, p = _setup(problem, kw)
    p.precision = H.PREC["fp64fast"]
    print(f"[parity fp64fast] case={case} shape=({p.nx},{p.ny},{p.nz}) steps={nsteps}", flush=True)
    ref = O.RefSolver(p)
    gpu = H.GpuSolver(p)
    for s in (ref, gpu):
        s.upload_q0(q0)
        s.init_faces()
    for step in range(nsteps):
        dt = ref.tstep()
        # the same CFL kernel, on this form's own (slightly different) state
        assert abs(gpu.tstep() - dt) <= 1e-6 * dt
        ref.ssprk(dt)
        gpu.ssprk(dt)
    l1 = H.rel_l1(gpu.q0(), ref.q0())
    print("rel L1 per variable:", l1)
    assert np.all(l1 <= (1e-12 if case in SMOOTH else 1e-4)), l1


def test_orszag_tang_128_ten_steps(require_gpu):
    """Orszag-Tang 128x128x1, gam 5/3, cour 0.8: plain FP64 form vs the
    restatement after 10 steps: relative L1 <= 1e-12 for all 8 conserved
    variables (bit for bit in fact)."""
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


def test_mhd_blast3d_ten_steps(require_gpu):
    """MHD blast 40x36x40 (periodic; ragged against the 64x8 EMF and 32x8 face
    tiles, two z chunks of the marching kernels): plain FP64 form vs the
    restatement over 10 steps, each taking its own dt: the dt sequences are
    equal, the relative L1 per conserved variable is <= 1e-12, and q and the
    faces are in fact equal bit for bit."""
    run, q0, p = _setup("mhd_blast3d", dict(nx=40, ny=36, nz=40))
    print(f"[parity fp64 10 steps] mhd_blast3d shape=({p.nx},{p.ny},{p.nz})", flush=True)
    ref = O.RefSolver(p)
    gpu = H.GpuSolver(p)
    for s in (ref, gpu):
        s.upload_q0(q0)
        s.init_faces()
    for step in range(10):
        dt = ref.tstep()
        assert gpu.tstep() == dt, f"dt at step {step}"
        ref.ssprk(dt)
        gpu.ssprk(dt)
    l1 = H.rel_l1(gpu.q0(), ref.q0())
    print("MHD blast 40x36x40 rel L1 per variable after 10 steps:", l1)
    assert np.all(l1 <= 1e-12), l1
    assert _bitwise_mismatch(gpu.q0(), ref.q0()) == 0
    assert _bitwise_mismatch(gpu.faces(), ref.faces()) == 0


def test_forms_stay_distinct_in_one_library(require_gpu):
    """The three forms live in one library: fp64 is the restatement bit for
    bit, fp64fast and fp32 run their own kernels (they differ in the last bits
    and at the FP32 level respectively), whatever form ran first."""
    run, q0, p = _setup("orszag_tang", dict(nx=32, ny=32, nz=1, cour=0.8))
    solvers = {}
    for form in ("fp64fast", "fp32", "fp64"):
        pf = H.gpu_params(run)
        pf.precision = H.PREC[form]
        solvers[form] = H.GpuSolver(pf)
    ref = O.RefSolver(p)
    for s in list(solvers.values()) + [ref]:
        s.upload_q0(q0)
        s.init_faces()
    for _ in range(2):
        dt = ref.tstep()
        for s in list(solvers.values()) + [ref]:
            s.ssprk(dt)
    qr = ref.q0()
    assert _bitwise_mismatch(solvers["fp64"].q0(), qr) == 0
    l_fast = H.rel_l1(solvers["fp64fast"].q0(), qr)
    l_32 = H.rel_l1(solvers["fp32"].q0(), qr)
    print("fp64fast rel L1", l_fast, "\nfp32 rel L1", l_32)
    assert _bitwise_mismatch(solvers["fp64fast"].q0(), qr) > 0 and l_fast.max() <= 1e-5
    # density and energy (Mz and Bz stay ~0 here, so their relative L1 is not meaningful)
    assert l_fast[0] < l_32[0] <= 1e-3 and l_32[7] <= 1e-3


def _random_state(shape, seed, bc_open):
    """Cell-by-cell random states (rho ~ U[0.5, 2], p ~ U[0.05, 2], v and B
    ~ U[-1, 1]^3, independent per cell): every interface has a jump, so the
    WENO weights leave their ideal values everywhere and the eigensystem
    meets every ordering of the wave speeds and both signs of B_n and B_t."""
    rng = np.random.default_rng(seed)
    n6 = tuple(n + 6 for n in shape)
    rho = rng.uniform(0.5, 2.0, n6)
    pg = rng.uniform(0.05, 2.0, n6)
    v = rng.uniform(-1.0, 1.0, (3,) + n6)
    B = rng.uniform(-1.0, 1.0, (3,) + n6)
    if seed % 2:       # weak transverse fields and B_n = 0 planes: the degenerate branches
        B[1:] *= 1e-9
        B[0][::3] = 0.0
    q = np.zeros((8,) + n6)
    q[0] = rho
    q[1:4] = rho * v
    q[4:7] = B
    q[7] = pg / (5.0 / 3.0 - 1.0) + 0.5 * rho * (v * v).sum(0) + 0.5 * (B * B).sum(0)
    return q


@pytest.mark.parametrize("seed", range(4))
def test_fp64_bitwise_random_states(seed, require_gpu):
    """Random cell states on a ragged 3-D grid (13 x 11 x 9; periodic for even
    seeds, open on every axis for odd ones), two steps: bit for bit."""
    shape = (13, 11, 9)
    bc_open = seed % 2 == 1
    bcs = dict(x1bc="open", x2bc="open", x3bc="open") if bc_open else {}
    run, _, p = _setup("uniform", dict(nx=shape[0], ny=shape[1], nz=shape[2], cour=0.4,
                                      gam="1.6666666666666667", **bcs))
    q0 = _random_state(shape, seed, bc_open)
    print(f"[parity fp64 random] seed={seed} shape={shape} open={bc_open}", flush=True)
    ref = O.RefSolver(p)
    gpu = H.GpuSolver(p)
    for s in (ref, gpu):
        s.upload_q0(q0)
        s.init_faces()
    for step in range(2):
        dt = ref.tstep()
        assert gpu.tstep() == dt, f"dt differs at step {step}"
        ref.ssprk(dt)
        gpu.ssprk(dt)
    qr = ref.q0()
    assert np.all(np.isfinite(H.active(qr)))
    assert _bitwise_mismatch(gpu.q0(), qr) == 0
    assert _bitwise_mismatch(gpu.faces(), ref.faces()) == 0


@pytest.mark.parametrize("march", [2, 3, 8])
@pytest.mark.parametrize("bc_open", [False, True])
def test_fp64_bitwise_marching_sweeps(march, bc_open, monkeypatch, require_gpu):
    """The y and z sweeps with blocks marching through `march` consecutive
    tiles (HOWMHD_SWEEP_MARCH; large grids choose it from the launch size):
    random states on 12 x 40 x 33 (5 or 6 tiles along y and z, so the last
    block of a march of 2 or 3 holds fewer tiles and a march of 8 is cut to
    the tiles there are), two steps, bit for bit against the restatement."""
    shape = (12, 40, 33)
    bcs = dict(x1bc="open", x2bc="open", x3bc="open") if bc_open else {}
    run, _, p = _setup("uniform", dict(nx=shape[0], ny=shape[1], nz=shape[2], cour=0.4,
                                      gam="1.6666666666666667", **bcs))
    q0 = _random_state(shape, 5 if bc_open else 4, bc_open)
    print(f"[parity fp64 march] march={march} shape={shape} open={bc_open}", flush=True)
    monkeypatch.setenv("HOWMHD_SWEEP_MARCH", str(march))
    ref = O.RefSolver(p)
    gpu = H.GpuSolver(p)
    for s in (ref, gpu):
        s.upload_q0(q0)
        s.init_faces()
    for step in range(2):
        dt = ref.tstep()
        assert gpu.tstep() == dt, f"dt differs at step {step}"
        ref.ssprk(dt)
        gpu.ssprk(dt)
    assert _bitwise_mismatch(gpu.q0(), ref.q0()) == 0
    assert _bitwise_mismatch(gpu.faces(), ref.faces()) == 0


def test_fp64fast_marching_sweeps_change_no_bit(monkeypatch, require_gpu):
    """fp64fast with marching y/z sweeps (3 tiles per block) equals the same
    form with single tiles bit for bit: the march only moves cells within LDS."""
    run, q0, p = _setup("mhd_blast3d", dict(nx=16, ny=40, nz=33, x3bc="open", cour=0.4))
    p.precision = H.PREC["fp64fast"]
    print("[fp64fast march] 16 x 40 x 33", flush=True)
    solvers = []
    for m in ("3", "1"):
        monkeypatch.setenv("HOWMHD_SWEEP_MARCH", m)
        solvers.append(H.GpuSolver(p))
    for s in solvers:
        s.upload_q0(q0)
        s.init_faces()
    for _ in range(2):
        dt = solvers[1].tstep()
        for s in solvers:
            s.ssprk(dt)
    assert _bitwise_mismatch(solvers[0].q0(), solvers[1].q0()) == 0
    assert _bitwise_mismatch(solvers[0].faces(), solvers[1].faces()) == 0
