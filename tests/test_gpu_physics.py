"""Physics-level checks of the HIP time step (needs a GPU).

  * FP32 precision form against the FP64 restatement: density relative L1
    <= 1e-3 after 10 steps, CT divergence unchanged to <= 1e-10;
  * conservation of mass, momentum and energy and of the CT face divergence
    by the product build on a periodic 3-D problem, and at 512^3 (the
    largest benchmark size) through size-independent properties;
  * convergence order on the circularly polarised Alfven wave equals the
    restatement's (about 2nd order: the reference's CT flux terms use a
    two-point average);
  * the `how_mhd` driver reproduces the reference main loop (dt clipping,
    output cadence, VTS files) and its final state matches the restatement.
"""
import math
import re
import struct
import subprocess

import numpy as np
import pytest

import howmhd as H
import oracle_ctypes as O

pytestmark = pytest.mark.gpu


def _setup(problem, **kw):
    kw.setdefault("nt", 1)
    kw.setdefault("tend", 1.0)
    run = H.parse(H.make_input(problem=problem, **kw))
    return run, H.init_problem(run), H.gpu_params(run)


def _d4_div(b, p):
    aa, bb = 9.0 / 8.0, -1.0 / 24.0
    act = b[:, 3:-3, 3:-3, 3:-3]
    div = 0.0
    for d, h in enumerate((p.dx, p.dy, p.dz)):
        f = act[d]
        if f.shape[d] == 1:
            continue
        div = div + (aa * (f - np.roll(f, 1, d)) + bb * (np.roll(f, -1, d) - np.roll(f, 2, d))) / h
    return div


@pytest.mark.parametrize("problem,kw", [
    ("mhd_blast3d", dict(nx=32, ny=32, nz=32, cour=0.4)),
    ("turbulence3d", dict(nx=32, ny=32, nz=32, cour=0.4, seed=3)),
])
def test_fp32_form_error_against_fp64(problem, kw, require_gpu):
    run, q0, p64 = _setup(problem, **kw)
    print(f"[fp32 form] {problem} {kw}", flush=True)
    p32 = H.gpu_params(run)
    p32.precision = H.PREC["fp32"]
    ref = O.RefSolver(p64)
    g32 = H.GpuSolver(p32)
    for s in (ref, g32):
        s.upload_q0(q0)
        s.init_faces()
    div0 = _d4_div(g32.faces(), p64)
    for _ in range(10):
        dt = ref.tstep()
        ref.ssprk(dt)
        g32.ssprk(dt)
    qr, qg = ref.q0(), g32.q0()
    l1 = H.rel_l1(qg, qr)
    a_r, a_g = H.active(qr), H.active(qg)
    emag = lambda a: 0.5 * (a[4] ** 2 + a[5] ** 2 + a[6] ** 2).sum()
    print("FP32 form rel L1 per variable:", l1, " magnetic energy rel err:",
          abs(emag(a_g) - emag(a_r)) / emag(a_r))
    assert l1[0] <= 1e-3, l1
    assert l1[7] <= 1e-3, l1
    ddiv = np.abs(_d4_div(g32.faces(), p64) - div0).max()
    print("FP32 form max |div_CT change|:", ddiv)
    assert ddiv <= 1e-10


def test_conservation_and_ct_divergence_product_build(require_gpu):
    run, q0, p = _setup("mpi_smooth3d", nx=32, ny=24, nz=16, cour=0.8)
    g = H.GpuSolver(p)
    g.upload_q0(q0)
    g.init_faces()
    d0 = g.diag()
    b0 = g.faces()
    for _ in range(5):
        g.ssprk(g.tstep())
    d1 = g.diag()
    a0 = np.abs(H.active(q0)).reshape(8, -1).sum(axis=1)
    for m in (0, 1, 2, 3, 7):
        scale = a0[1:4].sum() if m in (1, 2, 3) else a0[m]
        assert abs(d1.sum_cons[m] - d0.sum_cons[m]) <= 1e-12 * scale, m
    b1 = g.faces()
    ddiv = np.abs(_d4_div(b1, p) - _d4_div(b0, p)).max()
    scale = np.abs(H.active(b1)).max() / min(p.dx, p.dy, p.dz)
    assert ddiv <= 1e-13 * scale


def _shift(a, axis, off, periodic):
    """a at index i + off along axis, through the boundary map: wrap
    (periodic) or clamp (open), as the ghost fills of src/bound.cpp leave it."""
    n = a.shape[axis]
    idx = np.arange(n) + off
    idx = idx % n if periodic else np.clip(idx, 0, n - 1)
    return np.take(a, idx, axis=axis)


def _diag_from_state(q, b, p):
    """The diagnostics of howmhd_gpu_diag computed in NumPy from a downloaded
    state, each value formed as the reference forms it (src/output.cpp:154-188,
    300-347; the D4 face divergence of src/fluxct.cpp:848)."""
    a, f = H.active(q), H.active(b)
    per = [p.bc[d] == H.BC["periodic"] for d in range(3)]
    h = (p.dx, p.dy, p.dz)
    n = (p.nx, p.ny, p.nz)
    rho = a[0]
    vx, vy, vz = a[1] / rho, a[2] / rho, a[3] / rho
    vv2 = vx * vx + vy * vy + vz * vz
    BB2 = a[4] * a[4] + a[5] * a[5] + a[6] * a[6]
    pg = (p.gam - 1.0) * (a[7] - 0.5 * rho * vv2 - 0.5 * BB2)
    aa, bb = 9.0 / 8.0, -1.0 / 24.0
    dv = np.zeros_like(rho)
    dc = np.zeros_like(rho)
    for d in range(3):
        if n[d] == 1:
            continue
        fd = f[d]
        dv = dv + (aa * (fd - _shift(fd, d, -1, per[d])) + bb * (_shift(fd, d, 1, per[d]) - _shift(fd, d, -2, per[d]))) / h[d]
        B = a[4 + d]
        i = np.arange(n[d])
        im = i - 1 if per[d] else np.clip(i - 1, 0, n[d] - 1)
        ip = i + 1 if per[d] else np.clip(i + 1, 0, n[d] - 1)
        one_sided = ((im % n[d]) == i) | ((ip % n[d]) == i)
        den = np.where(one_sided, h[d], 2.0 * h[d])
        shape = [1, 1, 1]
        shape[d] = n[d]
        dc = dc + (_shift(B, d, 1, per[d]) - _shift(B, d, -1, per[d])) / den.reshape(shape)
    return {"sum": [math.fsum(a[m].ravel()) for m in range(8)],
            "abs_sum": [float(np.abs(a[m]).sum()) for m in range(8)],
            "max_divb_ct": float(np.abs(dv).max()), "max_divb_cell": float(np.abs(dc).max()),
            "min_rho": float(rho.min()), "min_p": float(pg.min())}


def _floor_blast():
    """MHD blast 16^3 with the default floors (rhomin = pgmin = 1e-12), the
    ambient pressure set to 1e-13 and a 3^3 pocket of density 1e-13: prot
    applies both floors (the restatement counts 1 and 1137 cells at the first
    step's final stage).  One step only: the second step drives the pocket to
    non-numbers, in the restatement as on the GPU, and sums of those compare
    nothing."""
    run, q0, p = _setup("mhd_blast3d", nx=16, ny=16, nz=16)
    a = H.active(q0)
    BB2 = a[4] ** 2 + a[5] ** 2 + a[6] ** 2
    amb = (p.gam - 1.0) * (a[7] - 0.5 * BB2) < 1.0
    a[7][amb] = 1e-13 / (p.gam - 1.0) + 0.5 * BB2[amb]
    a[0][2:5, 2:5, 2:5] = 1e-13
    return run, q0, p


DIAG_CASES = {
    "floors_default": (_floor_blast, 1),
    "blast_open": (lambda: _setup("mhd_blast3d", nx=12, ny=14, nz=16, x1bc="open", x3bc="open"), 2),
    "ot2d": (lambda: _setup("orszag_tang", nx=32, ny=32, nz=1, cour=0.8), 3),
    "briowu_open": (lambda: _setup("brio_wu", nx=64, ny=1, nz=1, x1bc="open", x2bc="open", x3bc="open"), 4),
}


@pytest.mark.parametrize("case", list(DIAG_CASES))
def test_diag_against_oracle_state(case, require_gpu):
    """Every field of howmhd_gpu_diag (the FP64 device reduction) against
    values computed from the restatement's state after each step (the plain
    FP64 path equals it bit for bit): maxima, minima and the CT divergence
    exactly, the sums to within the rounding of a reordered sum, and the floor
    counts of the last stage-5 prot equal the restatement's counts, taken
    before its fmax, with the default floors."""
    make, nsteps = DIAG_CASES[case]
    run, q0, p = make()
    print(f"[diag] case={case} shape=({p.nx},{p.ny},{p.nz})", flush=True)
    ref = O.RefSolver(p)
    g = H.GpuSolver(p)
    for s in (ref, g):
        s.upload_q0(q0)
        s.init_faces()
    seen_floors = [0, 0]
    for step in range(nsteps + 1):
        if step:
            dt = ref.tstep()
            ref.ssprk(dt)
            g.ssprk(dt)
        d = g.diag()
        w = _diag_from_state(ref.q0(), ref.faces(), p)
        for m in range(8):
            assert abs(d.sum_cons[m] - w["sum"][m]) <= 1e-13 * w["abs_sum"][m], (step, m)
        assert d.max_divb_ct == w["max_divb_ct"], (step, d.max_divb_ct, w["max_divb_ct"])
        assert d.max_divb_cell == w["max_divb_cell"], (step, d.max_divb_cell, w["max_divb_cell"])
        assert d.min_rho == w["min_rho"] and d.min_p == w["min_p"], step
        want = ref.floor_counts() if step else (0, 0)
        assert (d.n_rho_floor, d.n_p_floor) == want, (step, d.n_rho_floor, d.n_p_floor, want)
        assert d.flagged_fraction == 0.0
        seen_floors[0] += want[0]
        seen_floors[1] += want[1]
    print("floor counts over the steps (rho, p):", seen_floors, " last diag:", d.max_divb_ct, d.max_divb_cell)
    if case == "floors_default":
        assert seen_floors[0] > 0 and seen_floors[1] > 0


def _alfven_error(n, solver_cls, **kw):
    run, q0, p = _setup("cpaw3d", nx=n, ny=n, nz=n, xsize=3.0, ysize=1.5, zsize=1.5, cour=0.4)
    s = solver_cls(p, **kw)
    s.upload_q0(q0)
    s.init_faces()
    t = 0.0
    while t < 1.0:
        dt = min(s.tstep(), 1.0 - t)
        s.ssprk(dt)
        t += dt
    return np.abs(H.active(s.q0())[5] - H.active(q0)[5]).mean()


def test_alfven_wave_convergence_matches_restatement(require_gpu):
    e16 = _alfven_error(16, H.GpuSolver)
    e32 = _alfven_error(32, H.GpuSolver)
    order = math.log2(e16 / e32)
    print(f"GPU Alfven wave L1(By): 16^3 {e16:.4e}  32^3 {e32:.4e}  order {order:.3f}")
    assert 1.6 < order < 2.6
    r16 = _alfven_error(16, O.RefSolver)
    assert abs(e16 - r16) <= 1e-10 * r16


def test_driver_matches_restatement(tmp_path, require_gpu):
    inp = tmp_path / "input.in"
    inp.write_text("""# Orszag-Tang, two outputs (nstep is parsed and not used, as in the reference)
problem orszag_tang
output_format vts
nstep 1
nx 32
ny 32
nz 1
xsize 1.0
ysize 1.0
zsize 1.0
gam 1.6666666666666667
cour 0.8
tend 0.02
nt 2
""")
    out = subprocess.run([str(H.DRIVER), str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    log = out.stdout
    assert "output step = 2" in log and "[advance done ]" in log
    files = sorted((tmp_path / "output").glob("dump*_rank000000.vts"))
    assert [f.name for f in files] == [f"dump{i:06d}_rank000000.vts" for i in range(3)]
    # same loop on the restatement: dt clipped to the output times
    run = H.parse(inp.read_text())
    q0 = H.init_problem(run)
    ref = O.RefSolver(H.gpu_params(run))
    ref.upload_q0(q0)
    ref.init_faces()
    time, dtout = 0.0, run.tend / run.nt
    for out_step in (1, 2):
        target = dtout * out_step
        while time < target:
            dt = ref.tstep()
            if time + dt > target:
                dt = target - time
            ref.ssprk(dt)
            time += dt
    raw = files[-1].read_bytes()
    head, _, tail = raw.partition(b'<AppendedData encoding="raw">\n_')
    off = int(re.search(rb'Name="rho" NumberOfComponents="1" format="appended" offset="(\d+)"', head).group(1))
    nbytes = struct.unpack_from("<Q", tail, off)[0]
    rho = np.frombuffer(tail, dtype="<f8", count=nbytes // 8, offset=off + 8).reshape(1, 32, 32)
    want = H.active(ref.q0())[0].transpose(2, 1, 0)
    assert np.abs(rho - want).sum() / np.abs(want).sum() <= 1e-12
    t_file = float(re.search(rb'Name="time" NumberOfTuples="1" format="ascii">\s*([0-9.e+-]+)', head).group(1))
    assert abs(t_file - time) <= 1e-14 * time


def test_driver_max_steps(tmp_path, require_gpu):
    """max_steps stops the run after that many hydro steps and writes the
    output of the interval it stopped in at the time reached."""
    inp = tmp_path / "input.in"
    inp.write_text(H.make_input(problem="orszag_tang", nx=32, ny=32, nz=1, cour=0.8, tend=0.5, nt=1,
                                output_format="none", max_steps=3))
    out = subprocess.run([str(H.DRIVER), str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "hydro_steps = 3" in out.stdout
    m = re.search(r"output step = 1 time = ([0-9.e+-]+)", out.stdout)
    assert m and 0.0 < float(m.group(1)) < 0.5


def test_fp32_turbulence_384_stays_close_to_fp64(require_gpu):
    """Decaying turbulence at 384^3 for 100 steps (the precision-form
    configuration): the FP32 form stays finite, within density and energy
    relative L1 1e-3 of the plain FP64 path, and its kinetic and magnetic
    energies within 1e-6 of FP64's.  (Regression: with FP32 interface wave speeds
    and eigenvector normalisations this run went non-finite near step 84.)"""
    run, q0, p64 = _setup("turbulence3d", nx=384, ny=384, nz=384, cour=0.4, seed=1)
    p32 = H.gpu_params(run)
    p32.precision = H.PREC["fp32"]
    g64, g32 = H.GpuSolver(p64), H.GpuSolver(p32)
    for s in (g64, g32):
        s.upload_q0(q0)
        s.init_faces()
    del q0
    for step in range(100):
        dt = g64.tstep()
        g64.ssprk(dt)
        g32.ssprk(dt)
        if step % 20 == 19:
            print(f"[turb384] step {step + 1}", flush=True)
    d32 = g32.diag()
    assert np.isfinite(d32.sum_cons[7]) and d32.min_rho > 0.5
    q32, q64 = g32.q0(), g64.q0()
    l1 = H.rel_l1(q32, q64)

    def energies(q):
        a = H.active(q)
        kin = 0.5 * float(np.sum((a[1] ** 2 + a[2] ** 2 + a[3] ** 2) / a[0]))
        mag = 0.5 * float(np.sum(a[4] ** 2 + a[5] ** 2 + a[6] ** 2))
        return kin, mag

    (k32, m32), (k64, m64) = energies(q32), energies(q64)
    ek, em = abs(k32 - k64) / k64, abs(m32 - m64) / m64
    print("turbulence 384^3 FP32 rel L1:", l1, "kinetic / magnetic energy rel err:", ek, em,
          "divB_CT", d32.max_divb_ct)
    assert l1[0] <= 1e-3 and l1[7] <= 1e-3
    # the configuration's three reported errors: density, total and magnetic
    # energy (measured: ~1e-9 and below)
    assert ek <= 1e-6 and em <= 1e-6
    assert d32.max_divb_ct <= 1e-10


def test_driver_with_indicator_and_fp32(tmp_path, require_gpu):
    """`how_mhd input.in` with the added keys: precision fp32, indicator mlp
    and indicator_weights run the routed FP32 form; the final density stays
    within 1e-3 of the plain FP64 restatement."""
    weights = H.PKG_DIR / "weights" / "indicator_mlp.txt"
    inp = tmp_path / "input.in"
    inp.write_text(f"""problem mhd_blast3d
output_format vts
nx 32
ny 32
nz 32
cour 0.4
tend 0.004
nt 1
precision fp32
indicator mlp
indicator_weights {weights}
quiet 1
""")
    out = subprocess.run([str(H.DRIVER), str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "indicator weights" in out.stdout and "precision          = fp32" in out.stdout
    files = sorted((tmp_path / "output").glob("dump*_rank000000.vts"))
    assert len(files) == 2
    run = H.parse(inp.read_text())
    p = H.gpu_params(run)
    p.precision = H.PREC["fp64"]
    p.indicator = H.INDICATOR["off"]
    ref = O.RefSolver(p)
    ref.upload_q0(H.init_problem(run))
    ref.init_faces()
    t = 0.0
    while t < run.tend:
        dt = min(ref.tstep(), run.tend - t)
        ref.ssprk(dt)
        t += dt
    raw = files[-1].read_bytes()
    head, _, tail = raw.partition(b'<AppendedData encoding="raw">\n_')
    off = int(re.search(rb'Name="rho" NumberOfComponents="1" format="appended" offset="(\d+)"', head).group(1))
    nbytes = struct.unpack_from("<Q", tail, off)[0]
    rho = np.frombuffer(tail, dtype="<f8", count=nbytes // 8, offset=off + 8).reshape(32, 32, 32)
    want = H.active(ref.q0())[0].transpose(2, 1, 0)
    l1 = np.abs(rho - want).sum() / np.abs(want).sum()
    print("driver fp32+indicator vs FP64 restatement, density rel L1:", l1)
    assert l1 <= 1e-3


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_driver_restart_continues_bit_for_bit(tmp_path, precision, require_gpu):
    """how_mhd with `checkpoint_every 2` over 4 outputs, and a second run that
    resumes from the output-2 checkpoint: outputs 3 and 4 of the two runs are
    the same bytes (q(0) and faces travel through the checkpoint exactly; the
    first tstep after the restart takes its CFL maxima from q(0) itself)."""
    text = f"""problem mhd_blast3d
output_format vts
nx 16
ny 12
nz 10
gam 1.6666666666666667
cour 0.4
tend 0.02
nt 4
precision {precision}
quiet 1
"""
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    (a / "input.in").write_text(text + "checkpoint_every 2\n")
    ra = subprocess.run([str(H.DRIVER), "input.in"], cwd=a, capture_output=True, text=True, timeout=120)
    assert ra.returncode == 0, ra.stderr
    ckpt = a / "checkpoint" / "ckpt000002.bin"
    assert ckpt.exists() and (a / "checkpoint" / "ckpt000004.bin").exists()
    (b / "input.in").write_text(text + f"restart {ckpt}\n")
    rb = subprocess.run([str(H.DRIVER), "input.in"], cwd=b, capture_output=True, text=True, timeout=120)
    assert rb.returncode == 0, rb.stderr
    assert "restart" in rb.stdout and "output step = 3" in rb.stdout
    assert sorted(f.name for f in (b / "output").glob("dump*.vts")) == [
        "dump000003_rank000000.vts", "dump000004_rank000000.vts"]
    for k in (3, 4):
        name = f"dump{k:06d}_rank000000.vts"
        assert (a / "output" / name).read_bytes() == (b / "output" / name).read_bytes(), name


def test_fp64_512_invariants(require_gpu):
    """The largest benchmark size, MHD blast 512^3 in the plain FP64 form, the
    configuration's fixed 50 steps (the restatement would take minutes per
    step here, so the check is by size-independent properties): mass and
    momentum conserved on the periodic grid, energy up to prot's rebuild from
    the face-interpolated B (src/prot.cpp:45-46), max |div_CT| <= 1e-10, the
    state finite and positive, and dt from the fused CFL maxima equal to the
    CFL kernel's dt of the same state re-uploaded."""
    run, q0, p = _setup("mhd_blast3d", nx=512, ny=512, nz=512, cour=0.4)
    print("[512^3 fp64] mhd_blast3d, 50 steps", flush=True)
    g = H.GpuSolver(p)
    try:
        g.upload_q0(q0)
        g.init_faces()
        d0 = g.diag()
        for _ in range(50):
            g.ssprk(g.tstep())
        d1 = g.diag()
        dt_fused = g.tstep()
        q1 = g.q0()
    finally:
        g.close()
    print("512^3 fp64: divB_CT", d1.max_divb_ct, "min rho", d1.min_rho, "min p", d1.min_p, flush=True)
    rho0 = np.abs(H.active(q0)[0]).sum()
    e0 = np.abs(H.active(q0)[7]).sum()
    del q0
    # the blast starts at rest: momentum is measured against what it grew to
    mscale = np.abs(H.active(q1)[1:4]).sum()
    assert abs(d1.sum_cons[0] - d0.sum_cons[0]) <= 1e-12 * rho0
    for m in (1, 2, 3):
        assert abs(d1.sum_cons[m] - d0.sum_cons[m]) <= 1e-12 * mscale, m
    assert abs(d1.sum_cons[7] - d0.sum_cons[7]) <= 1e-10 * e0
    assert d1.max_divb_ct <= 1e-10
    assert np.isfinite(d1.min_rho) and np.isfinite(d1.min_p) and d1.min_rho > 0 and d1.min_p > 0
    # the same dt from the CFL kernel over the re-uploaded state
    g2 = H.GpuSolver(p)
    try:
        g2.upload_q0(q1)
        assert g2.tstep() == dt_fused
    finally:
        g2.close()
