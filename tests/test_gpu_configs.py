"""The reference configurations at their own sizes (needs a GPU).

  * Orszag-Tang 128 x 128 to t = 0.5 through the drop-in driver: the VTS
    output (all 8 conserved fields and the cell divB) equals, byte for byte,
    the file the host writer makes from the CPU restatement's state;
  * 3-D circularly polarised Alfven wave, plain FP64, 32^3 / 64^3 / 128^3 /
    256^3, one period: L1 error of every conserved variable against the
    exact solution (the initial state) and the convergence orders;
  * MHD blast 256^3, plain FP64, the fixed 50 steps: conservation, CT
    divergence, positivity; and the FP32 + indicator form on the same dt
    sequence;
  * learned-indicator evaluation on the 512^2 MHD rotor to t = 0.15 and on
    256^3 decaying turbulence over 50 steps: flagged fraction, density and
    energy L1 against the plain FP64 run, CT divergence, and what the trained
    network gains over routing nothing.

The plain FP64 form equals the CPU restatement bit for bit on every case the
restatement can run in seconds (test_gpu_parity.py); here the sizes are those
of the configurations, and the checks are the properties their reports state.
"""
import math
import re
import subprocess

import numpy as np
import pytest

import eval_indicator as EV
import howmhd as H
import oracle_ctypes as O

pytestmark = pytest.mark.gpu

NAMES = ["rho", "Mx", "My", "Mz", "Bx", "By", "Bz", "E"]


def _run(problem, **kw):
    kw.setdefault("nt", 1)
    kw.setdefault("tend", 1.0)
    run = H.parse(H.make_input(problem=problem, **kw))
    return run, H.init_problem(run), H.gpu_params(run)


def test_orszag_tang_128_to_t05_driver_matches_restatement(tmp_path, require_gpu):
    """Orszag-Tang 128 x 128 x 1 (gam 5/3, cour 0.8, tend 0.5, periodic)
    through `how_mhd input.in` with one VTS output at t = 0.5, against the CPU
    restatement run through the reference's main loop (dt from tstep, clipped
    to the output time, src/main.cpp:165-223).  The host writer turns the
    restatement's q(0) into a VTS file of its own; the two files are equal
    byte for byte: rho, momentum, magnetic field, energy, the derived fields
    and the cell-centred divB of src/output.cpp:314-347, and the time."""
    inp = tmp_path / "input.in"
    inp.write_text(H.make_input(problem="orszag_tang", nx=128, ny=128, nz=1, xsize=1.0, ysize=1.0, zsize=1.0,
                                gam="1.6666666666666667", cour=0.8, tend=0.5, nt=1, output_format="vts"))
    out = subprocess.run([str(H.DRIVER), str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    gpu_file = tmp_path / "output" / "dump000001_rank000000.vts"
    run = H.parse(inp.read_text())
    q0 = H.init_problem(run)
    ref = O.RefSolver(H.gpu_params(run))
    ref.upload_q0(q0)
    ref.init_faces()
    time, steps = 0.0, 0
    while time < run.tend:
        dt = ref.tstep()
        if time + dt > run.tend:
            dt = run.tend - time
        ref.ssprk(dt)
        time += dt
        steps += 1
    ref_dir = tmp_path / "ref"
    H.write_output(run, ref.q0(), 1, time, str(ref_dir))
    ref.close()
    ref_file = ref_dir / "output" / "dump000001_rank000000.vts"
    print(f"Orszag-Tang 128^2 to t = {time}: {steps} steps; {gpu_file.stat().st_size} bytes", flush=True)
    m = re.search(r"output step = 1 time = \S+ last_dt = \S+ hydro_steps = (\d+)", out.stdout)
    assert m and int(m.group(1)) == steps, out.stdout[-500:]
    assert gpu_file.read_bytes() == ref_file.read_bytes()


def test_cpaw_convergence_32_to_256(require_gpu):
    """Circularly polarised Alfven wave along (1,2,2)/3 in [0,3]x[0,1.5]^2,
    rho 1, p 0.1, B_par 1, B_perp 0.1, v_perp = -B_perp, gam 5/3, cour 0.4,
    one period, at 32^3, 64^3, 128^3 and 256^3.  The wave's own components
    (momentum and field) converge at 2nd order (the reference's CT flux terms
    average B_n v_t over two cells, src/ssprk.cpp:979-989); density and
    energy, which the wave leaves unchanged to first order, at about 5th
    order.  Measured: B and M 1.98-2.00, rho 4.98-5.12, E 5.3-5.6."""
    errs = []
    for n in (32, 64, 128, 256):
        run, q0, p = _run("cpaw3d", nx=n, ny=n, nz=n, xsize=3.0, ysize=1.5, zsize=1.5, cour=0.4)
        g = H.GpuSolver(p)
        try:
            g.upload_q0(q0)
            g.init_faces()
            t, steps = 0.0, 0
            while t < 1.0:
                dt = min(g.tstep(), 1.0 - t)
                g.ssprk(dt)
                t += dt
                steps += 1
            q = g.q0()
            d = g.diag()
        finally:
            g.close()
        e = np.abs(H.active(q) - H.active(q0)).reshape(8, -1).mean(axis=1)
        errs.append(e)
        print(f"cpaw {n}^3: {steps} steps, L1 " + " ".join(f"{nm} {v:.3e}" for nm, v in zip(NAMES, e)),
              f"max|div_CT| {d.max_divb_ct:.2e}", flush=True)
        assert np.all(np.isfinite(e))
        assert d.max_divb_ct <= 1e-10
    for i, pair in enumerate(("32->64", "64->128", "128->256")):
        order = [math.log2(errs[i][m] / errs[i + 1][m]) for m in range(8)]
        print("cpaw order", pair, " ".join(f"{nm} {o:.3f}" for nm, o in zip(NAMES, order)), flush=True)
        for m in range(1, 7):
            assert 1.9 <= order[m] <= 2.1, (pair, NAMES[m], order[m])
        assert order[0] >= 4.8, (pair, order[0])
        assert order[7] >= 4.8, (pair, order[7])


def test_blast_256_fifty_steps(require_gpu):
    """MHD blast at 256^3 (box [-0.5,0.5]^3, rho 1, p 0.1 / 10 for r < 0.1,
    B = (1,1,0)/sqrt2, gam 5/3, cour 0.4), the configuration's fixed 50 steps
    in the plain FP64 form: mass and momentum conserved to 1e-12, energy to
    1e-10 (prot rebuilds E from the face-interpolated B, src/prot.cpp:45-46),
    max |div_CT| <= 1e-10, the state finite with positive density and
    pressure."""
    run, q0, p = _run("mhd_blast3d", nx=256, ny=256, nz=256, cour=0.4)
    g = H.GpuSolver(p)
    try:
        g.upload_q0(q0)
        g.init_faces()
        d0 = g.diag()
        t = 0.0
        for _ in range(50):
            dt = g.tstep()
            g.ssprk(dt)
            t += dt
        d1 = g.diag()
        q = g.q0()
    finally:
        g.close()
    print(f"blast 256^3, 50 steps to t = {t:.5f}: max|div_CT| {d1.max_divb_ct:.2e}, "
          f"min rho {d1.min_rho:.4f}, min p {d1.min_p:.4f}", flush=True)
    a0, a1 = H.active(q0), H.active(q)
    assert np.all(np.isfinite(a1))
    assert abs(d1.sum_cons[0] - d0.sum_cons[0]) <= 1e-12 * np.abs(a0[0]).sum()
    mscale = np.abs(a1[1:4]).sum()
    for m in (1, 2, 3):
        assert abs(d1.sum_cons[m] - d0.sum_cons[m]) <= 1e-12 * mscale, m
    assert abs(d1.sum_cons[7] - d0.sum_cons[7]) <= 1e-10 * np.abs(a0[7]).sum()
    assert d1.max_divb_ct <= 1e-10
    assert d1.min_rho > 0 and d1.min_p > 0
    assert d1.n_rho_floor == 0 and d1.n_p_floor == 0


def _check_product_form(out, expect_routed):
    f64, f32, ind = out["forms"]["fp64"], out["forms"]["fp32"], out["forms"]["fp32+indicator"]
    assert f64["rel_l1_rho"] == 0.0 and f64["rel_l1_E"] == 0.0   # the same run twice: bit for bit
    for form in ("fp32", "fp32+indicator"):
        e = out["forms"][form]
        assert e["finite"], form
        assert e["rel_l1_rho"] <= 1e-3, (form, e["rel_l1_rho"])
        assert e["divb_ct_max"] <= 1e-10, (form, e["divb_ct_max"])
    if expect_routed:
        assert ind["routed_steps"] > 0
    assert 0.0 < ind["flagged_fraction"] < 1.0
    return f32, ind


def test_indicator_rotor_512(require_gpu):
    """Learned-indicator evaluation on the 512^2 MHD rotor to t = 0.15 (rho 10
    inside r < 0.1 tapered to 1 at r = 0.115, omega 20, p 1, Bx 5/sqrt(4 pi),
    gam 1.4, cour 0.8; src/problems/mhd_rotor.cpp:43-108).  Every form runs
    the plain FP64 run's dt sequence.  The FP32 + indicator form keeps
    density within 1e-3 (relative L1) of plain FP64 and max |div_CT| at
    round-off; routing every interface to upwind-5 instead (no network) is
    at least 10x further from plain FP64 in density, which is what the
    trained network buys.  Measured: about 700 steps, 24% of interfaces
    flagged, rho L1 4.9e-5 (indicator) vs 1.9e-2 (no network)."""
    out = EV.run_case("rotor512", time_steps=5, forms=["fp64", "fp32", "fp32+indicator", "fp32+linear"])
    assert abs(out["t_end"] - 0.15) <= 1e-12
    f32, ind = _check_product_form(out, expect_routed=True)
    lin = out["forms"]["fp32+linear"]
    print(f"rotor 512^2: {out['steps']} steps, flagged {ind['flagged_fraction']:.3f}, rho L1: fp32 "
          f"{f32['rel_l1_rho']:.2e}, fp32+indicator {ind['rel_l1_rho']:.2e}, no network {lin['rel_l1_rho']:.2e}",
          flush=True)
    assert 0.05 <= ind["flagged_fraction"] <= 0.6
    assert lin["rel_l1_rho"] >= 10 * ind["rel_l1_rho"]


def test_indicator_turbulence_256(require_gpu):
    """Learned-indicator evaluation on 256^3 decaying turbulence (seed 1),
    50 steps along the plain FP64 run's dt sequence: the FP32 + indicator form
    within 1e-3 of plain FP64 in density (relative L1), max |div_CT| at
    round-off.  Nearly every wavefront is flagged on this flow, so the
    product form spends most steps on the plain-WENO fallback; the flagged
    fraction and the step counts are printed.  Measured: 92% of interfaces
    flagged, 7 routed and 43 fallback steps, rho L1 5.1e-8."""
    out = EV.run_case("turb256", time_steps=3, forms=["fp64", "fp32", "fp32+indicator"])
    assert out["steps"] == 50
    f32, ind = _check_product_form(out, expect_routed=False)
    assert ind["routed_steps"] + ind["fallback_steps"] == 50
    print(f"turbulence 256^3: flagged {ind['flagged_fraction']:.3f}, routed {ind['routed_steps']}, "
          f"fallback {ind['fallback_steps']}, rho L1: fp32 {f32['rel_l1_rho']:.2e}, "
          f"fp32+indicator {ind['rel_l1_rho']:.2e}", flush=True)


def test_indicator_blast_256(require_gpu):
    """The FP32 + indicator form on the 50-step 256^3 MHD blast (the headline
    configuration) along the plain FP64 dt sequence: density within 1e-3
    (relative L1), CT divergence at round-off, and a small flagged fraction
    (the flow is smooth away from the blast front).  Measured: 0.8% of
    interfaces flagged, rho L1 1.5e-6."""
    out = EV.run_case("blast256", time_steps=3, forms=["fp64", "fp32", "fp32+indicator"])
    assert out["steps"] == 50
    f32, ind = _check_product_form(out, expect_routed=True)
    assert ind["flagged_fraction"] <= 0.1
    print(f"blast 256^3: flagged {ind['flagged_fraction']:.4f}, rho L1: fp32 {f32['rel_l1_rho']:.2e}, "
          f"fp32+indicator {ind['rel_l1_rho']:.2e}", flush=True)
