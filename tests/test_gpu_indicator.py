"""The learned smoothness indicator on the GPU (needs a GPU).

  * howmhd_mlp_infer / howmhd_mlp_grad / howmhd_mlp_adam against the NumPy
    restatement of the same numerics (this also pins the MFMA operand layout:
    a transposed fragment gives O(1) differences);
  * sample collection: features and labels of every interface against the
    restatement's primitives and WENO weights (periodic and open axes);
  * routing: every interface flagged == the plain method bit for bit, none
    flagged == the restatement's linear route bit for bit (plain FP64 form,
    reference rounding); a mixed routing keeps mass, momentum, energy and the CT face
    divergence to round-off, and FP32 sweeps run routed.
"""
import numpy as np
import pytest

import howmhd as H
import mlp_oracle as M
import oracle_ctypes as O

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _params(seed, b3=None):
    rng = np.random.default_rng(seed)
    P = M.init_params(seed)
    P[M.OFF_B1:M.OFF_W2] = rng.normal(0, 0.2, 32)
    P[M.OFF_B2:M.OFF_W3] = rng.normal(0, 0.2, 32)
    if b3 is not None:
        P[M.OFF_W3:M.OFF_B3] = 0.0
        P[M.OFF_B3] = b3
    return P


@pytest.mark.parametrize("n", [64, 1000, 10007])
def test_mlp_infer_matches_restatement(n, require_gpu):
    torch = _torch()
    print(f"[mlp infer] n={n}", flush=True)
    rng = np.random.default_rng(n)
    X = rng.uniform(0, 1, (n, 18)).astype(np.float32)
    P = _params(1)
    z = H.mlp_infer(torch.from_numpy(X).cuda(), torch.from_numpy(P).cuda()).cpu().numpy()
    zr, _ = M.forward(P, X)
    d = np.abs(z - zr)
    scale = 1 + np.abs(zr)
    print("infer max |dz|:", d.max(), " mean:", d.mean(), " mean |z|:", np.abs(zr).mean())
    assert np.all(d <= 2e-2 * scale)
    assert np.mean(d <= 1e-4 * scale) >= 0.99
    assert d.mean() <= 1e-3


def test_mlp_grad_matches_restatement(require_gpu):
    torch = _torch()
    n = 5000
    rng = np.random.default_rng(7)
    X = rng.uniform(0, 1, (n, 18)).astype(np.float32)
    y = (rng.uniform(size=n) < 0.3).astype(np.float32)
    P = _params(2)
    G, loss = H.mlp_grad(torch.from_numpy(X).cuda(), torch.from_numpy(y).cuda(), torch.from_numpy(P).cuda(),
                         pos_weight=2.5, scale=1.0 / n)
    G, loss = G.cpu().numpy().astype(np.float64), float(loss.cpu().numpy()[0])
    lr, Gr = M.loss_and_grad(P, X, y, pos_weight=2.5, scale=1.0 / n)
    print("loss gpu", loss, "restatement", lr)
    assert abs(loss - lr) <= 1e-5 * abs(lr)
    groups = {"W1": (M.OFF_W1, M.OFF_B1), "b1": (M.OFF_B1, M.OFF_W2), "W2": (M.OFF_W2, M.OFF_B2),
              "b2": (M.OFF_B2, M.OFF_W3), "w3": (M.OFF_W3, M.OFF_B3), "b3": (M.OFF_B3, M.NPARAM)}
    for name, (a, b) in groups.items():
        err = np.linalg.norm(G[a:b] - Gr[a:b]) / np.linalg.norm(Gr[a:b])
        print(f"grad {name}: rel L2 {err:.3e}")
        assert err <= 5e-3, name


def test_mlp_adam_matches_restatement(require_gpu):
    torch = _torch()
    rng = np.random.default_rng(3)
    P = _params(3)
    Mm = rng.normal(0, 1e-3, M.NPARAM).astype(np.float32)
    V = rng.uniform(0, 1e-5, M.NPARAM).astype(np.float32)
    G = rng.normal(0, 1e-2, M.NPARAM).astype(np.float32)
    dev = [torch.from_numpy(a.copy()).cuda() for a in (P, Mm, V, G)]
    H.mlp_adam(*dev, t=5, lr=1e-3)
    M.adam(P, Mm, V, G, 5, 1e-3)
    for a, b in zip(dev[:3], (P, Mm, V)):
        np.testing.assert_allclose(a.cpu().numpy(), b, rtol=2e-6, atol=1e-9)


def _setup(problem, **kw):
    kw.setdefault("nt", 1)
    kw.setdefault("tend", 1.0)
    run = H.parse(H.make_input(problem=problem, **kw))
    return run, H.init_problem(run), H.gpu_params(run)


@pytest.mark.parametrize("problem,kw,steps,mixed", [
    ("orszag_tang", dict(nx=32, ny=32, nz=1, cour=0.8), 4, False),   # under-resolved: nearly all flagged
    ("mhd_blast3d", dict(nx=12, ny=14, nz=16, x1bc="open", cour=0.4), 2, True),
])
def test_collect_samples_match_restatement(problem, kw, steps, mixed, require_gpu):
    torch = _torch()
    run, q0, p = _setup(problem, **kw)
    print(f"[collect] {problem} {kw}", flush=True)
    g = H.GpuSolver(p)
    g.upload_q0(q0)
    g.init_faces()
    for _ in range(steps):
        g.ssprk(g.tstep())
    qa = H.active(g.q0())
    for d in range(3):
        n = g.n_samples(d)
        if n == 0:
            continue
        X = torch.zeros((n, 18), dtype=torch.float32, device="cuda")
        y = torch.zeros(n, dtype=torch.float32, device="cuda")
        assert g.collect_samples(d, X, y) == n
        X, y = X.cpu().numpy(), y.cpu().numpy()
        per = p.bc[d] == H.BC["periodic"]
        fr = M.interface_features(qa, d, per, p.gam, p.rhomin, p.pgmin)
        lr, dev = M.interface_labels(qa, d, per, p.gam, O.line_maxdev, p.rhomin, p.pgmin)
        assert X.shape == fr.shape
        # float32 casts of FP64 values that differ in the last bits can put a
        # stencil on either side of the flatness floor: allow 0.1 % of samples
        bad = np.any(np.abs(X - fr) > 2e-5, axis=1)
        assert bad.mean() <= 1e-3, bad.mean()
        clear = np.abs(dev - 0.05) > 1e-9
        print(f"dir {d}: {n} samples, flagged fraction {y.mean():.3f}")
        assert np.array_equal(y[clear], lr[clear])
        assert y.sum() > 0
        if mixed:
            assert y.sum() < n


@pytest.mark.parametrize("problem,kw", [
    ("orszag_tang", dict(nx=32, ny=32, nz=1, cour=0.8)),
    ("mhd_blast3d", dict(nx=12, ny=14, nz=16, x1bc="open", cour=0.4)),
])
def test_routing_extremes_plain_fp64(problem, kw, require_gpu):
    run, q0, p = _setup(problem, **kw)
    print(f"[routing extremes] {problem} {kw}", flush=True)
    ref_w, ref_l = O.RefSolver(p), O.RefSolver(p)
    ref_l.set_route(1)
    g_all, g_none = H.GpuSolver(p), H.GpuSolver(p)
    g_all.set_indicator(_params(0, b3=50.0), 0.5, enable=2)
    g_none.set_indicator(_params(0, b3=-50.0), 0.5, enable=2)
    for s in (ref_w, ref_l, g_all, g_none):
        s.upload_q0(q0)
        s.init_faces()
    for _ in range(3):
        dt = ref_w.tstep()
        for s in (ref_w, ref_l, g_all, g_none):
            s.ssprk(dt)
    for g, r in ((g_all, ref_w), (g_none, ref_l)):
        a, b = H.active(g.q0()), H.active(r.q0())
        assert np.count_nonzero(a.view(np.int64) != b.view(np.int64)) == 0
        fa, fb = H.active(g.faces()), H.active(r.faces())
        assert np.count_nonzero(fa.view(np.int64) != fb.view(np.int64)) == 0
    st_all, st_none = g_all.indicator_stats(), g_none.indicator_stats()
    assert st_all["flagged"] == st_all["total"] > 0
    assert st_none["flagged"] == 0 and st_none["total"] == st_all["total"]
    # interfaces evaluated: 15 sweeps (5 stages x 3 steps) over every direction
    per_step = sum(g_all.n_samples(d) for d in range(3))
    assert st_all["total"] == 15 * per_step


def test_routing_mixed_conserves(require_gpu):
    run, q0, p = _setup("mpi_smooth3d", nx=32, ny=24, nz=16, cour=0.8)
    g = H.GpuSolver(p)
    # b3 = 0 with random hidden layers: a mixed routing
    P = _params(5)
    P[M.OFF_B3] = 0.0
    g.set_indicator(P, 0.5, enable=2)
    g.upload_q0(q0)
    g.init_faces()
    d0 = g.diag()
    b0 = g.faces()
    for _ in range(5):
        g.ssprk(g.tstep())
    st = g.indicator_stats()
    print("mixed routing stats:", st)
    assert 0.02 < st["flagged_fraction"] < 0.98
    d1 = g.diag()
    a0 = np.abs(H.active(q0)).reshape(8, -1).sum(axis=1)
    for m in (0, 1, 2, 3, 7):
        scale = a0[1:4].sum() if m in (1, 2, 3) else a0[m]
        assert abs(d1.sum_cons[m] - d0.sum_cons[m]) <= 1e-12 * scale, m
    aa, bb = 9.0 / 8.0, -1.0 / 24.0

    def div(b):
        act = H.active(b)
        out = 0.0
        for dd, h in enumerate((p.dx, p.dy, p.dz)):
            f = act[dd]
            out = out + (aa * (f - np.roll(f, 1, dd)) + bb * (np.roll(f, -1, dd) - np.roll(f, 2, dd))) / h
        return out
    b1 = g.faces()
    scale = np.abs(H.active(b1)).max() / min(p.dx, p.dy, p.dz)
    assert np.abs(div(b1) - div(b0)).max() <= 1e-13 * scale
    # the diagnostics report the flagged fraction of the most recent routed step
    g.ssprk(g.tstep())
    st1 = g.indicator_stats()
    assert g.diag().flagged_fraction == st1["flagged"] / st1["total"]


def test_routing_fp32_all_flagged_matches_plain_fp32(require_gpu):
    run, q0, p = _setup("mhd_blast3d", nx=32, ny=32, nz=32, cour=0.4)
    p.precision = H.PREC["fp32"]
    a, b = H.GpuSolver(p), H.GpuSolver(p)
    b.set_indicator(_params(0, b3=50.0), 0.5, enable=1)
    for s in (a, b):
        s.upload_q0(q0)
        s.init_faces()
    for _ in range(3):
        dt = a.tstep()
        a.ssprk(dt)
        b.ssprk(dt)
    l1 = H.rel_l1(b.q0(), a.q0())
    print("fp32 routed(all) vs plain fp32 rel L1:", l1)
    assert np.all(l1 <= 1e-12)


def test_indicator_mode_requires_weights(require_gpu):
    run, q0, p = _setup("orszag_tang", nx=16, ny=16, nz=1)
    p.indicator = H.INDICATOR["mlp"]
    g = H.GpuSolver(p)
    g.upload_q0(q0)
    g.init_faces()
    with pytest.raises(H.HowmhdError, match="needs weights"):
        g.ssprk(1e-3)
    with pytest.raises(H.HowmhdError, match="threshold"):
        g.set_indicator(_params(0), 1.5)
    g.set_indicator(_params(0), 0.5)
    g.ssprk(g.tstep())


def test_trained_indicator_ablation(require_gpu):
    """What the trained network gains over the method without it, on the MHD
    rotor (128^2, gam 1.4, to t = 0.05; the reference is the plain FP64 method):
      * FP32 + trained indicator: density within 1e-3 of the reference, with a
        minority of interfaces on full WENO;
      * every interface on linear upwind-5 (no indicator): >= 10x its error;
      * an untrained network flagging the same fraction: >= 10x its error."""
    import eval_indicator as E
    rep = E.run_case("rotor128", time_steps=3,
                     forms=["fp64", "fp32", "fp32+indicator", "fp32+linear", "fp32+untrained"])
    f = rep["forms"]
    tr, lin, un = f["fp32+indicator"], f["fp32+linear"], f["fp32+untrained"]
    print("trained", tr, "\nlinear", lin, "\nuntrained", un)
    assert tr["finite"] and tr["rel_l1_rho"] <= 1e-3 and tr["divb_ct_max"] <= 1e-10
    assert tr["flagged_fraction"] < 0.5
    assert (not lin["finite"]) or lin["rel_l1_rho"] >= 10 * tr["rel_l1_rho"]
    assert (not un["finite"]) or un["rel_l1_rho"] >= 10 * tr["rel_l1_rho"]


def test_training_script_learns_the_labels(tmp_path, require_gpu):
    """train_indicator.py end to end at 1/8 of its grids (Orszag-Tang 32^2,
    blast and turbulence 8^3): samples collected by the FP64 sweep, 4 epochs
    on the library's gradient and Adam kernels, threshold for held-out recall
    0.99.  The network must beat a flagger that knows only the positive rate
    (its precision would equal that rate), and the weights file it writes must
    load through the host library and route a step."""
    import train_indicator as TR
    out = tmp_path / "w.txt"
    st = TR.main(["--out", str(out), "--scale", "0.125", "--samples", "4e5", "--epochs", "4",
                  "--batch", "2048", "--recall", "0.99", "--seed", "3"])
    print("held-out:", st, flush=True)
    assert st["samples"] >= 1000
    assert 0.05 < st["positive_fraction"] < 0.95
    assert st["recall"] >= 0.98
    assert st["precision"] >= st["positive_fraction"] + 0.1
    assert st["flagged_fraction"] < 0.95
    P, thr = H.load_indicator(out)
    assert P.shape == (H.MLP_NPARAM,) and np.all(np.isfinite(P)) and abs(thr - st["threshold"]) < 1e-6
    run = H.parse(H.make_input(problem="orszag_tang", nx=32, ny=32, nz=1, cour=0.8, nt=1, tend=1.0))
    p = H.gpu_params(run)
    p.precision = H.PREC["fp32"]
    g = H.GpuSolver(p)
    try:
        g.set_indicator(P, thr, enable=2)
        g.upload_q0(H.init_problem(run))
        g.init_faces()
        for _ in range(3):
            g.ssprk(g.tstep())
        assert np.all(np.isfinite(H.active(g.q0())))
        assert g.indicator_stats()["total"] > 0
    finally:
        g.close()


def test_plain_weno_fallback(require_gpu):
    """Routing mode 1: where nearly every wavefront is flagged (decaying
    turbulence) the steps fall back to the plain sweep, re-checking every 9th
    step; where few are flagged (blast) every step routes.  Steps in fallback
    equal the plain method."""
    import eval_indicator as E
    P, thr = H.load_indicator(E.WEIGHTS)
    # turbulence: nearly everything flagged
    run, q0, p = _setup("turbulence3d", nx=32, ny=32, nz=32, cour=0.4, seed=2)
    p.precision = H.PREC["fp32"]
    g, plain = H.GpuSolver(p), H.GpuSolver(p)
    g.set_indicator(P, thr, enable=1)
    for s in (g, plain):
        s.upload_q0(q0)
        s.init_faces()
    for _ in range(20):
        dt = plain.tstep()  # g steps with the same dt, without a tstep of its own
        plain.ssprk(dt)
        g.ssprk(dt)
    st = g.indicator_stats()
    print("turbulence routing:", st)
    assert st["fallback_steps"] >= 8 and st["routed_steps"] >= 2
    assert st["routed_steps"] + st["fallback_steps"] == 20
    assert H.rel_l1(g.q0(), plain.q0())[0] <= 1e-4
    # blast: few interfaces flagged, no fallback
    run, q0, p = _setup("mhd_blast3d", nx=32, ny=32, nz=32, cour=0.4)
    p.precision = H.PREC["fp32"]
    g = H.GpuSolver(p)
    g.set_indicator(P, thr, enable=1)
    g.upload_q0(q0)
    g.init_faces()
    for _ in range(6):
        g.ssprk(g.tstep())
    st = g.indicator_stats()
    print("blast routing:", st)
    assert st["fallback_steps"] == 0 and st["routed_steps"] == 6


def test_in_sweep_network_matches_standalone_kernel(require_gpu):
    """The routed sweep's network (transposed MFMA products on the sweep's
    LDS tile) gives the standalone inference kernel's logits on the same
    features."""
    torch = _torch()
    import eval_indicator as E
    P, thr = H.load_indicator(E.WEIGHTS)
    run, q0, p = _setup("mhd_blast3d", nx=24, ny=20, nz=16, x1bc="open", cour=0.4)
    g = H.GpuSolver(p)
    g.set_indicator(P, thr, enable=2)
    g.upload_q0(q0)
    g.init_faces()
    for _ in range(3):
        g.ssprk(g.tstep())
    Pd = torch.from_numpy(P).cuda()
    for d in range(3):
        n = g.n_samples(d)
        X = torch.zeros((n, 18), dtype=torch.float32, device="cuda")
        y = torch.zeros(n, dtype=torch.float32, device="cuda")
        zs = torch.zeros(n, dtype=torch.float32, device="cuda")
        assert g.collect_samples(d, X, y) == n
        assert g.indicator_logits(d, zs) == n
        zr = H.mlp_infer(X, Pd).cpu().numpy()
        zs = zs.cpu().numpy()
        dz = np.abs(zs - zr)
        scale = 1 + np.abs(zr)
        print(f"dir {d}: max |dz| {dz.max():.3e}, mean {dz.mean():.3e}")
        assert np.all(dz <= 2e-2 * scale)
        assert np.mean(dz <= 1e-4 * scale) >= 0.99


def test_fp32_indicator_full_size_invariants(require_gpu):
    """The product form at the benchmark size: MHD blast 256^3, FP32 with the
    trained indicator (plain-WENO fallback on), 3 steps.  Size-independent
    properties: mass and momentum are conserved on the periodic grid (one flux
    per interface, FP64 stage sums), energy up to prot's rebuild from the
    face-interpolated B, max |div_CT| stays below 1e-10
    (FP64 face update of FP32 EMFs), the state stays finite, and the routing
    is active (a small flagged fraction)."""
    import eval_indicator as E
    run, q0, p = _setup("mhd_blast3d", nx=256, ny=256, nz=256, cour=0.4)
    p.precision = H.PREC["fp32"]
    P, thr = H.load_indicator(E.WEIGHTS)
    g = H.GpuSolver(p)
    try:
        g.set_indicator(P, thr, enable=1)
        g.upload_q0(q0)
        g.init_faces()
        d0 = g.diag()
        for _ in range(3):
            g.ssprk(g.tstep())
        d1 = g.diag()
        st = g.indicator_stats()
        print("256^3 fp32+indicator:", st, "divB_CT", d1.max_divb_ct, "min rho", d1.min_rho)
        a0 = np.abs(H.active(q0)).reshape(8, -1).sum(axis=1)
        # the blast starts at rest: momentum is measured against what it grew to
        a1 = np.abs(H.active(g.q0())).reshape(8, -1).sum(axis=1)
        for m in (0, 1, 2, 3):
            scale = a1[1:4].sum() if m in (1, 2, 3) else a0[m]
            assert abs(d1.sum_cons[m] - d0.sum_cons[m]) <= 1e-12 * scale, m
        # energy: prot rebuilds E from p with the face-interpolated B
        # (src/prot.cpp:45-46, src/fluxct.cpp:891-896), which moves E by the
        # difference of the two magnetic energies; measured 2.5e-12 here
        assert abs(d1.sum_cons[7] - d0.sum_cons[7]) <= 1e-10 * a0[7]
        assert d1.max_divb_ct <= 1e-10
        assert np.isfinite(d1.min_rho) and np.isfinite(d1.min_p) and d1.min_rho > 0 and d1.min_p > 0
        assert st["routed_steps"] == 3 and 0.0 < st["flagged_fraction"] < 0.05
    finally:
        g.close()


def test_routed_fp64_marching_sweeps_change_no_bit(monkeypatch, require_gpu):
    """Mixed routing in the plain FP64 form with marching y/z sweeps (3 tiles
    per block) equals single tiles bit for bit, flags and counts included."""
    run, q0, p = _setup("mpi_smooth3d", nx=16, ny=40, nz=33, cour=0.8)
    P = _params(5)
    P[M.OFF_B3] = 0.0
    print("[routed fp64 march] 16 x 40 x 33", flush=True)
    solvers = []
    for m in ("3", "1"):
        monkeypatch.setenv("HOWMHD_SWEEP_MARCH", m)
        g = H.GpuSolver(p)
        g.set_indicator(P, 0.5, enable=2)
        g.upload_q0(q0)
        g.init_faces()
        solvers.append(g)
    for _ in range(2):
        dt = solvers[1].tstep()
        for g in solvers:
            g.ssprk(dt)
    a, b = (H.active(g.q0()) for g in solvers)
    assert np.count_nonzero(a.view(np.int64) != b.view(np.int64)) == 0
    sa, sb = (g.indicator_stats() for g in solvers)
    print("stats", sa)
    assert sa == sb and 0 < sa["flagged"] < sa["total"]
