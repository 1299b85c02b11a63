"""Pins the restatement of the learned-indicator form (CPU only).

  upwind5          == the textbook 5th-order upwind reconstruction of the LLF
                      split fluxes (Jiang & Shu 1996, ideal weights)
  linear route     a whole step of the restatement with every interface on
                      upwind5 converges at better than 4th order on a smooth
                      wave and stays conservative
  labels           line_maxdev == max |w - d| recomputed from the
                      restatement's own eigensystem and weno_wave weights;
                      smooth data -> 0, a jump -> 1
  bf16             == torch's float32 -> bfloat16 conversion (round to nearest even)
  MLP backward     float64 form == central finite differences of its loss
  features         constant / linear stencils, affine invariance
"""
import math

import numpy as np
import pytest

import howmhd as H
import mlp_oracle as M
import oracle_ctypes as O

GAM = 5.0 / 3.0


def upwind5_textbook(F, U, alpha):
    """f+ = (F + alpha U)/2 reconstructed from the left of the face between
    cells 2 and 3 with (2, -13, 47, 27, -3)/60, f- from the right with
    (-3, 27, 47, -13, 2)/60."""
    fp = 0.5 * (F + alpha * U)
    fm = 0.5 * (F - alpha * U)
    left = (2 * fp[0] - 13 * fp[1] + 47 * fp[2] + 27 * fp[3] - 3 * fp[4]) / 60
    right = (-3 * fm[1] + 27 * fm[2] + 47 * fm[3] - 13 * fm[4] + 2 * fm[5]) / 60
    return left + right


@pytest.mark.parametrize("seed", range(5))
def test_upwind5_matches_textbook(seed):
    rng = np.random.default_rng(seed)
    for _ in range(50):
        F = rng.normal(size=6)
        U = rng.normal(size=6)
        a = abs(rng.normal())
        assert abs(O.upwind5(F, U, a) - upwind5_textbook(F, U, a)) <= 1e-13 * (1 + np.abs(F).sum() + a * np.abs(U).sum())


def test_upwind5_is_weno_with_ideal_weights():
    """On smooth data the WENO weights approach the ideal ones and the two
    fluxes agree to the weights' deviation."""
    x = np.arange(6) * 1e-3
    F = 1 + np.sin(x)
    U = 2 + np.cos(x)
    f_weno, wp, wm = O.weno_wave(F, U, 0.7, weights=True)
    assert np.allclose(wp, [0.1, 0.6, 0.3], atol=1e-3) and np.allclose(wm, [0.1, 0.6, 0.3], atol=1e-3)
    assert abs(O.upwind5(F, U, 0.7) - f_weno) <= 1e-9


def _entropy_wave(n):
    L = (3.0, 1.5, 1.5)
    xs = [(np.arange(n) + 0.5) * L[d] / n for d in range(3)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    rho = 1 + 0.2 * np.sin(2 * np.pi * (X + 2 * Y + 2 * Z) / 3)
    v = np.array([1.0, 2.0, 2.0]) / 3.0
    B = np.array([0.3, -0.2, 0.4])
    E = 1.0 / (GAM - 1) + 0.5 * rho * (v @ v) + 0.5 * (B @ B)
    q0 = np.zeros((8, n + 6, n + 6, n + 6))
    for m, f in enumerate([rho, rho * v[0], rho * v[1], rho * v[2], B[0] + 0 * X, B[1] + 0 * X, B[2] + 0 * X, E]):
        q0[m, 3:-3, 3:-3, 3:-3] = f
    return q0


def _solver(q0, n, route):
    run = H.parse(H.make_input(problem="uniform", nx=n, ny=n, nz=n, xsize=3.0, ysize=1.5, zsize=1.5, cour=0.4,
                               nt=1, tend=1.0))
    s = O.RefSolver(H.gpu_params(run))
    s.set_route(route)
    s.upload_q0(q0)
    s.init_faces()
    return s


def test_linear_route_converges_and_conserves():
    errs = []
    for n in (12, 24):
        q0 = _entropy_wave(n)
        s = _solver(q0, n, 1)
        t = 0.0
        while t < 1.0:
            dt = min(s.tstep(), 1.0 - t)
            s.ssprk(dt)
            t += dt
        q = s.q0()
        errs.append(np.abs(H.active(q)[0] - H.active(q0)[0]).mean())
        a0, a1 = H.active(q0).reshape(8, -1).sum(1), H.active(q).reshape(8, -1).sum(1)
        assert np.all(np.abs(a1 - a0) <= 1e-11 * np.abs(H.active(q0)).reshape(8, -1).sum(1).max())
    order = math.log2(errs[0] / errs[1])
    print("linear route entropy wave L1(rho):", errs, "order", order)
    assert order > 4.0


def test_linear_route_approaches_weno_on_smooth_data():
    """On a smooth wave the two routes differ by a high-order term: the
    difference after one step falls by more than 2^4 from 16^3 to 32^3 and is
    a small part of the step's own change."""
    diffs = []
    for n in (16, 32):
        q0 = _entropy_wave(n)
        a, b = _solver(q0, n, 0), _solver(q0, n, 1)
        for s in (a, b):
            s.ssprk(0.01)
        diffs.append(H.rel_l1(b.q0(), a.q0())[0])
        change = H.rel_l1(a.q0(), q0)[0]
    print("linear vs WENO after one step:", diffs, "step change", change)
    assert diffs[0] / diffs[1] > 16
    assert diffs[1] <= 1e-4 * change


def _line(rho_l, rho_r, nn=16):
    rho = np.where(np.arange(nn + 6) < (nn + 6) // 2, rho_l, rho_r).astype(float)
    v = np.array([0.1, 0.0, 0.0])
    B = np.array([0.75, 1.0, 0.0])
    p = rho.copy()
    E = p / (GAM - 1) + 0.5 * rho * (v @ v) + 0.5 * (B @ B)
    return np.stack([rho, rho * v[0], rho * v[1], rho * v[2], B[0] + 0 * rho, B[1] + 0 * rho, B[2] + 0 * rho, E])


def test_line_maxdev_recomputed_from_weights():
    """maxdev from the restatement == max |w - d| over waves and sides of
    weno_wave, with the projections formed here from the eigensystem."""
    rng = np.random.default_rng(3)
    nn = 10
    q = _line(1.0, 0.4, nn)
    q[0] *= 1 + 0.05 * rng.normal(size=nn + 6)
    q[7] *= 1 + 0.05 * rng.normal(size=nn + 6)
    dev = O.line_maxdev(q, GAM)
    rho = q[0]
    v = q[1:4] / rho
    B = q[4:7]
    p = (GAM - 1) * (q[7] - 0.5 * rho * (v * v).sum(0) - 0.5 * (B * B).sum(0))
    H_ = (q[7] + p) / rho
    prim = np.stack([rho, v[0], v[1], v[2], B[0], B[1], B[2], p, H_])
    Bn = B[0]
    pt = p + 0.5 * (B * B).sum(0)
    Fphys = np.stack([q[0] * v[0], q[1] * v[0] + pt - Bn * Bn, q[2] * v[0] - Bn * B[1], q[3] * v[0] - Bn * B[2],
                      B[1] * v[0] - Bn * v[1], B[2] * v[0] - Bn * v[2],
                      (q[7] + pt) * v[0] - Bn * (v * B).sum(0)])
    U7 = q[[0, 1, 2, 3, 5, 6, 7]]
    cs2 = GAM * p / rho
    bb = (B * B).sum(0) / rho
    bn2 = Bn * Bn / rho
    root = np.sqrt((bb + cs2) ** 2 - 4 * bn2 * cs2)
    lf, la, ls = np.sqrt(0.5 * (bb + cs2 + root)), np.sqrt(bn2), np.sqrt(0.5 * (bb + cs2 - root))
    lam = np.stack([v[0] - lf, v[0] - la, v[0] - ls, v[0], v[0] + ls, v[0] + la, v[0] + lf])
    for i in range(nn + 1):
        L, _ = O.eigensystem(GAM, prim[:, i + 2], prim[:, i + 3])
        want = 0.0
        for m in range(7):
            amx = np.abs(lam[m, i:i + 6]).max()
            _, wp, wm = O.weno_wave(L[m] @ Fphys[:, i:i + 6], L[m] @ U7[:, i:i + 6], amx, weights=True)
            for w in (wp, wm):
                want = max(want, abs(w[0] - 0.1), abs(w[2] - 0.3), abs(1 - w[0] - w[2] - 0.6))
        assert abs(dev[i] - want) <= 1e-9, (i, dev[i], want)


def test_labels_smooth_and_jump():
    smooth = O.line_maxdev(_line(1.0, 1.0), GAM)
    assert np.all(smooth <= 1e-12)
    jump = O.line_maxdev(_line(1.0, 0.125), GAM)
    nn = 16
    mid = (nn + 6) // 2 - 2          # interface index of the jump
    assert jump[mid] > 0.05
    assert np.all(jump[:mid - 3] <= 1e-12) and np.all(jump[mid + 4:] <= 1e-12)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=10000).astype(np.float32) * np.float32(10.0) ** rng.integers(-8, 8, 10000),
                        np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0, -0.0], dtype=np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(M.bf16(x), want)


def test_mlp_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    P = M.init_params(2).astype(np.float64)
    P[M.OFF_B1:M.OFF_W2] = rng.normal(0, 0.1, 32)
    P[M.OFF_B2:M.OFF_W3] = rng.normal(0, 0.1, 32)
    X = rng.uniform(0, 1, (40, 18))
    y = (rng.uniform(size=40) > 0.6).astype(float)
    loss, G = M.loss_and_grad(P, X, y, pos_weight=3.0, scale=0.5, quantize=False)
    h = 1e-6
    idx = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 20), np.arange(M.OFF_B1, M.OFF_B1 + 4),
                          rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 20), np.arange(M.OFF_B2, M.OFF_B2 + 4),
                          np.arange(M.OFF_W3, M.OFF_W3 + 6), [M.OFF_B3]])
    for j in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[j] += h
        Pm[j] -= h
        lp, _ = M.loss_and_grad(Pp, X, y, 3.0, 0.5, quantize=False)
        lm, _ = M.loss_and_grad(Pm, X, y, 3.0, 0.5, quantize=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - G[j]) <= 1e-6 * (1 + abs(fd)), (j, fd, G[j])


def test_stencil_features():
    n = 4
    const = np.full((n, 6), 2.5, np.float32)
    ramp = np.tile(np.arange(6, dtype=np.float32), (n, 1))
    f = M.stencil_features(const, ramp, 3 * ramp + 7)
    assert np.all(f[:, :6] == 0)
    np.testing.assert_allclose(f[:, 6:12], np.tile(np.arange(6) / 5, (n, 1)), atol=1e-7)
    np.testing.assert_allclose(f[:, 12:], np.tile(np.arange(6) / 5, (n, 1)), atol=1e-6)


def test_interface_features_order():
    """Sample order and stencils of interface_features on a 3-D index field."""
    nx, ny, nz = 5, 4, 3
    qa = np.zeros((8, nx, ny, nz))
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    qa[0] = 1 + X + 10 * Y + 100 * Z           # rho encodes the cell
    qa[7] = 10.0                               # p > 0
    for d, per in ((0, True), (1, False), (2, True)):
        f = M.interface_features(qa, d, per, GAM)
        n = (nx, ny, nz)[d]
        nint = n if per else n + 1
        assert f.shape == (nx * ny * nz // n * nint, 18)
        # rho features of the first sample (interface tbeg of the first line)
        t = 1 if per else 0
        cells = [(c % n) if per else min(max(c, 0), n - 1) for c in range(t - 3, t + 3)]
        r = np.array([1 + (c if d == 0 else 0) + 10 * (c if d == 1 else 0) + 100 * (c if d == 2 else 0)
                      for c in cells], np.float32)
        np.testing.assert_allclose(f[0, :6], (r - r.min()) / (r.max() - r.min()), atol=1e-6)
