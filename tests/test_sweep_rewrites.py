"""The HIP sweep's rewrites of the characteristic flux change no bit (CPU).

The sweep kernel leaves out arithmetic of src/ssprk.cpp:602-994 that cannot
change the interface flux dF:
  * the sign flips of L rows / R columns 2, 4 (or 0, 6) by sign(B_t)
    (ssprk.cpp:789-818) cancel;
  * products with the structural zeros of L (rows 1, 5, columns 1, 6), of R
    (columns 1, 5 rows 0, 1; column 3 rows 4, 5) and of the CT weights -0, 0
    (ssprk.cpp:979-989) are left out;
  * rows 0/6 and 2/4 share their products with entries 4..6, and row 5's
    projection is row 1's shared partial sum (entries 0, 2, 3) minus row 1's
    last two products;
  * a - 2b is one fused multiply-add (2b is exact, so no test needs it).
Here one interface's dF and CT terms are computed both ways from the
restatement's eigensystem and WENO flux (the reference's operation order,
IEEE doubles in Python and C without contraction) and compared bit for bit,
signed zeros included, on random states and on states built to produce exact
zeros: vanishing transverse field, vanishing normal field, 1-D flows, signed
zero stencil entries.
"""
import math
import struct

import numpy as np
import pytest

import oracle_ctypes as O

GAM = 5.0 / 3.0


def bits(x):
    return struct.pack("<d", float(x))


def proj(row, X, cols):
    """sum over cols of row[j] * X[j], left to right (the reference's order)"""
    s = None
    for j in cols:
        p = float(row[j]) * float(X[j])
        s = p if s is None else s + p
    return s


def reference_flux(ul, ur, U, F, amx, zero_waves=None):
    """dF of one interface as src/ssprk.cpp:825-946 forms it (zero_waves:
    list that receives the number of waves whose flux is exactly +-0)."""
    L, R = O.eigensystem(GAM, ul, ur)
    dF = [0.0] * 7
    for m in range(7):
        Fs_k = [proj(L[m], F[k], range(7)) for k in range(6)]
        qs_k = [proj(L[m], U[k], range(7)) for k in range(6)]
        Fs = O.weno_wave(Fs_k, qs_k, amx[m])
        if zero_waves is not None and Fs == 0.0:
            zero_waves.append(m)
        for r in range(7):
            dF[r] = dF[r] + float(R[r, m]) * Fs
    return dF


def unflipped(ul, ur):
    """L and R of the restatement with the sign(B_t) flips undone (x +-1 is exact)."""
    L, R = O.eigensystem(GAM, ul, ur)
    rho = 0.5 * (ul[0] + ur[0])
    vx, vy, vz = (0.5 * (ul[i] + ur[i]) for i in (1, 2, 3))
    Bx, By, Bz = (0.5 * (ul[i] + ur[i]) for i in (4, 5, 6))
    HH = 0.5 * (ul[8] + ur[8])
    vv2 = vx * vx + vy * vy + vz * vz
    BB2 = Bx * Bx + By * By + Bz * Bz
    cs2 = (GAM - 1.0) * (HH - 0.5 * (vv2 + BB2 / rho))
    cs = math.sqrt(max(0.0, cs2))
    la = math.sqrt(max(0.0, Bx * Bx / rho))

    def sign1(x):
        return 1.0 if x >= 0.0 else -1.0

    s = sign1(By) if By != 0.0 else sign1(Bz)
    rows = (2, 4) if cs >= la else (0, 6)
    L, R = L.copy(), R.copy()
    for m in rows:
        L[m] *= s
        R[:, m] *= s
    return L, R


def rewritten_flux(ul, ur, U, F, amx):
    """dF as the HIP sweep forms it: passes (0,6), (1,5), (2,4), (3)."""
    L, R = unflipped(ul, ur)
    # the structure the passes rely on, bit for bit
    for a, b in ((0, 6), (2, 4)):
        assert all(bits(L[a, j]) == bits(L[b, j]) for j in (4, 5, 6))
    assert all(bits(L[1, j]) == bits(L[5, j]) for j in (0, 2, 3))
    assert all(bits(L[1, j]) == bits(-L[5, j]) for j in (4, 5))
    assert all(L[m, j] == 0.0 for m in (1, 5) for j in (1, 6))
    assert all(R[r, m] == 0.0 for m in (1, 5) for r in (0, 1)) and R[4, 3] == 0.0 and R[5, 3] == 0.0
    Fs = [None] * 7
    for a, b in ((0, 6), (2, 4)):
        pa, pb = [], []
        for X in (F, U):
            ya, yb = [], []
            for k in range(6):
                sh = [float(L[a, j]) * float(X[k][j]) for j in (4, 5, 6)]
                ya.append(proj(L[a], X[k], range(4)) + sh[0] + sh[1] + sh[2])
                yb.append(proj(L[b], X[k], range(4)) + sh[0] + sh[1] + sh[2])
            pa.append(ya)
            pb.append(yb)
        Fs[a] = O.weno_wave(pa[0], pa[1], amx[a])
        Fs[b] = O.weno_wave(pb[0], pb[1], amx[b])
    p1, p5 = [], []
    for X in (F, U):
        y1, y5 = [], []
        for k in range(6):
            part = proj(L[1], X[k], (0, 2, 3))
            f4 = float(L[1, 4]) * float(X[k][4])
            f5 = float(L[1, 5]) * float(X[k][5])
            y1.append(part + f4 + f5)
            y5.append(part - f4 - f5)
        p1.append(y1)
        p5.append(y5)
    Fs[1] = O.weno_wave(p1[0], p1[1], amx[1])
    Fs[5] = O.weno_wave(p5[0], p5[1], amx[5])
    Fs[3] = O.weno_wave([proj(L[3], F[k], range(7)) for k in range(6)],
                        [proj(L[3], U[k], range(7)) for k in range(6)], amx[3])
    dF = [0.0] * 7
    for m in range(7):
        for r in range(7):
            if (m in (1, 5) and r < 2) or (m == 3 and r in (4, 5)):
                continue
            dF[r] = dF[r] + float(R[r, m]) * Fs[m]
    return dF


def ct_terms(bv, dF4, zero_weights):
    """bsy = dF(4) + (-0 a + 9 b + 9 c - 0 d)/18 (ssprk.cpp:979-991), or without the zero-weight products"""
    a, b, c, d = bv
    if zero_weights:
        return dF4 + (-0.0 * a + 9.0 * b + 9.0 * c - 0.0 * d) / 18.0
    return dF4 + (9.0 * b + 9.0 * c) / 18.0


def make_case(rng, kind):
    """(ul, ur, U[6][7], F[6][7], amx[7], bn*vt of the four CT cells)."""
    def cell():
        rho = rng.uniform(0.3, 3.0)
        v = rng.normal(0, 0.7, 3)
        B = rng.normal(0, 1.0, 3)
        if kind in ("bt0", "oned"):
            B[1:] = 0.0
        if kind == "bn0":
            B[0] = 0.0
        if kind in ("oned", "negzero"):
            v[1:] = 0.0
        if kind == "negzero":
            B[1:] = -0.0
            v[1:] = -0.0
        if kind == "weak_bt":
            B[1:] *= 1e-9
        cs2 = rng.uniform(0.05, 2.0)
        vv2 = float(v @ v)
        H = 0.5 * (vv2 + float(B @ B) / rho) + cs2 / (GAM - 1.0)
        p = cs2 * rho / GAM
        return np.array([rho, v[0], v[1], v[2], B[0], B[1], B[2], p, H])

    ul, ur = cell(), cell()
    U = rng.normal(0, 1.0, (6, 7))
    F = rng.normal(0, 1.0, (6, 7))
    if kind in ("bt0", "oned", "negzero"):
        # transverse components of the stencil exactly zero (or -0): the
        # Alfven waves' projections vanish
        z = -0.0 if kind == "negzero" else 0.0
        U[:, [2, 3, 4, 5]] = z
        F[:, [2, 3, 4, 5]] = z
    if kind == "zerostencil":
        U[:] = 0.0
        F[:] = -0.0
    amx = np.abs(rng.normal(0, 1.0, 7))
    if kind == "amx0":
        amx[:] = 0.0
    bv = rng.normal(0, 1.0, 4)
    if kind in ("oned", "negzero", "bt0"):
        bv[:] = [0.0, -0.0, 0.0, -0.0] if kind == "negzero" else 0.0
    return ul, ur, U, F, amx, bv


KINDS = ["general", "bt0", "bn0", "oned", "negzero", "weak_bt", "zerostencil", "amx0"]


@pytest.mark.parametrize("kind", KINDS)
def test_rewritten_flux_is_bitwise_equal(kind):
    rng = np.random.default_rng(KINDS.index(kind) + 11)
    n = 300
    zero_waves = []
    for _ in range(n):
        ul, ur, U, F, amx, bv = make_case(rng, kind)
        ref = reference_flux(ul, ur, U, F, amx, zero_waves)
        new = rewritten_flux(ul, ur, U, F, amx)
        assert [bits(x) for x in new] == [bits(x) for x in ref], (kind, ref, new)
        for dF4 in (ref[4], ref[5]):
            assert bits(ct_terms(bv, dF4, False)) == bits(ct_terms(bv, dF4, True))
    if kind in ("oned", "negzero", "zerostencil"):
        # waves with an exactly zero flux: the signed-zero paths were taken
        assert len(zero_waves) >= n, len(zero_waves)


def test_planted_slip_is_caught():
    """The comparison detects a wrong sign in row 5's shared projection."""
    rng = np.random.default_rng(3)
    ul, ur, U, F, amx, _ = make_case(rng, "general")
    ref = reference_flux(ul, ur, U, F, amx)
    L, R = unflipped(ul, ur)
    bad = L.copy()
    bad[5, 4] = -bad[5, 4]   # row 5 built with row 1's sign at entry 4
    dF = [0.0] * 7
    for m in range(7):
        Fs = O.weno_wave([proj(bad[m], F[k], range(7)) for k in range(6)],
                         [proj(bad[m], U[k], range(7)) for k in range(6)], amx[m])
        for r in range(7):
            dF[r] = dF[r] + float(R[r, m]) * Fs
    assert [bits(x) for x in dF] != [bits(x) for x in ref]
