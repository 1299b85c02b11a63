"""NumPy restatement of the learned smoothness indicator (TEST INFRASTRUCTURE ONLY).

Loaded by tests/ and by nothing in the product.  It states the numerics the
HIP kernels implement (k_mlp_infer, k_mlp_grad, k_mlp_adam and the sweep's
feature and label code), so the GPU results can be checked against it:

  features  for each of rho, p, |B| over the 6-cell WENO stencil (cells
            t-3 .. t+2 of interface t), in float32:
            (x - min) / (max - min), or 0 when max - min <= 1e-5 max|x| + 1e-30
  forward   x, W1, W2, w3 rounded to bf16 (round to nearest even);
            z1 = x W1^T + b1; a1 = bf16(relu(z1)); z2 = a1 W2^T + b2;
            a2 = bf16(relu(z2)); z3 = b3 + a2 . w3
  loss      sum_i w_i (softplus(z3_i) - y_i z3_i) * scale, w_i = pos_weight if y_i = 1
  backward  dz3 = w (sigmoid(z3) - y) * scale; dw3 = sum dz3 a2; db3 = sum dz3;
            dz2 = bf16(dz3 w3) where a2 != 0; dW2 = dz2^T a1; db2 = sum dz2;
            dz1 = bf16(dz2 W2) where a1 != 0; dW1 = dz1^T x;  db1 = sum dz1
  label     1 when the largest |w_r - d_r| of the FP64 WENO5 nonlinear weights
            (both sides, all 7 waves; d = 0.1, 0.6, 0.3) exceeds 0.05
            (src/ssprk.cpp:903-908, 926-931 give the weights)

With quantize=False every rounding step is the identity and the arithmetic is
float64: that form is checked against finite differences of its own loss.
"""
from __future__ import annotations

import numpy as np

NIN, NH = 18, 32
FLAT = np.float32(1e-5)
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3, NPARAM = 0, 576, 608, 1632, 1664, 1696, 1697


def bf16(x) -> np.ndarray:
    """float32 -> bf16 -> float32, round to nearest even (finite inputs)."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(P):
    P = np.asarray(P)
    return (P[OFF_W1:OFF_B1].reshape(NH, NIN), P[OFF_B1:OFF_W2], P[OFF_W2:OFF_B2].reshape(NH, NH),
            P[OFF_B2:OFF_W3], P[OFF_W3:OFF_B3], P[OFF_B3])


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation (float32)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAM, dtype=np.float32)
    P[OFF_W1:OFF_B1] = rng.normal(0, np.sqrt(2.0 / NIN), NH * NIN)
    P[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2.0 / NH), NH * NH)
    P[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(1.0 / NH), NH)
    return P


def _q(quantize):
    return bf16 if quantize else (lambda a: np.asarray(a, dtype=np.float64))


def _f32(quantize):
    return (lambda a: np.asarray(a, dtype=np.float32)) if quantize else (lambda a: a)


def forward(P, X, quantize=True):
    """Returns (z3, cache).  Sums are formed in float64 and rounded to float32
    (the MFMA accumulates in float32 in an order of its own)."""
    q, f32 = _q(quantize), _f32(quantize)
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(P, dtype=np.float32 if quantize else np.float64))
    x = q(X).astype(np.float64)
    W1q, W2q, w3q = q(W1).astype(np.float64), q(W2).astype(np.float64), q(w3).astype(np.float64)
    z1 = f32(x @ W1q.T + np.asarray(b1, np.float64))
    a1 = q(np.maximum(z1, 0)).astype(np.float64)
    z2 = f32(a1 @ W2q.T + np.asarray(b2, np.float64))
    a2 = q(np.maximum(z2, 0)).astype(np.float64)
    z3 = f32(float(b3) + a2 @ w3q)
    return z3, dict(x=x, a1=a1, a2=a2, W2q=W2q, w3q=w3q)


def loss_and_grad(P, X, y, pos_weight=1.0, scale=1.0, quantize=True):
    q, f32 = _q(quantize), _f32(quantize)
    z3, c = forward(P, X, quantize)
    z = np.asarray(z3, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    w = np.where(y > 0.5, pos_weight, 1.0)
    loss = float(np.sum(w * (np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z))) - y * z)) * scale)
    p = 1.0 / (1.0 + np.exp(-z))
    dz3 = np.asarray(f32(w * (p - y) * scale), dtype=np.float64)
    G = np.zeros(NPARAM, dtype=np.float64)
    G[OFF_W3:OFF_B3] = dz3 @ c["a2"]
    G[OFF_B3] = dz3.sum()
    dz2 = np.where(c["a2"] != 0, q(f32(dz3[:, None] * c["w3q"][None, :])), 0.0).astype(np.float64)
    G[OFF_W2:OFF_B2] = (dz2.T @ c["a1"]).ravel()
    G[OFF_B2:OFF_W3] = dz2.sum(axis=0)
    dz1 = np.where(c["a1"] != 0, q(f32(dz2 @ c["W2q"])), 0.0).astype(np.float64)
    G[OFF_W1:OFF_B1] = (dz1.T @ c["x"]).ravel()
    G[OFF_B1:OFF_W2] = dz1.sum(axis=0)
    return loss, G


def adam(P, M, V, G, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (float32, bias-corrected), in place; as k_mlp_adam."""
    f = np.float32
    c1, c2 = f(1) - f(b1) ** f(t), f(1) - f(b2) ** f(t)
    M[:] = f(b1) * M + (f(1) - f(b1)) * G
    V[:] = f(b2) * V + (f(1) - f(b2)) * G * G
    P[:] = P - f(lr) * (M / c1) / (np.sqrt(V / c2) + f(eps))


def stencil_features(rho6, p6, b6) -> np.ndarray:
    """Features from (n, 6) float32 stencils of rho, p, |B| -> (n, 18) float32."""
    out = []
    for g in (rho6, p6, b6):
        g = np.asarray(g, dtype=np.float32)
        lo, hi = g.min(axis=1, keepdims=True), g.max(axis=1, keepdims=True)
        span = hi - lo
        thr = FLAT * np.maximum(np.abs(hi), np.abs(lo)) + np.float32(1e-30)
        inv = np.where(span > thr, np.float32(1) / np.where(span > 0, span, np.float32(1)), np.float32(0))
        out.append(((g - lo) * inv).astype(np.float32))
    return np.concatenate(out, axis=1)


def cell_primitives(qa, gam, rhomin, pgmin, d=0):
    """rho, p (float64, with the floors of src/ssprk.cpp:458-516) and |B|
    (float32) of the active cells (8, nx, ny, nz).  |B| is formed as the
    sweep along axis d forms it: B_n^2 + B_t1^2 + B_t2^2 summed in float64 in
    that order, rounded to float32, then a float32 square root."""
    rho = qa[0]
    v2 = (qa[1] ** 2 + qa[2] ** 2 + qa[3] ** 2) / rho ** 2
    BB2 = qa[4] ** 2 + qa[5] ** 2 + qa[6] ** 2
    pg = (gam - 1.0) * (qa[7] - 0.5 * (rho * v2 + BB2))
    fl = (rho < rhomin) | (pg < pgmin)
    rho_f = np.where(fl, np.maximum(rho, rhomin), rho)
    pg_f = np.where(fl, np.maximum(pg, pgmin), pg)
    bn, b1, b2 = (qa[4 + (d + k) % 3] for k in range(3))
    bmag = np.sqrt((bn * bn + b1 * b1 + b2 * b2).astype(np.float32))
    return rho_f, pg_f, bmag


def _axis_map(a, n, per):
    if per:
        return np.mod(a, n)
    return np.clip(a, 0, n - 1)


def interface_features(qa, d, per, gam, rhomin=1e-12, pgmin=1e-12):
    """Features of every interface of the sweep along axis d (0 x, 1 y, 2 z) of
    the active state qa (8, nx, ny, nz), in the sample order of
    howmhd_gpu_collect_samples: x lines (z, y, t), y (z, t, x), z (t, y, x)."""
    rho, pg, bm = (np.asarray(a, dtype=np.float32) for a in cell_primitives(qa, gam, rhomin, pgmin, d))
    n = qa.shape[1 + d]
    tbeg = 1 if per else 0
    ts = np.arange(tbeg, n + 1)
    cells = _axis_map(ts[:, None] + np.arange(-3, 3)[None, :], n, per)   # (nint, 6)
    feats = []
    for g in (rho, pg, bm):
        st = np.take(g, cells, axis=d)          # axis d replaced by (nint, 6)
        # -> (x?, ...) arrange to sample order with a trailing 6 axis
        st = np.moveaxis(st, d + 1, -1)          # stencil axis last; interface axis at d
        # st axes: positions (x, y, z) with d replaced by t; reorder to (z, y, x)
        st = np.transpose(st, (2, 1, 0, 3))
        feats.append(st.reshape(-1, 6))
    return stencil_features(*feats)


def interface_labels(qa, d, per, gam, line_maxdev, rhomin=1e-12, pgmin=1e-12):
    """Labels (0/1) and max deviations of every interface along axis d in the
    same order, from the restatement's per-line weights (line_maxdev)."""
    perm = [(0, 1, 2, 3, 4, 5, 6, 7), (0, 2, 3, 1, 5, 6, 4, 7), (0, 3, 1, 2, 6, 4, 5, 7)][d]
    n = qa.shape[1 + d]
    tbeg = 1 if per else 0
    idx = _axis_map(np.arange(-3, n + 3), n, per)
    qp = np.take(qa[list(perm)], idx, axis=1 + d)        # (8, ..., n+6 along d, ...)
    qp = np.moveaxis(qp, 1 + d, -1)                       # (8, a, b, n+6)
    # remaining spatial axes in (x, y, z) order without d
    shp = qp.shape[1:3]
    dev = np.zeros(shp + (n + 1,))
    for a in range(shp[0]):
        for b in range(shp[1]):
            dev[a, b] = line_maxdev(qp[:, a, b, :], gam, rhomin, pgmin)
    dev = dev[..., tbeg:]
    # dev axes: the two transverse axes in increasing axis order, then t.
    # Reorder to the sample order (z, y, x) with t at axis d.
    full = np.moveaxis(dev, -1, d)                        # (x, y, z) with t at d
    full = np.transpose(full, (2, 1, 0)).reshape(-1)
    return (full > 0.05).astype(np.float32), full
