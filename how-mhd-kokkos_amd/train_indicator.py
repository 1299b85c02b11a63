"""Train the learned smoothness indicator on the GPU and write its weights file.

    python how-mhd-kokkos_amd/train_indicator.py --out how-mhd-kokkos_amd/weights/indicator_mlp.txt

Data: seeded Orszag-Tang (2-D), MHD blast (3-D) and decaying turbulence (3-D)
runs of the plain FP64 method.  At snapshot steps every interface of every
sweep direction becomes one sample (howmhd_gpu_collect_samples): 18 features
(6-cell stencils of rho, p, |B|, each mapped to [0, 1]) and the label "the
FP64 WENO5 nonlinear weights deviate from the ideal ones by more than 0.05".

Training: weighted binary cross-entropy (positives weighted by --pos-weight),
Adam, minibatches drawn from a device-resident shuffle; the forward, backward
and Adam steps are the library's HIP kernels (howmhd_mlp_grad,
howmhd_mlp_adam); torch only holds the device buffers and shuffles indices.

Threshold: the largest flag probability at which the held-out recall of the
positive class stays >= --recall.

The weights file is text: a header line "howmhd-mlp 1697 <threshold>", then
the 1697 FP32 parameters W1[32][18], b1[32], W2[32][32], b2[32], w3[32], b3,
one per line; lines starting with '#' are comments.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent))
import howmhd as H  # noqa: E402

NPARAM = H.MLP_NPARAM
OFF_B1, OFF_W2, OFF_B2, OFF_W3 = 576, 608, 1632, 1664

# (problem, input keys, hydro steps, snapshot every k steps)
RUNS = [
    ("orszag_tang", dict(nx=256, ny=256, nz=1, cour=0.8), 600, 40),
    ("mhd_blast3d", dict(nx=64, ny=64, nz=64, cour=0.4), 120, 30),
    ("turbulence3d", dict(nx=64, ny=64, nz=64, cour=0.4, seed=11), 150, 50),
    ("turbulence3d", dict(nx=64, ny=64, nz=64, cour=0.4, seed=12), 150, 50),
]


def log(*a):
    print(*a, flush=True)


def collect(runs, torch, max_samples):
    feats, labels = [], []
    total = 0
    for problem, kw, steps, every in runs:
        kw = dict(kw)
        kw.setdefault("nt", 1)
        kw.setdefault("tend", 100.0)
        run = H.parse(H.make_input(problem=problem, **kw))
        g = H.GpuSolver(H.gpu_params(run))
        g.upload_q0(H.init_problem(run))
        g.init_faces()
        t0 = time.time()
        for step in range(1, steps + 1):
            g.ssprk(g.tstep())
            if step % every:
                continue
            for d in range(3):
                n = g.n_samples(d)
                if n == 0:
                    continue
                X = torch.empty((n, H.MLP_NFEAT), dtype=torch.float32, device="cuda")
                y = torch.empty(n, dtype=torch.float32, device="cuda")
                g.collect_samples(d, X, y)
                feats.append(X)
                labels.append(y)
                total += n
        g.close()
        log(f"collect {problem} {kw}: {steps} steps in {time.time() - t0:.1f} s, samples so far {total}")
        if total >= max_samples:
            break
    X = torch.cat(feats)
    y = torch.cat(labels)
    return X, y


def init_params(seed):
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAM, dtype=np.float32)
    P[0:OFF_B1] = rng.normal(0, np.sqrt(2.0 / 18), 576)
    P[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2.0 / 32), 1024)
    P[OFF_W3:OFF_W3 + 32] = rng.normal(0, np.sqrt(1.0 / 32), 32)
    return P


def train(X, y, torch, epochs, batch, lr, pos_weight, seed):
    n = X.shape[0]
    P = torch.from_numpy(init_params(seed)).cuda()
    Mm = torch.zeros_like(P)
    V = torch.zeros_like(P)
    G = torch.zeros_like(P)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    step = 0
    nb = n // batch
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda", generator=gen)
        ep_loss = 0.0
        lr_ep = lr * (0.5 ** (ep // max(1, epochs // 3)))
        loss.zero_()
        for b in range(nb):
            idx = perm[b * batch:(b + 1) * batch]
            Xb = X.index_select(0, idx).contiguous()
            yb = y.index_select(0, idx).contiguous()
            G.zero_()
            H.mlp_grad(Xb, yb, P, pos_weight=pos_weight, scale=1.0 / batch, G=G, loss=loss)
            step += 1
            H.mlp_adam(P, Mm, V, G, step, lr_ep)
        torch.cuda.synchronize()
        ep_loss = float(loss.item()) / nb
        log(f"epoch {ep}: lr {lr_ep:.2e} mean weighted BCE {ep_loss:.5f}")
    return P


def evaluate(P, X, y, torch, recall):
    z = H.mlp_infer(X, P)
    pos = y > 0.5
    zp = torch.sort(z[pos]).values
    k = int(np.floor((1.0 - recall) * zp.numel()))
    zthr = float(zp[k].item()) if zp.numel() else 0.0
    # flag when z > zthr' with zthr' just below the k-th positive logit
    zthr = np.nextafter(np.float32(zthr), np.float32(-np.inf))
    flag = z > float(zthr)
    tp = int((flag & pos).sum())
    stats = {
        "samples": int(y.numel()),
        "positive_fraction": float(pos.float().mean()),
        "flagged_fraction": float(flag.float().mean()),
        "recall": tp / max(1, int(pos.sum())),
        "precision": tp / max(1, int(flag.sum())),
        "logit_threshold": float(zthr),
        "threshold": float(1.0 / (1.0 + np.exp(-float(zthr)))),
    }
    return stats


def write_weights(path: Path, P: np.ndarray, threshold: float, meta: dict):
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w") as f:
        f.write(f"# learned smoothness indicator, 18-32-32-1 MLP (train_indicator.py)\n")
        for k, v in meta.items():
            f.write(f"# {k}: {json.dumps(v)}\n")
        f.write(f"howmhd-mlp {NPARAM} {threshold:.9g}\n")
        for v in P:
            f.write(f"{float(v):.9g}\n")


def read_weights(path) -> tuple[np.ndarray, float]:
    """(params, threshold) of a weights file (same format the host library reads)."""
    vals, thr = [], None
    for line in Path(path).read_text().splitlines():
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        if thr is None:
            tag, n, t = line.split()
            assert tag == "howmhd-mlp" and int(n) == NPARAM
            thr = float(t)
            continue
        vals.append(float(line))
    P = np.array(vals, dtype=np.float32)
    assert P.shape == (NPARAM,)
    return P, thr


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=str(Path(__file__).resolve().parent / "weights" / "indicator_mlp.txt"))
    ap.add_argument("--samples", type=float, default=1e7)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--pos-weight", type=float, default=4.0)
    ap.add_argument("--recall", type=float, default=0.995)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--scale", type=float, default=1.0, help="grid scale factor for the collection runs")
    a = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("train_indicator.py needs a GPU")
    runs = []
    for problem, kw, steps, every in RUNS:
        kw = dict(kw)
        for k in ("nx", "ny", "nz"):
            if kw[k] > 1:
                kw[k] = max(8, int(kw[k] * a.scale))
        runs.append((problem, kw, max(every, int(steps * a.scale)), max(1, int(every * a.scale))))
    t0 = time.time()
    X, y = collect(runs, torch, a.samples)
    log(f"collected {X.shape[0]} samples ({float(y.mean()):.3f} positive) in {time.time() - t0:.1f} s")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(a.seed + 1)
    perm = torch.randperm(X.shape[0], device="cuda", generator=gen)
    nval = max(1, X.shape[0] // 20)
    Xv, yv = X[perm[:nval]].contiguous(), y[perm[:nval]].contiguous()
    Xt, yt = X[perm[nval:]].contiguous(), y[perm[nval:]].contiguous()
    del X, y
    t1 = time.time()
    P = train(Xt, yt, torch, a.epochs, a.batch, a.lr, a.pos_weight, a.seed)
    log(f"trained in {time.time() - t1:.1f} s")
    st_train = evaluate(P, Xt[:2_000_000], yt[:2_000_000], torch, a.recall)
    st = evaluate(P, Xv, yv, torch, a.recall)
    log("held-out:", json.dumps(st))
    meta = {"runs": [[p, kw, s, e] for p, kw, s, e in runs], "samples": int(Xt.shape[0] + nval),
            "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "pos_weight": a.pos_weight, "seed": a.seed,
            "target_recall": a.recall, "held_out": st, "train_subset": st_train}
    write_weights(Path(a.out), P.cpu().numpy(), st["threshold"], meta)
    log(f"wrote {a.out}")
    return st


if __name__ == "__main__":
    main()
