"""Runs the reference configurations that are reports rather than the headline
benchmark (needs a GPU) and prints one JSON document.

    python how-mhd-kokkos_amd/report_configs.py --out report.json [--only cpaw,turb384,blast512]

  cpaw      3-D circularly polarised Alfven wave (problem cpaw3d) on
            [0,3]x[0,1.5]x[0,1.5], k along (1,2,2)/3, rho 1, p 0.1, B_par 1,
            B_perp 0.1, gam 5/3, cour 0.4, one period: L1 error of every
            conserved variable against the exact solution (the initial state)
            at 32^3, 64^3, 128^3, 256^3 and the observed orders (plain FP64).
  turb384   decaying turbulence (turbulence3d, seed 1) at 384^3, 100 steps:
            the FP32 form against the plain FP64 path on the same dt sequence:
            density, energy and magnetic-energy errors, CT divergence, speedup.
  blast512  MHD blast (mhd_blast3d) at 512^3, 50 steps: zone-cycles/s of the
            plain FP64 path and of the FP32 + indicator form.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
sys.path.insert(0, str(PKG))
import howmhd as H  # noqa: E402


def log(*a):
    print(*a, flush=True)


def _run(problem, **kw):
    kw.setdefault("nt", 1)
    kw.setdefault("tend", 100.0)
    return H.parse(H.make_input(problem=problem, **kw))


def _solver(run, precision="fp64"):
    p = H.gpu_params(run)
    p.precision = H.PREC[precision]
    return H.GpuSolver(p)


def cpaw(sizes=(32, 64, 128, 256)):
    out = {"sizes": list(sizes), "l1": {}, "order": {}}
    names = ["rho", "Mx", "My", "Mz", "Bx", "By", "Bz", "E"]
    errs = []
    for n in sizes:
        run = _run("cpaw3d", nx=n, ny=n, nz=n, xsize=3.0, ysize=1.5, zsize=1.5, cour=0.4, tend=1.0)
        q0 = H.init_problem(run)
        g = _solver(run)
        g.upload_q0(q0)
        g.init_faces()
        t, steps = 0.0, 0
        t0 = time.perf_counter()
        while t < 1.0:
            dt = min(g.tstep(), 1.0 - t)
            g.ssprk(dt)
            t += dt
            steps += 1
        el = time.perf_counter() - t0
        a, b = H.active(g.q0()), H.active(q0)
        e = np.abs(a - b).reshape(8, -1).mean(axis=1)
        errs.append(e)
        out["l1"][n] = {nm: float(v) for nm, v in zip(names, e)}
        log(f"[cpaw] {n}^3: {steps} steps, {el:.1f} s, L1(By) {e[5]:.4e}, L1(rho) {e[0]:.4e}")
        g.close()
    for i in range(len(sizes) - 1):
        key = f"{sizes[i]}->{sizes[i + 1]}"
        out["order"][key] = {nm: (math.log2(errs[i][m] / errs[i + 1][m]) if errs[i + 1][m] > 0 else None)
                             for m, nm in enumerate(names) if errs[i][m] > 1e-14}
    return out


def turb384(n=384, steps=100):
    run = _run("turbulence3d", nx=n, ny=n, nz=n, cour=0.4, seed=1)
    q0 = H.init_problem(run)
    res = {"grid": n, "steps": steps}
    speeds = {}
    for prec in ("fp64", "fp32"):
        g = _solver(run, prec)
        g.upload_q0(q0)
        g.init_faces()
        g.ssprk(g.tstep())
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            g.ssprk(g.tstep())
        g.synchronize()
        speeds[prec] = n ** 3 * 10 / (time.perf_counter() - t0)
        g.close()
    g64, g32 = _solver(run, "fp64"), _solver(run, "fp32")
    for g in (g64, g32):
        g.upload_q0(q0)
        g.init_faces()
    del q0
    t = 0.0
    for _ in range(steps):
        dt = g64.tstep()
        g64.ssprk(dt)
        g32.ssprk(dt)
        t += dt
    a64, a32 = g64.q0(), g32.q0()
    l1 = H.rel_l1(a32, a64)
    em = lambda a: 0.5 * float((H.active(a)[4:7] ** 2).sum())
    ek = lambda a: 0.5 * float(((H.active(a)[1:4] ** 2).sum(axis=0) / H.active(a)[0]).sum())
    res.update({
        "t_end": t,
        "zone_cycles_per_s": {k: round(v, 1) for k, v in speeds.items()},
        "speedup_fp32_vs_fp64": round(speeds["fp32"] / speeds["fp64"], 3),
        "rel_l1": {nm: float(v) for nm, v in zip(["rho", "Mx", "My", "Mz", "Bx", "By", "Bz", "E"], l1)},
        "magnetic_energy_rel_err": abs(em(a32) - em(a64)) / em(a64),
        "kinetic_energy_rel_err": abs(ek(a32) - ek(a64)) / ek(a64),
        "divb_ct_max": {"fp64": g64.diag().max_divb_ct, "fp32": g32.diag().max_divb_ct},
    })
    log(f"[turb384] {json.dumps(res)}")
    return res


def blast512(n=512, steps=50):
    run = _run("mhd_blast3d", nx=n, ny=n, nz=n, cour=0.4)
    q0 = H.init_problem(run)
    res = {"grid": n, "steps": steps}
    weights = PKG / "weights" / "indicator_mlp.txt"
    for form in ("fp64", "fp32+indicator"):
        g = _solver(run, "fp64" if form == "fp64" else "fp32")
        if form != "fp64":
            P, thr = H.load_indicator(weights)
            g.set_indicator(P, thr, enable=1)
        g.upload_q0(q0)
        g.init_faces()
        g.ssprk(g.tstep())
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.ssprk(g.tstep())
        g.synchronize()
        el = time.perf_counter() - t0
        res[form] = {"zone_cycles_per_s": round(n ** 3 * steps / el, 1), "ms_per_step": round(el / steps * 1e3, 2),
                     "divb_ct_max": g.diag().max_divb_ct}
        log(f"[blast512] {form}: {res[form]}")
        g.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--only", default="cpaw,turb384,blast512")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("report_configs.py needs a GPU")
    rep = {}
    for name in a.only.split(","):
        rep[name] = {"cpaw": cpaw, "turb384": turb384, "blast512": blast512}[name]()
    if a.out:
        Path(a.out).write_text(json.dumps(rep, indent=1))
    return rep


if __name__ == "__main__":
    main()
