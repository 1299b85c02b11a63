"""Evaluate the learned smoothness indicator against plain WENO (needs a GPU).

    python how-mhd-kokkos_amd/eval_indicator.py --out report.json [--cases rotor512,turb256,blast256]

For each case the plain FP64 method (reference rounding) is the reference run: its dt
sequence drives every form from the same initial state, and every form's
density and energy are compared with it at the end (relative L1 over active
cells).  Forms:

  fp32                 FP32 sweeps/CT, WENO at every interface
  fp32+indicator       the trained network routes unflagged interfaces to
                       linear upwind-5 (the product form: with the
                       plain-WENO fallback for steps whose wavefronts are
                       nearly all flagged)
  fp32+linear          every interface on linear upwind-5 (no indicator)
  fp32+untrained       a randomly initialised network whose threshold gives
                       the trained network's flagged fraction (ablation)
  fp64+indicator       the trained network on the FP64 method

Speed: zone-cycles/s of each form over --time-steps steps from the initial
state (the same state for every form), timed with the device synchronised;
with --speed-end also from the reference run's final q(0) and faces (the
developed flow: zone_cycles_per_s_end, speedup_vs_fp64_end).  The accuracy
pass over the whole dt sequence is timed too (ms_per_step_run,
speedup_vs_fp64_run; no CFL reduction in that loop).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
sys.path.insert(0, str(PKG))
import howmhd as H  # noqa: E402

WEIGHTS = PKG / "weights" / "indicator_mlp.txt"

CASES = {
    # MHD rotor of the indicator evaluation: rho 10 inside r < 0.1 tapered to 1
    # at r = 0.115, omega 20, p 1, Bx 5/sqrt(4 pi), gam 1.4, tend 0.15
    "rotor512": ("mhd_rotor", dict(nx=512, ny=512, nz=1, gam=1.4, cour=0.8), dict(tend=0.15)),
    "turb256": ("turbulence3d", dict(nx=256, ny=256, nz=256, cour=0.4, seed=1), dict(steps=50)),
    "blast256": ("mhd_blast3d", dict(nx=256, ny=256, nz=256, cour=0.4), dict(steps=50)),
    # the rotor at 16x the cells: sweep launches of many dispatch rounds
    # instead of about one
    "rotor2048": ("mhd_rotor", dict(nx=2048, ny=2048, nz=1, gam=1.4, cour=0.8), dict(tend=0.15)),
    # small versions (tests, quick looks)
    "rotor128": ("mhd_rotor", dict(nx=128, ny=128, nz=1, gam=1.4, cour=0.8), dict(tend=0.05)),
    "turb64": ("turbulence3d", dict(nx=64, ny=64, nz=64, cour=0.4, seed=1), dict(steps=30)),
}


def log(*a):
    print(*a, flush=True)


def _solver(run, precision, P=None, thr=0.5, enable=1):
    p = H.gpu_params(run)
    p.precision = H.PREC[precision]
    g = H.GpuSolver(p)
    if P is not None:
        g.set_indicator(P, thr, enable=enable)
    return g


def untrained_params(seed=123):
    rng = np.random.default_rng(seed)
    P = np.zeros(H.MLP_NPARAM, dtype=np.float32)
    P[0:576] = rng.normal(0, np.sqrt(2.0 / 18), 576)
    P[576:608] = rng.normal(0, 0.1, 32)
    P[608:1632] = rng.normal(0, np.sqrt(2.0 / 32), 1024)
    P[1632:1664] = rng.normal(0, 0.1, 32)
    P[1664:1696] = rng.normal(0, np.sqrt(1.0 / 32), 32)
    return P


def matched_threshold(ref_run, P_trained, thr_trained, P_other):
    """Threshold of P_other that flags the same fraction of the reference run's
    current interfaces as the trained network does."""
    import torch
    feats = []
    for d in range(3):
        n = ref_run.n_samples(d)
        if n == 0:
            continue
        X = torch.empty((n, H.MLP_NFEAT), dtype=torch.float32, device="cuda")
        y = torch.empty(n, dtype=torch.float32, device="cuda")
        ref_run.collect_samples(d, X, y)
        feats.append(X)
    X = torch.cat(feats)
    zt = H.mlp_infer(X, torch.from_numpy(P_trained).cuda())
    frac = float((zt > float(np.log(thr_trained / (1 - thr_trained)))).float().mean())
    zo = H.mlp_infer(X, torch.from_numpy(P_other).cuda())
    k = min(max(int(round((1 - frac) * zo.numel())), 0), zo.numel() - 1)
    zthr = float(torch.sort(zo).values[k])
    thr = 1.0 / (1.0 + np.exp(-zthr))
    return float(np.clip(thr, 1e-6, 1 - 1e-6)), frac


def run_case(name, time_steps=10, forms=None, speed_end=False):
    import torch
    problem, kw, stop = CASES[name]
    kw = dict(kw)
    kw.setdefault("nt", 1)
    kw.setdefault("tend", stop.get("tend", 100.0))
    run = H.parse(H.make_input(problem=problem, **kw))
    q0 = H.init_problem(run)
    zones = run.gnx * run.gny * run.gnz
    P, thr = H.load_indicator(WEIGHTS)

    # reference trajectory: plain FP64 dt sequence
    ref_run = _solver(run, "fp64")
    ref_run.upload_q0(q0)
    ref_run.init_faces()
    dts, t = [], 0.0
    tend = stop.get("tend")
    nsteps = stop.get("steps")
    P_un = untrained_params()
    thr_un = None
    while True:
        if nsteps is not None and len(dts) >= nsteps:
            break
        if tend is not None and t >= tend * (1 - 1e-12):
            break
        dt = ref_run.tstep()
        if tend is not None:
            dt = min(dt, tend - t)
        ref_run.ssprk(dt)
        dts.append(dt)
        t += dt
        half = (nsteps is not None and len(dts) == nsteps // 2) or (tend is not None and t >= 0.5 * tend)
        if thr_un is None and half:
            # the untrained network's threshold, matched on the mid-run state
            thr_un, frac_cal = matched_threshold(ref_run, P, thr, P_un)
    if thr_un is None:
        thr_un, frac_cal = matched_threshold(ref_run, P, thr, P_un)
    q_ref = ref_run.q0()
    b_ref = ref_run.faces() if speed_end else None
    ref_run.close()
    log(f"[{name}] reference: {len(dts)} steps to t = {t:.5g}")

    forms = forms or ["fp64", "fp32", "fp32+indicator", "fp32+linear", "fp32+untrained", "fp64+indicator"]
    P_lin = P.copy()       # output weights 0, bias -50: nothing is flagged
    P_lin[1664:1696] = 0.0
    P_lin[1696] = -50.0
    spec = {
        "fp64": ("fp64", None, None),
        "fp32": ("fp32", None, None),
        "fp32+indicator": ("fp32", P, thr),
        "fp32+linear": ("fp32", P_lin, 0.5),
        "fp32+untrained": ("fp32", P_un, thr_un),
        "fp64+indicator": ("fp64", P, thr),
    }
    out = {"case": name, "problem": problem, "input": kw, "steps": len(dts), "t_end": t, "zones": zones,
           "untrained_threshold": thr_un, "calibration_flagged_fraction": frac_cal, "forms": {}}
    base_speed, base_end, base_run = None, None, None

    def timed(g):
        g.ssprk(g.tstep())
        g.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(time_steps):
            g.ssprk(g.tstep())
        g.synchronize()
        return time.perf_counter() - t0
    for form in forms:
        prec, Pf, thf = spec[form]
        # the product forms route with the plain-WENO fallback; the ablation
        # forms route every step
        mode = 1 if form in ("fp32+indicator", "fp64+indicator") else 2
        # speed from the initial state (and from the reference's final state)
        g = _solver(run, prec, Pf, thf, enable=mode)
        speed_e, st_e = None, None
        if speed_end:
            g.upload_q0(q_ref)
            g.upload_faces(b_ref)
            speed_e = zones * time_steps / timed(g)
            if Pf is not None:
                st_e = g.indicator_stats()
                g.set_indicator(Pf, thf, enable=mode)   # clears the counters and the fallback state
        g.upload_q0(q0)
        g.init_faces()
        el = timed(g)
        speed = zones * time_steps / el
        if form == "fp64":
            base_speed, base_end = speed, speed_e
        # accuracy along the reference run's dt sequence (same routing mode)
        if Pf is not None:
            g.set_indicator(Pf, thf, enable=mode)
        g.upload_q0(q0)
        g.init_faces()
        ok = True
        g.synchronize()
        t0 = time.perf_counter()
        for dt in dts:
            g.ssprk(dt)
        g.synchronize()
        el_run = time.perf_counter() - t0
        if form == "fp64":
            base_run = el_run
        qf = g.q0()
        if not np.all(np.isfinite(H.active(qf))):
            ok = False
        l1 = H.rel_l1(qf, q_ref)
        ent = {"zone_cycles_per_s": round(speed, 1), "ms_per_step": round(el / time_steps * 1e3, 3),
               "rel_l1_rho": float(l1[0]), "rel_l1_E": float(l1[7]), "finite": ok,
               "divb_ct_max": g.diag().max_divb_ct}
        if base_speed:
            ent["speedup_vs_fp64"] = round(speed / base_speed, 3)
        # the whole run along the reference's dt sequence (no tstep)
        ent["ms_per_step_run"] = round(el_run / len(dts) * 1e3, 3)
        if base_run:
            ent["speedup_vs_fp64_run"] = round(base_run / el_run, 3)
        if speed_e is not None:
            ent["zone_cycles_per_s_end"] = round(speed_e, 1)
            ent["ms_per_step_end"] = round(zones / speed_e * 1e3, 3)
            if base_end:
                ent["speedup_vs_fp64_end"] = round(speed_e / base_end, 3)
        if st_e is not None:
            ent["flagged_fraction_end"] = round(st_e["flagged_fraction"], 4)
            ent["wavefront_flagged_fraction_end"] = round(st_e["wave_fraction"], 4)
        if Pf is not None:
            st = g.indicator_stats()
            ent["flagged_fraction"] = round(st["flagged_fraction"], 4)
            ent["wavefront_flagged_fraction"] = round(st["wave_fraction"], 4)
            ent["routed_steps"] = st["routed_steps"]
            ent["fallback_steps"] = st["fallback_steps"]
        g.close()
        out["forms"][form] = ent
        log(f"[{name}] {form}: {json.dumps(ent)}")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--cases", default="rotor512,turb256,blast256")
    ap.add_argument("--time-steps", type=int, default=10)
    ap.add_argument("--out", default=None)
    ap.add_argument("--speed-end", action="store_true",
                    help="also time every form from the reference run's final state")
    a = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("eval_indicator.py needs a GPU")
    rep = {"weights": str(WEIGHTS.name), "cases": {}}
    for c in a.cases.split(","):
        rep["cases"][c] = run_case(c, a.time_steps, speed_end=a.speed_end)
    if a.out:
        Path(a.out).write_text(json.dumps(rep, indent=1))
    return rep


if __name__ == "__main__":
    main()
