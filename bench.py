#!/usr/bin/env python3
"""Headline benchmark: zone-cycles/s of the HIP HOW-MHD time step on one MI355X.

Workload: 3-D MHD blast (problem mhd_blast3d) on a 256^3 periodic unit box,
rho=1, p=0.1 ambient and p=10 for r<0.1, B=(1/sqrt2, 1/sqrt2, 0), v=0,
gam=5/3, cour=0.4, plain method in FP64 with the reference's rounding (bit for
bit equal to the CPU restatement).  One "step" is one hydro cycle: the CFL
reduction (tstep) plus one SSPRK(5,4) step with its 15 directional sweeps,
5 CT updates and 5 stage combinations.  q(0) and the faces are resident in
HBM before the timed region; the per-step dt read-back is part of the loop,
as in the reference driver.

value = 256^3 * steps / wall time of the timed steps.  After the timed loop a
short run with per-kernel HIP-event timing on the library's stream gives the
roofline entry of the dominant kernel.  The CPU baseline is the C restatement
of the reference time step (oracle/) on the host's physical cores with
OMP_PROC_BIND=spread and OMP_PLACES=cores (the reference README's settings),
in a child process, on a bounded sample (the benchmark workload, two steps at
256^3).

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "how-mhd-kokkos_amd"))

METRIC = "zone-cycles/s per SSPRK(5,4) step at 256³; speedup vs plain FP64 at fixed L1"

# Algorithmic work per unit, 3-D FP64 (add/sub/mul/div/sqrt = 1 FLOP):
#   directional sweep: 2,850 FLOP per cell (cell values 85, interface
#   eigensystem 540, 7 waves x 314 WENO/projection, 31 update/CT terms)
#   whole zone-step:   44,000 FLOP and 1,472 B of compulsory state traffic
FLOP_SWEEP_CELL = 2850.0
FLOP_ZONE_STEP = 44000.0
BYTES_ZONE_STEP = 1472.0
# bytes per cell of each memory-bound kernel (FP64 path, compulsory reads+writes
# of stages 2-4; CT active, so the transverse-B fluxes and q's cell-centred B
# are not read by the combination)
BYTES_CELL = {
    "emf2": 8 * (6 + 3),          # 6 CT flux arrays in, 3 EMF candidates out
    "emf": 8 * (3 + 3),           # candidates in, smoothed EMFs out (k_emf)
    "emf_fused": 8 * (6 + 3),     # 3-D grids: k_emf_march, CT fluxes in, smoothed EMFs out
    "faces": 8 * (3 + 6 + 3),     # O in, previous + initial faces in, faces out
    "combine": 8 * (15 + 5 + 5 + 3 + 8),  # 3x5 fluxes, q_init and q_prev without B, faces in; q out
    "cfl": 24.0 / 256,            # tstep folds 3 maxima per 256-cell block of the stage-5 combination
}
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "fp32", "fp64fast"], default="fp64")
    ap.add_argument("--indicator", action="store_true",
                    help="time the form with the learned indicator (weights/indicator_mlp.txt) instead")
    ap.add_argument("--cpu-size", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=3)
    ap.add_argument("--no-forms", action="store_true", help="skip the FP32 form leg")
    ap.add_argument("--err-steps", type=int, default=10)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (HBM-counter calibration under rocprofv3)")
    return ap.parse_args()


TRAFFIC_FILE = ROOT / "profiles" / "traffic_mhd_blast3d_256_fp64.txt"


def traffic_table() -> dict:
    """HBM bytes per launch of each kernel from the committed rocprofv3 PMC
    summary of the benchmark workload (profiles/traffic_mhd_blast3d_256_fp64.txt:
    '<kernel> <bytes_per_launch>' lines); empty when it is absent."""
    out = {}
    if not TRAFFIC_FILE.exists():
        return out
    for line in TRAFFIC_FILE.read_text().splitlines():
        parts = line.split()
        if len(parts) >= 2 and not line.startswith("#"):
            try:
                out[parts[0]] = float(parts[1])
            except ValueError:
                pass
    return out


def traffic_per_step(tab: dict, launches_per_step: dict):
    """Counter bytes of one step: per-launch bytes x launches per step, or
    None when a kernel of the step is missing from the table."""
    tot = 0.0
    for k, n in launches_per_step.items():
        if k not in tab:
            return None
        tot += tab[k] * n
    return tot


CPU_CHILD = r"""
import json, sys, time
sys.path[:0] = [sys.argv[1], sys.argv[2]]
import howmhd as H
import oracle_ctypes as O
n, steps, threads = int(sys.argv[3]), int(sys.argv[4]), int(sys.argv[5])
O.set_threads(threads)
run = H.parse(H.make_input(problem="mhd_blast3d", nx=n, ny=n, nz=n, nt=1, tend=1.0, cour=0.4,
                           gam="1.6666666666666667"))
q0 = H.init_problem(run)
ref = O.RefSolver(H.gpu_params(run))
ref.upload_q0(q0)
ref.init_faces()
del q0
t0 = time.perf_counter()
for _ in range(steps):
    ref.ssprk(ref.tstep())
print(json.dumps({"seconds": time.perf_counter() - t0}))
"""


def host_cpu():
    """(CPU model, physical cores of the machine, physical cores among the
    CPUs this process may run on) from lscpu."""
    import subprocess
    model, phys_all, phys_mine = "unknown", None, None
    try:
        info = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        for line in info.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
        rows = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True,
                              timeout=30).stdout.splitlines()
        allowed = os.sched_getaffinity(0)
        cores, mine = set(), set()
        for r in rows:
            if r.startswith("#"):
                continue
            cpu, core, sock = (int(x) if x else 0 for x in r.split(","))
            cores.add((sock, core))
            if cpu in allowed:
                mine.add((sock, core))
        phys_all, phys_mine = len(cores), len(mine)
    except Exception:
        pass
    return model, phys_all, phys_mine


def cpu_baseline(n: int, steps: int):
    """The oracle on the host's physical cores, OMP_PROC_BIND=spread,
    OMP_PLACES=cores (README.md:409-412 of the reference), in a child process
    so that the OpenMP runtime starts with those settings.  The thread count
    is the physical cores among the CPUs this process may use, capped by
    OMP_NUM_THREADS when the environment sets it (a machine that shares its
    cores among GPUs gives each its share that way)."""
    import subprocess
    model, phys_all, phys_mine = host_cpu()
    threads = phys_mine or len(os.sched_getaffinity(0))
    env_cap = os.environ.get("OMP_NUM_THREADS")
    if env_cap and env_cap.isdigit() and int(env_cap) > 0:
        threads = min(threads, int(env_cap))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread", OMP_PLACES="cores")
    res = subprocess.run([sys.executable, "-c", CPU_CHILD, str(ROOT / "how-mhd-kokkos_amd"), str(ROOT / "oracle"),
                          str(n), str(steps), str(threads)], env=env, capture_output=True, text=True, timeout=900)
    if res.returncode != 0:
        raise RuntimeError(f"CPU baseline child failed ({res.returncode}): {res.stderr[-2000:]}")
    el = json.loads(res.stdout.strip().splitlines()[-1])["seconds"]
    return {
        "value": n ** 3 * steps / el,
        "unit": "zone-cycles/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": model,
        "physical_cores_machine": phys_all,
        "physical_cores_allowed": phys_mine,
        "omp": "OMP_PROC_BIND=spread OMP_PLACES=cores",
        "sample": f"mhd_blast3d {n}^3, {steps} steps (tstep + SSPRK(5,4)) of the C restatement of the "
                  f"reference time step, OpenMP, {threads} threads on {model}; {el:.2f} s",
    }


WEIGHTS = ROOT / "how-mhd-kokkos_amd" / "weights" / "indicator_mlp.txt"
# conserved variables in the reference order (README.md:53-64)
VARS = ["rho", "Mx", "My", "Mz", "Bx", "By", "Bz", "E"]


def form_leg(H, torch, run, args, fp64_value, precision, indicator=False):
    """One form of the time step on the same workload: zone-cycles/s, speedup
    over the plain FP64 path, and relative L1 against the plain FP64 path
    after --err-steps steps driven with the plain path's dt sequence.
    precision: fp64fast (FMA rounding), fp32 (FP32 sweeps and CT EMFs; FP64
    stage sums, E->p, floors and face-B); indicator: the learned smoothness
    indicator routes unflagged interfaces to linear upwind-5."""
    import numpy as np

    n = args.size
    q0 = H.init_problem(run)
    p64 = H.gpu_params(run)
    pf = H.gpu_params(run)
    pf.precision = H.PREC[precision]
    gf = H.GpuSolver(pf)
    extra = {}
    if indicator:
        P, thr = H.load_indicator(WEIGHTS)
        gf.set_indicator(P, thr, enable=1)
        extra["indicator_weights"] = str(WEIGHTS.relative_to(ROOT))
        extra["threshold"] = thr
    gf.upload_q0(q0)
    gf.init_faces()
    for _ in range(args.warmup):
        gf.ssprk(gf.tstep())
    gf.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gf.ssprk(gf.tstep())
    gf.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = n ** 3 * args.steps / el
    # error against the plain FP64 path, same dt sequence, from the same q(0)
    if indicator:
        gf.set_indicator(P, thr, enable=2)
    g64 = H.GpuSolver(p64)
    for s in (gf, g64):
        s.upload_q0(q0)
        s.init_faces()
    del q0
    for _ in range(args.err_steps):
        dt = g64.tstep()
        g64.ssprk(dt)
        gf.ssprk(dt)
    af, a64 = gf.q0(), g64.q0()
    l1 = H.rel_l1(af, a64)
    emag = lambda a: 0.5 * float((H.active(a)[4:7] ** 2).sum())
    df = gf.diag()
    if indicator:
        st = gf.indicator_stats()
        extra["flagged_fraction"] = round(st["flagged_fraction"], 4)
        extra["wavefront_flagged_fraction"] = round(st["wave_fraction"], 4)
    gf.close()
    g64.close()
    out = {
        "value": round(value, 1),
        "unit": "zone-cycles/s",
        "ms_per_step": round(el / args.steps * 1e3, 3),
        "speedup_vs_fp64": round(value / fp64_value, 3),
        "rel_l1_vs_fp64_after_steps": args.err_steps,
        "rel_l1_rho": float(l1[0]),
        "rel_l1_E": float(l1[7]),
        "rel_l1_max": float(np.max(l1)),
        "rel_l1_max_component": VARS[int(np.argmax(l1))],
        "rel_l1_per_component": {VARS[m]: float(l1[m]) for m in range(8)},
        "magnetic_energy_rel_err": abs(emag(af) - emag(a64)) / emag(a64),
        "divb_ct_max": df.max_divb_ct,
    }
    out.update(extra)
    return out


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(single-GPU time step; sharding by 3-D blocks with halo exchange is not built)", file=sys.stderr)
        sys.exit(2)
    import torch  # device plumbing only: selects the GPU and synchronises

    import howmhd as H

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    n = args.size
    run = H.parse(H.make_input(problem="mhd_blast3d", nx=n, ny=n, nz=n, nt=1, tend=1.0, cour=0.4,
                               gam="1.6666666666666667", precision=args.precision))
    q0 = H.init_problem(run)
    gpu = H.GpuSolver(H.gpu_params(run))
    if args.indicator:
        P, thr = H.load_indicator(WEIGHTS)
        gpu.set_indicator(P, thr, enable=1)
    t_up = time.perf_counter()
    gpu.upload_q0(q0)            # host (reference layout, ghosts) -> HBM, synchronous
    t_up = time.perf_counter() - t_up
    gpu.init_faces()
    del q0
    for _ in range(args.warmup):
        gpu.ssprk(gpu.tstep())
    gpu.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gpu.ssprk(gpu.tstep())
    gpu.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    zones = n ** 3
    value = zones * args.steps / el
    t_dn = time.perf_counter()
    qd = gpu.q0()                # HBM -> host (reference layout, ghosts filled), synchronous
    t_dn = time.perf_counter() - t_dn
    host_bytes = qd.nbytes
    del qd

    if args.copy_probe_mib > 0:
        nel = args.copy_probe_mib * 1024 * 1024 // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # per-kernel device time on the library's stream
    gpu.set_timing(True)
    for _ in range(args.timing_steps):
        gpu.ssprk(gpu.tstep())
    tim = gpu.timing()
    gpu.set_timing(False)
    d = gpu.diag()
    gpu.close()

    flop_peak = PEAK_FP32_TFLOPS if args.precision == "fp32" else PEAK_FP64_TFLOPS
    plain64 = args.precision == "fp64" and not args.indicator   # the profiled traffic is of this form
    # 3-D grids build and smooth the EMFs in one launch (timed as "emf")
    emf_fused = tim.get("emf2", (0.0, 0))[1] == 0
    kernels = {}
    for k, (ms, cnt) in tim.items():
        if cnt == 0:
            continue
        avg_ms = ms / cnt
        ent = {"avg_ms": round(avg_ms, 4), "launches": cnt, "ms_per_step": round(ms / args.timing_steps, 4)}
        if k.startswith("sweep"):
            tf = FLOP_SWEEP_CELL * zones / (avg_ms * 1e-3) / 1e12
            ent.update(achieved_tflops=round(tf, 3), frac_fp_peak=round(tf / flop_peak, 4))
        elif k in BYTES_CELL:
            bk = "emf_fused" if (k == "emf" and emf_fused) else k
            gbs = BYTES_CELL[bk] * zones / (avg_ms * 1e-3) / 1e9
            ent.update(achieved_gbs=round(gbs, 1), frac_hbm_peak=round(gbs / PEAK_HBM_GBS, 4))
        kernels[k] = ent
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    dk = kernels[dom]
    tab = traffic_table() if (plain64 and n == 256) else {}
    if dom.startswith("sweep"):
        achieved = dk["achieved_tflops"]
        roof = {"bound": "valu", "achieved": achieved, "peak": flop_peak, "unit": "TFLOP/s",
                "frac": round(achieved / flop_peak, 4),
                "traffic": tab.get(dom),
                "kernel": f"k_sweep ({dom}); FP64 vector-ALU bound (no MFMA: the sweep has no dense "
                          f"contraction; MI355X FP64 vector peak = 78.6 TFLOP/s)",
                "per_launch": f"{FLOP_SWEEP_CELL:.0f} FLOP/cell x {zones} cells"}
    else:
        achieved = dk["achieved_gbs"]
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": tab.get(dom), "kernel": dom}
    # counter bytes of a whole step against its compulsory bytes (1,472 B per
    # zone-step): 5 launches per stage kernel, one CFL fold
    step_bytes = traffic_per_step(tab, {"sweep_x": 5, "sweep_y": 5, "sweep_z": 5, "emf": 5, "faces": 5,
                                        "combine": 5, "cfl": 1}) if tab else None
    roof["traffic_step"] = step_bytes
    roof["traffic_ratio"] = round(step_bytes / (BYTES_ZONE_STEP * zones), 3) if step_bytes else None
    whole = {
        "fp64_tflops_effective": round(FLOP_ZONE_STEP * value / 1e12, 3),
        "frac_fp64_roof": round(FLOP_ZONE_STEP * value / (PEAK_FP64_TFLOPS * 1e12), 4),
        "compulsory_hbm_gbs": round(BYTES_ZONE_STEP * value / 1e9, 1),
    }
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp64fast": "f64", "fp32": "f32 sweeps/CT, f64 stage sums"}[args.precision],
        "data": "synthetic: deterministic mhd_blast3d initial condition",
        "config": {"workload": f"mhd_blast3d {n}^3 periodic, gam 5/3, cour 0.4, p_in 10, p_out 0.1, "
                               f"B=(1/sqrt2,1/sqrt2,0), precision {args.precision}, "
                               + ("learned indicator" if args.indicator else "plain method"),
                   "grid": [n, n, n], "precision": args.precision, "parallelism": "single GPU"},
        "roofline": roof,
        "kernels": kernels,
        "whole_step": whole,
        "divb_ct_max": d.max_divb_ct,
        # the C-ABI hands over host buffers in the reference layout: one upload
        # of q(0) before and one download after the timed steps, host layout
        # transposition included (not part of `value`)
        "host_boundary": {
            "upload_s": round(t_up, 4), "download_s": round(t_dn, 4), "bytes_each": host_bytes,
            "pcie_inclusive_zone_cycles_per_s": round(zones * args.steps / (el + t_up + t_dn), 1),
        },
    }
    if not args.no_forms and args.precision == "fp64" and not args.indicator:
        forms = {"fp64fast": form_leg(H, torch, run, args, value, "fp64fast"),
                 "fp32": form_leg(H, torch, run, args, value, "fp32")}
        if WEIGHTS.exists():
            forms["fp32+indicator"] = form_leg(H, torch, run, args, value, "fp32", indicator=True)
        out["forms"] = forms
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_size, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
