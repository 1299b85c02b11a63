"""Writes tests/golden/oracle_regression.json: regression fixtures of the CPU
restatement of the time step (oracle/), not outputs of the reference (which
ships no tests or golden data and cannot be built here).

For each small case: the dt of every step (hex floats) and SHA-256 digests of
the active cells of q(0) and of the faces after the steps.  The CPU test
recomputes them, so any change of the restatement's arithmetic shows; the GPU
parity tests compare the plain FP64 GPU path with the restatement bit for bit,
so the digests pin that path too.

    python tests/golden/make_oracle_fixtures.py
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "how-mhd-kokkos_amd"), str(ROOT / "oracle")]
import howmhd as H  # noqa: E402
import oracle_ctypes as O  # noqa: E402

CASES = [
    ("orszag_tang", dict(nx=24, ny=24, nz=1, cour=0.8), 4),
    ("brio_wu", dict(nx=48, ny=1, nz=1, x1bc="open", x2bc="open", x3bc="open"), 6),
    ("mhd_blast3d", dict(nx=12, ny=10, nz=8), 2),
    ("mhd_blast3d", dict(nx=10, ny=9, nz=8, x1bc="open", x3bc="open"), 2),
    ("cpaw3d", dict(nx=16, ny=8, nz=8, xsize=3.0, ysize=1.5, zsize=1.5), 2),
    ("mhd_rotor", dict(nx=24, ny=24, nz=1, gam=1.4), 3),
    ("blast3d", dict(nx=12, ny=12, nz=6, rhomin=0.9999, pgmin=5.0), 2),
]


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(H.active(a), dtype="<f8").tobytes()).hexdigest()


def run_case(problem, kw, steps):
    kw = dict(kw)
    kw.setdefault("nt", 1)
    kw.setdefault("tend", 1.0)
    run = H.parse(H.make_input(problem=problem, **kw))
    s = O.RefSolver(H.gpu_params(run))
    s.upload_q0(H.init_problem(run))
    s.init_faces()
    dts = []
    for _ in range(steps):
        dt = s.tstep()
        s.ssprk(dt)
        dts.append(float(dt).hex())
    return {"problem": problem, "input": kw, "steps": steps, "dt": dts, "q_sha256": digest(s.q0()),
            "faces_sha256": digest(s.faces())}


def main():
    O.set_threads(1)
    out = {"note": "regression fixtures of the CPU restatement (oracle/), written by "
                   "tests/golden/make_oracle_fixtures.py; not outputs of the reference",
           "cases": [run_case(*c) for c in CASES]}
    path = Path(__file__).with_name("oracle_regression.json")
    path.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
