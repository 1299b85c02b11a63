"""ctypes binding of the CPU restatement (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as
the parity reference; never by the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB = ORACLE_DIR / "_build" / "libhowmhd_oracle.so"

_DP = C.POINTER(C.c_double)
_lib = None


def build() -> None:
    subprocess.run(["make", "-C", str(ORACLE_DIR)], check=True, capture_output=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        from_params = C.c_void_p  # opaque howmhd_params*
        table = {
            "howmhd_ref_create": (C.c_int, [from_params, C.POINTER(C.c_void_p)]),
            "howmhd_ref_destroy": (C.c_int, [C.c_void_p]),
            "howmhd_ref_upload_q0": (C.c_int, [C.c_void_p, _DP]),
            "howmhd_ref_upload_faces": (C.c_int, [C.c_void_p, _DP]),
            "howmhd_ref_init_faces": (C.c_int, [C.c_void_p]),
            "howmhd_ref_tstep": (C.c_int, [C.c_void_p, _DP]),
            "howmhd_ref_ssprk": (C.c_int, [C.c_void_p, C.c_double]),
            "howmhd_ref_download_q0": (C.c_int, [C.c_void_p, _DP]),
            "howmhd_ref_download_faces": (C.c_int, [C.c_void_p, _DP]),
            "howmhd_ref_set_threads": (C.c_int, [C.c_int]),
            "howmhd_ref_floor_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
            "howmhd_ref_ct_emf": (C.c_int, [C.c_void_p, _DP, _DP]),
            "howmhd_ref_ssprk_stages": (C.c_int, [C.c_void_p, C.c_double, C.c_int, _DP, _DP]),
            "howmhd_ref_line_sweep": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, _DP, _DP,
                                                _DP, _DP]),
            "howmhd_ref_eigensystem": (None, [C.c_double, _DP, _DP, _DP, _DP]),
            "howmhd_ref_weno_wave": (C.c_double, [_DP, _DP, C.c_double]),
            "howmhd_ref_weno_wave_weights": (C.c_double, [_DP, _DP, C.c_double, _DP, _DP]),
            "howmhd_ref_upwind5": (C.c_double, [_DP, _DP, C.c_double]),
            "howmhd_ref_set_route": (C.c_int, [C.c_void_p, C.c_int]),
            "howmhd_ref_large_t_count": (C.c_longlong, [C.c_int]),
            "howmhd_ref_odd_sqrt_count": (C.c_longlong, [C.c_int]),
            "howmhd_ref_line_maxdev": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int, _DP, _DP]),
        }
        for fn, (res, args) in table.items():
            f = getattr(_lib, fn)
            f.restype = res
            f.argtypes = args
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


def set_threads(n: int) -> int:
    return lib().howmhd_ref_set_threads(n)


class RefSolver:
    """Same call sequence as howmhd.GpuSolver, on the CPU restatement."""

    def __init__(self, params):
        self.lib = lib()
        self.p = params
        self.shape = (params.nx + 6, params.ny + 6, params.nz + 6)
        h = C.c_void_p()
        rc = self.lib.howmhd_ref_create(C.cast(C.pointer(params), C.c_void_p), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"howmhd_ref_create failed ({rc})")
        self.h = h

    def upload_q0(self, q0):
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        assert q0.shape == (8,) + self.shape
        self.lib.howmhd_ref_upload_q0(self.h, _ptr(q0))

    def upload_faces(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        self.lib.howmhd_ref_upload_faces(self.h, _ptr(b))

    def init_faces(self):
        self.lib.howmhd_ref_init_faces(self.h)

    def tstep(self) -> float:
        dt = C.c_double()
        self.lib.howmhd_ref_tstep(self.h, C.byref(dt))
        return dt.value

    def ssprk(self, dt: float):
        self.lib.howmhd_ref_ssprk(self.h, dt)

    def q0(self):
        out = np.empty((8,) + self.shape)
        self.lib.howmhd_ref_download_q0(self.h, _ptr(out))
        return out

    def faces(self):
        out = np.empty((3,) + self.shape)
        self.lib.howmhd_ref_download_faces(self.h, _ptr(out))
        return out

    def floor_counts(self) -> tuple[int, int]:
        """(density, pressure) floors applied by the stage-5 prot of the last ssprk."""
        nr, np_ = C.c_longlong(), C.c_longlong()
        assert self.lib.howmhd_ref_floor_counts(self.h, C.byref(nr), C.byref(np_)) == 0
        return nr.value, np_.value

    def ct_emf(self, ct6):
        """Smoothed corner EMFs (Ox, Oy, Oz) that the CT update builds from the
        six padded flux arrays ct6 = (fsy, fsz, gsx, gsz, hsx, hsy)."""
        ct6 = np.ascontiguousarray(ct6, dtype=np.float64)
        assert ct6.shape == (6,) + self.shape
        out = np.empty((3,) + self.shape)
        rc = self.lib.howmhd_ref_ct_emf(self.h, _ptr(ct6), _ptr(out))
        if rc != 0:
            raise RuntimeError(f"howmhd_ref_ct_emf failed ({rc})")
        return out

    def ssprk_stages(self, dt: float, nst: int):
        """(q, faces) of stage nst after the first nst stages of a step; q(0)
        and the stage-0 faces are unchanged."""
        q = np.empty((8,) + self.shape)
        b = np.empty((3,) + self.shape)
        rc = self.lib.howmhd_ref_ssprk_stages(self.h, dt, nst, _ptr(q), _ptr(b))
        if rc != 0:
            raise RuntimeError(f"howmhd_ref_ssprk_stages failed ({rc})")
        return q, b

    def set_route(self, route: int):
        """0: characteristic WENO5 everywhere; 1: linear upwind-5 everywhere."""
        assert self.lib.howmhd_ref_set_route(self.h, route) == 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.howmhd_ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def line_sweep(qline: np.ndarray, gam: float, dtdl: float, rhomin=1e-12, pgmin=1e-12):
    """One sweep of one line: returns (increments (8, nn+6), bsy, bsz, dF (7, nn+1))."""
    q = np.ascontiguousarray(qline, dtype=np.float64).copy()
    nn = q.shape[1] - 6
    bsy = np.zeros(nn + 1)
    bsz = np.zeros(nn + 1)
    dF = np.zeros((7, nn + 1))
    rc = lib().howmhd_ref_line_sweep(gam, rhomin, pgmin, nn, dtdl, _ptr(q), _ptr(bsy), _ptr(bsz), _ptr(dF))
    assert rc == 0
    return q, bsy, bsz, dF


def eigensystem(gam: float, ul, ur):
    ul = np.ascontiguousarray(ul, dtype=np.float64)
    ur = np.ascontiguousarray(ur, dtype=np.float64)
    L = np.zeros((7, 7))
    R = np.zeros((7, 7))
    lib().howmhd_ref_eigensystem(gam, _ptr(ul), _ptr(ur), _ptr(L), _ptr(R))
    return L, R


def weno_wave(Fsk, qsk, amx, weights=False):
    F = np.ascontiguousarray(Fsk, dtype=np.float64)
    q = np.ascontiguousarray(qsk, dtype=np.float64)
    if not weights:
        return lib().howmhd_ref_weno_wave(_ptr(F), _ptr(q), amx)
    wp = np.zeros(3)
    wm = np.zeros(3)
    fs = lib().howmhd_ref_weno_wave_weights(_ptr(F), _ptr(q), amx, _ptr(wp), _ptr(wm))
    return fs, wp, wm


def upwind5(F, U, alpha):
    F = np.ascontiguousarray(F, dtype=np.float64)
    U = np.ascontiguousarray(U, dtype=np.float64)
    return lib().howmhd_ref_upwind5(_ptr(F), _ptr(U), alpha)


def line_maxdev(qline: np.ndarray, gam: float, rhomin=1e-12, pgmin=1e-12):
    """Largest WENO-weight deviation from the ideal weights at the nn+1 interfaces of a line."""
    q = np.ascontiguousarray(qline, dtype=np.float64).copy()
    nn = q.shape[1] - 6
    out = np.zeros(nn + 1)
    assert lib().howmhd_ref_line_maxdev(gam, rhomin, pgmin, nn, _ptr(q), _ptr(out)) == 0
    return out


def large_t_count(reset=False):
    """WENO weight evaluations so far with some t_k = (eps + IS_k)^2 above
    1e200 or not finite."""
    return int(lib().howmhd_ref_large_t_count(int(reset)))


def odd_sqrt_count(reset=False):
    """Sweep square-root operands so far outside [2^-767, inf) and not +0."""
    return int(lib().howmhd_ref_odd_sqrt_count(int(reset)))
