"""ctypes binding of the HOW-MHD MI355X libraries (tests, bench and smoke use it).

The product is the C-ABI in include/howmhd_gpu.h (time step on the GPU) and
include/howmhd_host.h (parser, problem registry, writers).  This module only
loads the in-tree shared libraries built by the package Makefile and wraps
them for Python callers; it computes nothing itself.  A missing library is an
error: there is no fallback path.

Arrays cross in the reference layout: q0 is (8, nx+6, ny+6, nz+6) float64,
C-contiguous (k fastest), faces (3, nx+6, ny+6, nz+6).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_DIR = PKG_DIR / "lib"
BIN_DIR = PKG_DIR / "bin"
DRIVER = BIN_DIR / "how_mhd"

BC = {"periodic": 0, "open": 1}
PREC = {"fp64": 0, "fp32": 1}
INDICATOR = {"off": 0, "mlp": 1}


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("bc", C.c_int32 * 3),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("gam", C.c_double), ("cour", C.c_double), ("rhomin", C.c_double), ("pgmin", C.c_double),
        ("precision", C.c_int32), ("indicator", C.c_int32),
    ]


class Diag(C.Structure):
    _fields_ = [
        ("sum_cons", C.c_double * 8),
        ("max_divb_ct", C.c_double),
        ("max_divb_cell", C.c_double),
        ("min_rho", C.c_double), ("min_p", C.c_double),
        ("flagged_fraction", C.c_double),
    ]


KERNELS = ["sweep_x", "sweep_y", "sweep_z", "emf2", "emf", "faces", "combine", "cfl"]


class Timing(C.Structure):
    _fields_ = [("ms", C.c_double * 8), ("launches", C.c_int32 * 8)]


class Run(C.Structure):
    _fields_ = [
        ("gnx", C.c_int32), ("gny", C.c_int32), ("gnz", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("nt", C.c_int32), ("nstep", C.c_int32), ("ntime", C.c_int32),
        ("px", C.c_int32), ("py", C.c_int32), ("pz", C.c_int32),
        ("xsize", C.c_double), ("ysize", C.c_double), ("zsize", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("gam", C.c_double), ("cour", C.c_double), ("tend", C.c_double),
        ("rhomin", C.c_double), ("pgmin", C.c_double),
        ("problem", C.c_char * 64),
        ("x1bc", C.c_char * 32), ("x2bc", C.c_char * 32), ("x3bc", C.c_char * 32),
        ("output_format", C.c_char * 32),
        ("precision", C.c_char * 16),
        ("indicator", C.c_char * 16),
        ("indicator_weights", C.c_char * 256),
        ("seed", C.c_uint64),
        ("quiet", C.c_int32),
    ]


_DP = C.POINTER(C.c_double)


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


_libs: dict[str, C.CDLL] = {}


def _load(name: str) -> C.CDLL:
    if name not in _libs:
        path = LIB_DIR / name
        if not path.exists():
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build() (make -C how-mhd-kokkos_amd)")
        _libs[name] = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
    return _libs[name]


GPU_FUNCS = {
    "howmhd_gpu_create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "howmhd_gpu_destroy": (C.c_int, [C.c_void_p]),
    "howmhd_gpu_upload_q0": (C.c_int, [C.c_void_p, _DP]),
    "howmhd_gpu_init_faces": (C.c_int, [C.c_void_p]),
    "howmhd_gpu_upload_faces": (C.c_int, [C.c_void_p, _DP]),
    "howmhd_gpu_tstep": (C.c_int, [C.c_void_p, _DP]),
    "howmhd_gpu_ssprk": (C.c_int, [C.c_void_p, C.c_double]),
    "howmhd_gpu_download_q0": (C.c_int, [C.c_void_p, _DP]),
    "howmhd_gpu_download_faces": (C.c_int, [C.c_void_p, _DP]),
    "howmhd_gpu_diag": (C.c_int, [C.c_void_p, C.POINTER(Diag)]),
    "howmhd_gpu_synchronize": (C.c_int, [C.c_void_p]),
    "howmhd_gpu_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "howmhd_gpu_get_timing": (C.c_int, [C.c_void_p, C.POINTER(Timing)]),
    "howmhd_gpu_last_error": (C.c_char_p, [C.c_void_p]),
    "howmhd_gpu_build_kind": (C.c_char_p, []),
}

HOST_FUNCS = {
    "howmhd_host_parse": (C.c_int, [C.c_char_p, C.POINTER(Run), C.c_char_p, C.c_int]),
    "howmhd_host_gpu_params": (C.c_int, [C.POINTER(Run), C.POINTER(Params), C.c_char_p, C.c_int]),
    "howmhd_host_init_problem": (C.c_int, [C.POINTER(Run), _DP, C.c_char_p, C.c_int]),
    "howmhd_host_problem_names": (C.c_char_p, []),
    "howmhd_host_output": (C.c_int, [C.POINTER(Run), _DP, C.c_int, C.c_double, C.c_char_p, C.c_char_p, C.c_int]),
}


def _bind(lib: C.CDLL, table: dict) -> C.CDLL:
    for fn, (res, args) in table.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    return lib


def gpu_lib(exact: bool = False) -> C.CDLL:
    """The product library (fp-fast) or its FP-contraction-off twin (fp-exact)."""
    return _bind(_load("libhowmhd_gpu_exact.so" if exact else "libhowmhd_gpu.so"), GPU_FUNCS)


def host_lib() -> C.CDLL:
    return _bind(_load("libhowmhd_host.so"), HOST_FUNCS)


class HowmhdError(RuntimeError):
    pass


# ----------------------------------------------------------------- host shell
def parse(text: str) -> Run:
    lib = host_lib()
    run = Run()
    err = C.create_string_buffer(4096)
    rc = lib.howmhd_host_parse(text.encode(), C.byref(run), err, len(err))
    if rc != 0:
        raise HowmhdError(err.value.decode())
    return run


def gpu_params(run: Run) -> Params:
    lib = host_lib()
    p = Params()
    err = C.create_string_buffer(4096)
    if lib.howmhd_host_gpu_params(C.byref(run), C.byref(p), err, len(err)) != 0:
        raise HowmhdError(err.value.decode())
    return p


def init_problem(run: Run) -> np.ndarray:
    lib = host_lib()
    q0 = np.zeros((8, run.nx + 6, run.ny + 6, run.nz + 6))
    err = C.create_string_buffer(8192)
    if lib.howmhd_host_init_problem(C.byref(run), _ptr(q0), err, len(err)) != 0:
        raise HowmhdError(err.value.decode())
    return q0


def problem_names() -> list[str]:
    return host_lib().howmhd_host_problem_names().decode().split()


def write_output(run: Run, q0: np.ndarray, step: int, time: float, directory: str) -> None:
    lib = host_lib()
    err = C.create_string_buffer(4096)
    if lib.howmhd_host_output(C.byref(run), _ptr(np.ascontiguousarray(q0)), step, time, directory.encode(), err,
                              len(err)) != 0:
        raise HowmhdError(err.value.decode())


def make_input(**kw) -> str:
    """Input-file text from keyword=value pairs (reference keywords)."""
    return "\n".join(f"{k} {v}" for k, v in kw.items()) + "\n"


# ------------------------------------------------------------------ GPU solver
class GpuSolver:
    """One device context of the C-ABI: q(0) and faces stay resident."""

    def __init__(self, params: Params, exact: bool = False):
        self.lib = gpu_lib(exact)
        self.p = params
        self.shape = (params.nx + 6, params.ny + 6, params.nz + 6)
        h = C.c_void_p()
        rc = self.lib.howmhd_gpu_create(C.byref(params), C.byref(h))
        if rc != 0:
            raise HowmhdError(f"howmhd_gpu_create failed ({rc}): {self.lib.howmhd_gpu_last_error(None).decode()}")
        self.h = h

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise HowmhdError(f"{what} failed ({rc}): {self.lib.howmhd_gpu_last_error(self.h).decode()}")

    def upload_q0(self, q0: np.ndarray):
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        assert q0.shape == (8,) + self.shape
        self._check(self.lib.howmhd_gpu_upload_q0(self.h, _ptr(q0)), "upload_q0")

    def init_faces(self):
        self._check(self.lib.howmhd_gpu_init_faces(self.h), "init_faces")

    def upload_faces(self, b: np.ndarray):
        b = np.ascontiguousarray(b, dtype=np.float64)
        assert b.shape == (3,) + self.shape
        self._check(self.lib.howmhd_gpu_upload_faces(self.h, _ptr(b)), "upload_faces")

    def tstep(self) -> float:
        dt = C.c_double()
        self._check(self.lib.howmhd_gpu_tstep(self.h, C.byref(dt)), "tstep")
        return dt.value

    def ssprk(self, dt: float):
        self._check(self.lib.howmhd_gpu_ssprk(self.h, dt), "ssprk")

    def synchronize(self):
        self._check(self.lib.howmhd_gpu_synchronize(self.h), "synchronize")

    def q0(self) -> np.ndarray:
        out = np.empty((8,) + self.shape)
        self._check(self.lib.howmhd_gpu_download_q0(self.h, _ptr(out)), "download_q0")
        return out

    def faces(self) -> np.ndarray:
        out = np.empty((3,) + self.shape)
        self._check(self.lib.howmhd_gpu_download_faces(self.h, _ptr(out)), "download_faces")
        return out

    def diag(self) -> Diag:
        d = Diag()
        self._check(self.lib.howmhd_gpu_diag(self.h, C.byref(d)), "diag")
        return d

    def set_timing(self, on: bool):
        self._check(self.lib.howmhd_gpu_set_timing(self.h, int(on)), "set_timing")

    def timing(self) -> dict:
        """{kernel: (total ms, launches)} of the launches timed since set_timing(True)."""
        t = Timing()
        self._check(self.lib.howmhd_gpu_get_timing(self.h, C.byref(t)), "get_timing")
        return {k: (t.ms[i], t.launches[i]) for i, k in enumerate(KERNELS)}

    def close(self):
        if getattr(self, "h", None):
            self.lib.howmhd_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def build_kind(exact: bool = False) -> str:
    return gpu_lib(exact).howmhd_gpu_build_kind().decode()


def active(a: np.ndarray) -> np.ndarray:
    """Active cells of a reference-layout array (leading component axis kept)."""
    return a[:, 3:-3, 3:-3, 3:-3]


def rel_l1(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Relative L1 difference per component over active cells: sum|a-b| / sum|b|."""
    a, b = active(a), active(b)
    num = np.abs(a - b).reshape(a.shape[0], -1).sum(axis=1)
    den = np.abs(b).reshape(b.shape[0], -1).sum(axis=1)
    return np.where(den > 0, num / np.where(den > 0, den, 1.0), num)
