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
PREC = {"fp64": 0, "fp32": 1, "fp64fast": 2}
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
        ("n_rho_floor", C.c_int64), ("n_p_floor", C.c_int64),
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
        ("checkpoint_every", C.c_int32),
        ("restart", C.c_char * 256),
        ("max_steps", C.c_int32),
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
        # RTLD_LOCAL: no other library binds to this one's symbols
        _libs[name] = C.CDLL(str(path), mode=C.RTLD_LOCAL)
    return _libs[name]


GPU_FUNCS = {
    "howmhd_gpu_create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "howmhd_gpu_destroy": (C.c_int, [C.c_void_p]),
    "howmhd_gpu_upload_q0": (C.c_int, [C.c_void_p, _DP]),
    "howmhd_gpu_init_faces": (C.c_int, [C.c_void_p]),
    "howmhd_gpu_upload_faces": (C.c_int, [C.c_void_p, _DP]),
    "howmhd_gpu_tstep": (C.c_int, [C.c_void_p, _DP]),
    "howmhd_gpu_tstep_noalloc": (C.c_int, [C.POINTER(Params), C.c_void_p, C.c_void_p, _DP]),
    "howmhd_gpu_ssprk": (C.c_int, [C.c_void_p, C.c_double]),
    "howmhd_gpu_download_q0": (C.c_int, [C.c_void_p, _DP]),
    "howmhd_gpu_download_faces": (C.c_int, [C.c_void_p, _DP]),
    "howmhd_gpu_diag": (C.c_int, [C.c_void_p, C.POINTER(Diag)]),
    "howmhd_gpu_synchronize": (C.c_int, [C.c_void_p]),
    "howmhd_gpu_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "howmhd_gpu_get_timing": (C.c_int, [C.c_void_p, C.POINTER(Timing)]),
    "howmhd_gpu_last_error": (C.c_char_p, [C.c_void_p]),
    "howmhd_gpu_build_kind": (C.c_char_p, []),
    # learned indicator, include/howmhd_mlp.h
    "howmhd_mlp_infer": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "howmhd_mlp_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_float, C.c_float,
                                  C.c_void_p, C.c_void_p, C.c_void_p]),
    "howmhd_mlp_adam": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_float,
                                  C.c_float, C.c_float, C.c_float, C.c_void_p]),
    "howmhd_gpu_collect_samples": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int64,
                                             C.POINTER(C.c_int64)]),
    "howmhd_gpu_set_indicator": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_float, C.c_int]),
    "howmhd_gpu_indicator_stats": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_int64)] * 4),
    "howmhd_gpu_indicator_routing": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_int64)] * 2),
    "howmhd_gpu_indicator_logits": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
}

MLP_NPARAM = 1697
MLP_NFEAT = 18

HOST_FUNCS = {
    "howmhd_host_parse": (C.c_int, [C.c_char_p, C.POINTER(Run), C.c_char_p, C.c_int]),
    "howmhd_host_gpu_params": (C.c_int, [C.POINTER(Run), C.POINTER(Params), C.c_char_p, C.c_int]),
    "howmhd_host_init_problem": (C.c_int, [C.POINTER(Run), _DP, C.c_char_p, C.c_int]),
    "howmhd_host_problem_names": (C.c_char_p, []),
    "howmhd_host_output": (C.c_int, [C.POINTER(Run), _DP, C.c_int, C.c_double, C.c_char_p, C.c_char_p, C.c_int]),
    "howmhd_host_load_indicator": (C.c_int, [C.c_char_p, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float),
                                             C.c_char_p, C.c_int]),
    "howmhd_host_write_checkpoint": (C.c_int, [C.POINTER(Run), C.c_char_p, _DP, _DP, C.c_int32, C.c_int64,
                                               C.c_double, C.c_char_p, C.c_int]),
    "howmhd_host_read_checkpoint": (C.c_int, [C.POINTER(Run), C.c_char_p, _DP, _DP, C.POINTER(C.c_int32),
                                              C.POINTER(C.c_int64), C.POINTER(C.c_double), C.c_char_p, C.c_int]),
}


def _bind(lib: C.CDLL, table: dict) -> C.CDLL:
    for fn, (res, args) in table.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    return lib


def gpu_lib() -> C.CDLL:
    """The time-step library (every precision form)."""
    return _bind(_load("libhowmhd_gpu.so"), GPU_FUNCS)


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


def load_indicator(path) -> tuple[np.ndarray, float]:
    """(params, threshold) of an indicator weights file (howmhd_host_load_indicator)."""
    lib = host_lib()
    P = np.zeros(MLP_NPARAM, dtype=np.float32)
    thr = C.c_float()
    err = C.create_string_buffer(4096)
    if lib.howmhd_host_load_indicator(str(path).encode(), P.ctypes.data_as(C.POINTER(C.c_float)), MLP_NPARAM,
                                      C.byref(thr), err, len(err)) != 0:
        raise HowmhdError(err.value.decode())
    return P, thr.value


def write_checkpoint(run: Run, path, q0: np.ndarray, b: np.ndarray, out_step: int, total_steps: int,
                     time: float) -> None:
    """howmhd_host_write_checkpoint: q(0), faces and the loop position."""
    err = C.create_string_buffer(4096)
    q0 = np.ascontiguousarray(q0, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    if host_lib().howmhd_host_write_checkpoint(C.byref(run), str(path).encode(), _ptr(q0), _ptr(b), out_step,
                                               total_steps, time, err, len(err)) != 0:
        raise HowmhdError(err.value.decode())


def read_checkpoint(run: Run, path):
    """howmhd_host_read_checkpoint -> (q0, faces, out_step, total_steps, time)."""
    err = C.create_string_buffer(4096)
    shape = (run.nx + 6, run.ny + 6, run.nz + 6)
    q0 = np.zeros((8,) + shape)
    b = np.zeros((3,) + shape)
    o, n, t = C.c_int32(), C.c_int64(), C.c_double()
    if host_lib().howmhd_host_read_checkpoint(C.byref(run), str(path).encode(), _ptr(q0), _ptr(b), C.byref(o),
                                              C.byref(n), C.byref(t), err, len(err)) != 0:
        raise HowmhdError(err.value.decode())
    return q0, b, o.value, n.value, t.value


def make_input(**kw) -> str:
    """Input-file text from keyword=value pairs (reference keywords)."""
    return "\n".join(f"{k} {v}" for k, v in kw.items()) + "\n"


# ------------------------------------------------------------------ GPU solver
class GpuSolver:
    """One device context of the C-ABI: q(0) and faces stay resident."""

    def __init__(self, params: Params):
        self.lib = gpu_lib()
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

    # ---- learned indicator (include/howmhd_mlp.h)
    def n_samples(self, d: int) -> int:
        """Interfaces the sweep along axis d computes (0 for a degenerate axis)."""
        n = (self.p.nx, self.p.ny, self.p.nz)[d]
        if n <= 1:
            return 0
        cells = self.p.nx * self.p.ny * self.p.nz
        return cells // n * (n if self.p.bc[d] == BC["periodic"] else n + 1)

    def collect_samples(self, d: int, X, y) -> int:
        """Features (n, 18) and labels (n,) of the current state into the
        float32 device tensors X, y (torch); returns the sample count."""
        _check_dev(X, (None, MLP_NFEAT))
        _check_dev(y, (X.shape[0],))
        _torch().cuda.synchronize()
        cnt = C.c_int64()
        self._check(self.lib.howmhd_gpu_collect_samples(self.h, d, X.data_ptr(), y.data_ptr(), X.shape[0],
                                                        C.byref(cnt)), "collect_samples")
        return cnt.value

    def indicator_logits(self, d: int, z) -> int:
        """The in-sweep network's logits of every interface along axis d into
        the float32 device tensor z (torch); returns the count."""
        _check_dev(z, (None,))
        _torch().cuda.synchronize()
        cnt = C.c_int64()
        self._check(self.lib.howmhd_gpu_indicator_logits(self.h, d, z.data_ptr(), z.shape[0], C.byref(cnt)),
                    "indicator_logits")
        return cnt.value

    def set_indicator(self, params, threshold: float = 0.5, enable: int = 1):
        """Routing with the plain-WENO fallback (1), routing on every step (2)
        or off (0), with host weights."""
        if params is None:
            self._check(self.lib.howmhd_gpu_set_indicator(self.h, None, threshold, enable), "set_indicator")
            return
        P = np.ascontiguousarray(np.asarray(params, dtype=np.float32))
        assert P.shape == (MLP_NPARAM,)
        self._check(self.lib.howmhd_gpu_set_indicator(self.h, P.ctypes.data_as(C.POINTER(C.c_float)), threshold,
                                                      enable), "set_indicator")

    def indicator_stats(self) -> dict:
        v = [C.c_int64() for _ in range(4)]
        self._check(self.lib.howmhd_gpu_indicator_stats(self.h, *[C.byref(x) for x in v]), "indicator_stats")
        f, t, wf, wt = (x.value for x in v)
        r = [C.c_int64() for _ in range(2)]
        self._check(self.lib.howmhd_gpu_indicator_routing(self.h, *[C.byref(x) for x in r]), "indicator_routing")
        return {"flagged": f, "total": t, "waves_flagged": wf, "waves_total": wt,
                "flagged_fraction": f / t if t else 0.0, "wave_fraction": wf / wt if wt else 0.0,
                "routed_steps": r[0].value, "fallback_steps": r[1].value}

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


# ------------------------------------------------------- indicator kernels
def _torch():
    import torch
    return torch


def _check_dev(t, shape):
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise HowmhdError("expected a contiguous float32 CUDA tensor")
    if len(shape) != t.dim() or any(s is not None and s != d for s, d in zip(shape, t.shape)):
        raise HowmhdError(f"tensor shape {tuple(t.shape)} does not match {shape}")


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def mlp_infer(X, P):
    """Logits z3 of the samples X (n, 18) with parameters P (1697,), both
    float32 CUDA tensors; runs on torch's current stream."""
    torch = _torch()
    _check_dev(X, (None, MLP_NFEAT))
    _check_dev(P, (MLP_NPARAM,))
    out = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
    rc = gpu_lib().howmhd_mlp_infer(X.data_ptr(), X.shape[0], P.data_ptr(), out.data_ptr(), _stream())
    if rc:
        raise HowmhdError(f"howmhd_mlp_infer failed ({rc})")
    return out


def mlp_grad(X, y, P, pos_weight=1.0, scale=1.0, G=None, loss=None):
    """Adds the gradient (and loss) of the weighted BCE over (X, y) to G
    (and loss); zero-initialised tensors are made when not given."""
    torch = _torch()
    _check_dev(X, (None, MLP_NFEAT))
    _check_dev(y, (X.shape[0],))
    _check_dev(P, (MLP_NPARAM,))
    if G is None:
        G = torch.zeros(MLP_NPARAM, dtype=torch.float32, device=X.device)
    if loss is None:
        loss = torch.zeros(1, dtype=torch.float32, device=X.device)
    rc = gpu_lib().howmhd_mlp_grad(X.data_ptr(), y.data_ptr(), X.shape[0], P.data_ptr(), pos_weight, scale,
                                        G.data_ptr(), loss.data_ptr(), _stream())
    if rc:
        raise HowmhdError(f"howmhd_mlp_grad failed ({rc})")
    return G, loss


def mlp_adam(P, M, V, G, t: int, lr: float, beta1=0.9, beta2=0.999, eps=1e-8):
    for a in (P, M, V, G):
        _check_dev(a, (MLP_NPARAM,))
    rc = gpu_lib().howmhd_mlp_adam(P.data_ptr(), M.data_ptr(), V.data_ptr(), G.data_ptr(), t, lr, beta1, beta2,
                                        eps, _stream())
    if rc:
        raise HowmhdError(f"howmhd_mlp_adam failed ({rc})")


def tstep_noalloc(params: Params, q0_dev) -> float:
    """howmhd_gpu_tstep_noalloc: the CFL dt of q0_dev, a contiguous float64
    CUDA tensor of shape (8, nx+6, ny+6, nz+6) in the reference layout, on
    torch's current stream; allocates no device memory."""
    torch = _torch()
    shape = (8, params.nx + 6, params.ny + 6, params.nz + 6)
    if not (isinstance(q0_dev, torch.Tensor) and q0_dev.is_cuda and q0_dev.dtype == torch.float64
            and q0_dev.is_contiguous() and tuple(q0_dev.shape) == shape):
        raise HowmhdError(f"expected a contiguous float64 CUDA tensor of shape {shape}")
    dt = C.c_double()
    lib = gpu_lib()
    rc = lib.howmhd_gpu_tstep_noalloc(C.byref(params), C.c_void_p(q0_dev.data_ptr()), _stream(), C.byref(dt))
    if rc:
        raise HowmhdError(f"howmhd_gpu_tstep_noalloc failed ({rc}): {lib.howmhd_gpu_last_error(None).decode()}")
    return dt.value


def build_kind() -> str:
    return gpu_lib().howmhd_gpu_build_kind().decode()


def active(a: np.ndarray) -> np.ndarray:
    """Active cells of a reference-layout array (leading component axis kept)."""
    return a[:, 3:-3, 3:-3, 3:-3]


def rel_l1(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Relative L1 difference per component over active cells: sum|a-b| / sum|b|."""
    a, b = active(a), active(b)
    num = np.abs(a - b).reshape(a.shape[0], -1).sum(axis=1)
    den = np.abs(b).reshape(b.shape[0], -1).sum(axis=1)
    return np.where(den > 0, num / np.where(den > 0, den, 1.0), num)
