"""Host shell and C-ABI surface (CPU only).

The parser, problem registry and writers restate src/parameters.cpp,
src/problem.cpp, src/problems/*.cpp and src/output.cpp; these tests check their
documented behaviour on small inputs, that the libraries load, and that every
function declared in include/*.h is exported.
"""
import math
import re
import struct
from pathlib import Path

import numpy as np
import pytest

import howmhd as H

ROOT = Path(__file__).resolve().parents[1]


def declared_functions(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(howmhd_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,libs", [
    ("howmhd_gpu.h", ["libhowmhd_gpu.so"]),
    ("howmhd_host.h", ["libhowmhd_host.so"]),
    ("howmhd_mlp.h", ["libhowmhd_gpu.so"]),
])
def test_every_declared_symbol_is_exported(header, libs):
    import ctypes
    names = declared_functions(ROOT / "include" / header)
    assert names
    for lib in libs:
        dll = ctypes.CDLL(str(H.LIB_DIR / lib))
        missing = [n for n in names if not hasattr(dll, n)]
        assert not missing, (lib, missing)


def test_build_kinds():
    assert "fp64 reference rounding" in H.build_kind()


def test_create_fails_cleanly_without_gpu_or_with_bad_params():
    run = H.parse(H.make_input(problem="uniform", nx=2, ny=4, nz=4, nt=1, tend=1))
    p = H.gpu_params(run)
    with pytest.raises(H.HowmhdError, match="n == 1 or n >= 3"):
        H.GpuSolver(p)


@pytest.mark.parametrize("shape", [(4, 70000, 1), (4, 1, 65536), (1, 3, 70000)])
def test_create_rejects_axes_beyond_the_launch_grid(shape):
    """ny and nz go into gridDim.y / gridDim.z of the sweep, EMF and face
    launches; create refuses lengths above 65535 before any HIP call (so this
    runs without a GPU) instead of failing in the middle of a step."""
    nx, ny, nz = shape
    run = H.parse(H.make_input(problem="uniform", nx=nx, ny=ny, nz=nz, nt=1, tend=1))
    with pytest.raises(H.HowmhdError, match="must not exceed 65535"):
        H.GpuSolver(H.gpu_params(run))


# ----------------------------------------------------------------- parser
def test_parser_defaults_and_keys():
    text = """# comment line
problem orszag_tang
nx 16 ny 8
nz 1   # trailing comment is a token starting with #
nt 2
tend 0.5
gam 1.4
cour 0.8
rhomin 1e-8
x1bc open
output_format vts
"""
    r = H.parse(text)
    assert (r.gnx, r.gny, r.gnz) == (16, 8, 1)
    assert (r.nx, r.ny, r.nz) == (16, 8, 1)
    assert r.problem == b"orszag_tang"
    assert r.gam == 1.4 and r.cour == 0.8 and r.tend == 0.5 and r.nt == 2
    assert r.rhomin == 1e-8 and r.pgmin == 1e-12          # default pgmin
    assert r.x1bc == b"open" and r.x2bc == b"periodic"
    assert r.output_format == b"vts"
    assert r.dx == 1.0 / 16 and r.dy == 1.0 / 8 and r.dz == 1.0
    assert r.precision == b"fp64" and r.indicator == b"off"


def test_parser_nstep_parsed_and_max_steps():
    """nstep is read and otherwise unused, as in the reference
    (src/parameters.cpp:58-59; nothing reads par.nstep); the early stop of the
    driver has its own key, max_steps."""
    r = H.parse("nx 4 ny 4 nz 4 nt 1 tend 1 nstep 5")
    assert r.nstep == 5 and r.max_steps == 0
    r = H.parse("nx 4 ny 4 nz 4 nt 1 tend 1 max_steps 7")
    assert r.nstep == 0 and r.max_steps == 7


def test_parser_gnx_alias_and_unknown_key(capfd):
    r = H.parse("gnx 8 gny 4 gnz 2 nt 1 tend 1 bogus 3 4 5\nxsize 2\n")
    assert (r.nx, r.ny, r.nz) == (8, 4, 2)
    assert r.xsize == 2.0 and r.dx == 0.25
    assert "Warning: unknown parameter bogus" in capfd.readouterr().err


@pytest.mark.parametrize("text,msg", [
    ("nx 0 ny 1 nz 1 nt 1 tend 1", "global grid size must be positive"),
    ("nx 4 ny 4 nz 4 nt 0 tend 1", "nt must be positive"),
    ("nx 4 ny 4 nz 4 nt 1 tend 0", "tend must be positive"),
])
def test_parser_errors(text, msg):
    with pytest.raises(H.HowmhdError, match=msg):
        H.parse(text)


def test_unknown_bc_and_problem():
    r = H.parse("nx 4 ny 4 nz 4 nt 1 tend 1 x2bc reflecting problem nope")
    with pytest.raises(H.HowmhdError, match="unknown x2bc = reflecting"):
        H.gpu_params(r)
    with pytest.raises(H.HowmhdError, match="unknown problem = nope") as e:
        H.init_problem(r)
    for name in ("uniform", "orszag_tang", "brio_wu", "mhd_rotor", "blast3d", "mpi_smooth3d"):
        assert name in str(e.value)


# --------------------------------------------------------------- problems
def test_registry_names():
    names = H.problem_names()
    for name in ("uniform", "orszag_tang", "brio_wu", "mhd_rotor", "blast3d", "mpi_smooth3d",
                 "cpaw3d", "mhd_blast3d", "turbulence3d"):
        assert name in names


def test_orszag_tang_closed_form():
    r = H.parse(H.make_input(problem="orszag_tang", nx=8, ny=8, nz=1, nt=1, tend=1))
    q = H.init_problem(r)
    pi = math.pi
    i, j = 3 + 2, 3 + 5
    x, y = (2 + 0.5) / 8, (5 + 0.5) / 8
    rho = 25 / (36 * pi)
    assert q[0, i, j, 3] == rho
    assert q[1, i, j, 3] == rho * -math.sin(2 * pi * y)
    assert q[4, i, j, 3] == -math.sin(2 * pi * y) / math.sqrt(4 * pi)
    assert q[5, i, j, 3] == math.sin(4 * pi * x) / math.sqrt(4 * pi)
    vv2 = math.sin(2 * pi * y) ** 2 + math.sin(2 * pi * x) ** 2
    BB2 = q[4, i, j, 3] ** 2 + q[5, i, j, 3] ** 2
    E = 5 / (12 * pi) / (5 / 3 - 1) + 0.5 * rho * vv2 + 0.5 * BB2
    assert abs(q[7, i, j, 3] - E) < 1e-15
    # ghosts untouched
    assert np.all(q[:, :3] == 0) and np.all(q[:, :, :, :3] == 0)


def test_blast_variants():
    for prob, pin, pout, B in (("blast3d", 100.0, 1.0, (0.1, 0.0, 0.0)),
                               ("mhd_blast3d", 10.0, 0.1, (1 / math.sqrt(2), 1 / math.sqrt(2), 0.0))):
        r = H.parse(H.make_input(problem=prob, nx=20, ny=20, nz=20, nt=1, tend=1))
        q = H.active(H.init_problem(r))
        assert np.all(q[0] == 1.0)
        BB2 = sum(b * b for b in B)
        p = (2.0 / 3.0) * (q[7] - 0.5 * BB2)
        assert set(np.round(np.unique(p), 12)) == {round(pout, 12), round(pin, 12)}
        c = q[7, 10, 10, 10]   # cell centre at (10.5/20) is inside r < 0.1
        assert abs((2.0 / 3.0) * (c - 0.5 * BB2) - pin) < 1e-12
        assert np.allclose(q[4], B[0]) and np.allclose(q[5], B[1])


def test_cpaw3d_is_divergence_free_and_alfvenic():
    r = H.parse(H.make_input(problem="cpaw3d", nx=12, ny=12, nz=12, xsize=3, ysize=1.5, zsize=1.5, nt=1, tend=1))
    q = H.active(H.init_problem(r))
    B = q[4:7]
    v = q[1:4] / q[0]
    e1 = np.array([1, 2, 2]) / 3
    bpar = np.einsum("d...,d->...", B, e1)
    assert np.allclose(bpar, 1.0)
    assert np.allclose(v + (B - e1[:, None, None, None]), 0.0)      # v_perp = -B_perp
    perp = B - e1[:, None, None, None]
    assert np.allclose(np.sqrt((perp ** 2).sum(axis=0)), 0.1)


def test_turbulence3d_seeded_and_normalised():
    kw = dict(problem="turbulence3d", nx=16, ny=16, nz=16, nt=1, tend=1)
    a = H.active(H.init_problem(H.parse(H.make_input(**kw, seed=7))))
    b = H.active(H.init_problem(H.parse(H.make_input(**kw, seed=7))))
    c = H.active(H.init_problem(H.parse(H.make_input(**kw, seed=8))))
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    v = a[1:4] / a[0]
    assert abs(np.sqrt((v ** 2).sum(axis=0).mean()) - 1.0) < 1e-12
    # solenoidal: spectral divergence is zero
    vk = [np.fft.fftn(v[d]) for d in range(3)]
    k = np.fft.fftfreq(16) * 16
    KX, KY, KZ = np.meshgrid(k, k, k, indexing="ij")
    div = KX * vk[0] + KY * vk[1] + KZ * vk[2]
    assert np.abs(div).max() < 1e-9 * np.abs(vk[0]).max()
    p = (5 / 3 - 1) * (a[7] - 0.5 * (a[1] ** 2 + a[2] ** 2 + a[3] ** 2) / a[0] - 0.5 * a[6] ** 2)
    assert np.allclose(p, 0.6)                       # Mach 1 at v_rms = 1
    assert np.allclose(a[6], math.sqrt(2 * 0.6))     # beta = 2p/B^2 = 1


# ---------------------------------------------------------------- writers
def _state():
    r = H.parse(H.make_input(problem="mpi_smooth3d", nx=5, ny=4, nz=3, nt=1, tend=1, output_format="vts"))
    q = H.init_problem(r)
    # periodic ghosts, as bound(q,0) leaves them
    for ax in (1, 2, 3):
        n = q.shape[ax] - 6
        idx = [(c - 3) % n + 3 for c in range(n + 6)]
        q = np.take(q, idx, axis=ax)
    return r, np.ascontiguousarray(q)


def test_vts_writer_layout(tmp_path):
    r, q = _state()
    H.write_output(r, q, 7, 0.125, str(tmp_path))
    f = tmp_path / "output" / "dump000007_rank000000.vts"
    raw = f.read_bytes()
    head, _, tail = raw.partition(b'<AppendedData encoding="raw">\n_')
    text = head.decode()
    assert 'WholeExtent="0 4 0 3 0 2"' in text and 'header_type="UInt64"' in text
    assert "1.250000000000000e-01" in text
    names = re.findall(r'Name="(\w+)" NumberOfComponents="(\d)" format="appended" offset="(\d+)"', text)
    assert [n for n, _, _ in names] == ["rho", "pressure", "energy", "v2", "B2", "divB", "velocity",
                                        "magnetic_field", "momentum"]
    npts = 5 * 4 * 3
    a = H.active(q)

    def block(off, ncomp):
        nbytes = struct.unpack_from("<Q", tail, off)[0]
        assert nbytes == 8 * npts * ncomp
        return np.frombuffer(tail, dtype="<f8", count=npts * ncomp, offset=off + 8)

    offs = {n: (int(o), int(c)) for n, c, o in names}
    rho = block(*offs["rho"]).reshape(3, 4, 5)          # i fastest
    assert np.array_equal(rho, a[0].transpose(2, 1, 0))
    mom = block(*offs["momentum"]).reshape(3, 4, 5, 3)
    assert np.array_equal(mom[..., 1], a[2].transpose(2, 1, 0))
    pts_off = int(re.search(r'<Points>\s*<DataArray type="Float64" NumberOfComponents="3" format="appended" '
                            r'offset="(\d+)"', text).group(1))
    pts = block(pts_off, 3).reshape(3, 4, 5, 3)
    assert pts[0, 0, 1, 0] == 1.5 * r.dx and pts[2, 3, 4, 2] == 2.5 * r.dz
    assert raw.endswith(b"\n  </AppendedData>\n</VTKFile>\n")


def test_dat_writer_columns(tmp_path):
    r, q = _state()
    r.output_format = b"dat"
    H.write_output(r, q, 3, 0.5, str(tmp_path))
    lines = (tmp_path / "output" / "dump000003_rank000000.dat").read_text().splitlines()
    assert lines[0] == "# step = 3" and lines[1].startswith("# time = 5.000000000000000e-01")
    assert lines[6].startswith("# columns: ig jg kg il jl kl x y z rho")
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[7:]])
    assert rows.shape == (60, 22)
    a = H.active(q)
    # second row: ig=1, jg=0, kg=0
    assert rows[1, 0] == 1 and rows[1, 9] == float(f"{a[0, 1, 0, 0]:.15e}")
    # periodic centred divB of the cell-centred field
    i, j, k = 2, 1, 1
    dv = ((a[4, 3, j, k] - a[4, 1, j, k]) / (2 * r.dx) + (a[5, i, 2, k] - a[5, i, 0, k]) / (2 * r.dy)
          + (a[6, i, j, 2] - a[6, i, j, 0]) / (2 * r.dz))
    row = rows[(k * 4 + j) * 5 + i]
    assert abs(row[21] - dv) <= 1e-12 * max(1.0, abs(dv))


def test_output_none_and_unknown(tmp_path):
    r, q = _state()
    r.output_format = b"none"
    H.write_output(r, q, 0, 0.0, str(tmp_path))
    assert not (tmp_path / "output").exists()
    r.output_format = b"hdf5"
    with pytest.raises(H.HowmhdError, match="unknown output_format = hdf5"):
        H.write_output(r, q, 0, 0.0, str(tmp_path))


def test_indicator_weights_roundtrip(tmp_path):
    import train_indicator as T
    rng = np.random.default_rng(0)
    P = rng.normal(size=H.MLP_NPARAM).astype(np.float32)
    path = tmp_path / "w.txt"
    T.write_weights(path, P, 0.37, {"note": "test"})
    Q, thr = H.load_indicator(path)
    np.testing.assert_array_equal(P, Q)
    assert abs(thr - 0.37) < 1e-7
    Q2, thr2 = T.read_weights(path)
    np.testing.assert_array_equal(P, Q2)
    bad = tmp_path / "bad.txt"
    bad.write_text("howmhd-mlp 1697 0.5\n1.0\n")
    with pytest.raises(H.HowmhdError, match="1 of 1697"):
        H.load_indicator(bad)
    bad.write_text("howmhd-mlp 1697 0.5\n" + "1.0 2.0\n" * 1697)
    with pytest.raises(H.HowmhdError, match="bad or extra value"):
        H.load_indicator(bad)
    bad.write_text("howmhd-mlp 12 0.5\n")
    with pytest.raises(H.HowmhdError, match="expected 1697"):
        H.load_indicator(bad)
    bad.write_text("howmhd-mlp 1697 1.5\n")
    with pytest.raises(H.HowmhdError, match="threshold"):
        H.load_indicator(bad)


def test_checkpoint_roundtrip_and_checks(tmp_path):
    """Checkpoint files (no reference counterpart): the state and loop position
    come back bit for bit; a file of another grid, a file that is not a
    checkpoint and a truncated one are refused with a message."""
    run = H.parse(H.make_input(problem="orszag_tang", nx=10, ny=7, nz=1, nt=2, tend=1.0, checkpoint_every=1,
                               restart="prev.bin"))
    assert run.checkpoint_every == 1 and run.restart == b"prev.bin"
    rng = np.random.default_rng(7)
    shape = (run.nx + 6, run.ny + 6, run.nz + 6)
    q, b = rng.standard_normal((8,) + shape), rng.standard_normal((3,) + shape)
    q[0, 3, 3, 3] = -0.0
    path = tmp_path / "c.bin"
    H.write_checkpoint(run, path, q, b, 3, 1234567890123, 0.1 + 0.2)
    assert path.stat().st_size == 48 + 8 * 11 * q[0].size
    q2, b2, o, n, t = H.read_checkpoint(run, path)
    assert np.array_equal(q.view(np.int64), q2.view(np.int64)) and np.array_equal(b, b2)
    assert (o, n, t) == (3, 1234567890123, 0.1 + 0.2)
    other = H.parse(H.make_input(problem="orszag_tang", nx=10, ny=8, nz=1, nt=2, tend=1.0))
    with pytest.raises(H.HowmhdError, match="does not match"):
        H.read_checkpoint(other, path)
    junk = tmp_path / "junk.bin"
    junk.write_bytes(b"not a checkpoint at all" * 4)
    with pytest.raises(H.HowmhdError, match="not a checkpoint"):
        H.read_checkpoint(run, junk)
    cut = tmp_path / "cut.bin"
    cut.write_bytes(path.read_bytes()[:1000])
    with pytest.raises(H.HowmhdError, match="truncated"):
        H.read_checkpoint(run, cut)
    with pytest.raises(H.HowmhdError, match="cannot open"):
        H.read_checkpoint(run, tmp_path / "missing.bin")


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message before
    touching a device."""
    import subprocess
    import sys
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 2
    assert "measures one GPU" in out.stderr and out.stdout == ""


# ------------------------------------------------------------------ driver
DRIVER = ROOT / "how-mhd-kokkos_amd" / "bin" / "how_mhd"


@pytest.mark.skipif(not DRIVER.exists(), reason="driver not built")
@pytest.mark.parametrize("text,msg", [
    (None, "Error: cannot open input file"),                                    # src/main.cpp:64-79
    ("nx 0 ny 1 nz 1 nt 1 tend 1\n", "global grid size must be positive"),      # parameters.cpp:121-137
    ("nx 4 ny 4 nz 4 nt 1 tend 1 problem nope\n", "unknown problem = nope"),    # problem.cpp:24-34
    ("nx 4 ny 4 nz 4 nt 1 tend 1 problem uniform output_format xyz\n",
     "Error: unknown output_format = xyz"),                                     # output.cpp:565-569
    ("nx 4 ny 4 nz 4 nt 1 tend 1 problem uniform x3bc reflecting\n", "unknown x3bc = reflecting"),  # bound.cpp:305-315
])
def test_driver_aborts_on_bad_input(tmp_path, text, msg):
    """The drop-in driver stops with the reference's message and an abort
    (SIGABRT) on each input error, before it touches a GPU."""
    import signal
    import subprocess
    path = tmp_path / "input.in"
    if text is not None:
        path.write_text(text)
    r = subprocess.run([str(DRIVER), str(path)], capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert r.returncode == -signal.SIGABRT, (r.returncode, r.stdout[-500:], r.stderr[-500:])
    assert msg in r.stderr, r.stderr[-1000:]
    assert not (tmp_path / "output").exists()
