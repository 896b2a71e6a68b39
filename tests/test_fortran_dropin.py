"""The Fortran drop-in modules (fvs2d_amd/fortran) driven like the reference
drives its own: data_solution_init, gradient_init, runge_kutta_init,
compute_residual, time_integration, gradient_cellcntr_1var.  The Fortran test program is a
shared library (tests/fortran, built by __graft_entry__.build) called in this
process; its results must equal the Python binding's bit for bit (same library,
same options), and the Python binding is checked against the oracle elsewhere.
The log files the drop-in writes (log_res.plt, log_vortex_err.plt,
log_vortex_err_xy.plt) carry the reference's headers and rows, the vortex rows
checked against the oracle's error_isentropic_vortex; the phase timers
(cput_*) are filled.

Where the reference sources are present (this build container), the drop-in
also compiles against the reference's own mainparam and data_grid modules
(src/mainparam.f90, src/data_grid.f90, compiled where they lie into a
temporary directory); its input module needs MPI, so the test harness's
reduced input module stands beside them."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from fvs2d_amd import meshgen
from fvs2d_amd.solver import Solver
from oracle.oracle import Oracle
from tests.common import vortex_scheme

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "libfvs2d_fdriver.so")


def test_driver_built_and_exports():
    if not os.path.exists(DRIVER):
        pytest.skip("amdflang not available: Fortran harness not built")
    L = C.CDLL(DRIVER)
    assert hasattr(L, "fvs2d_fortran_driver")
    for name in ("fvs_create", "fvs_time_integration"):
        assert hasattr(L, name)  # resolved through libfvs2d_hip.so


@pytest.mark.gpu
@pytest.mark.parametrize("limiter", [0, 1])
def test_fortran_dropin_matches_c_abi(tmp_path, limiter):
    if not os.path.exists(DRIVER):
        pytest.skip("amdflang not available: Fortran harness not built")
    m = meshgen.jittered_square(24, seed=5)
    dt = 0.004
    nsteps = 20
    sch = vortex_scheme(rk_scheme=0, rk_order=4, dt=dt, limiter=limiter, reorder=1, kernel_path=0)
    py = Solver(m, sch)
    py.initialize_solution(1, 0.0)
    cv0 = py.cvar
    r_py, _ = py.compute_residual(0.0)
    py_norms = py.time_integration(0.0, nsteps, vortex_error=True)
    cv_py = py.cvar
    dp_py = py.gradient_cellcntr_1var(py.cvar2pvar()[:, 3])

    meshgen.write_grid(m, str(tmp_path / "mesh"))
    (tmp_path / "fvs2d.vortex").write_text("0.0, 0.0\n5.0\n1.0\n1.0\n1.0\n1.0\n")
    (tmp_path / "flags.nml").write_text(
        f"&flags grad_method=3, lsq_nn=1, lsq_p=0.0, limiter={limiter}, recon=2, kappa=0.0, mach=0.5, aoa=0.0,"
        f" gam=1.4, dtin={dt!r}, cfl=0.5, ssp=0, steady=0, vortex=1 /\n")
    cv0.astype("<f8").tofile(tmp_path / "cvar0.bin")
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        C.CDLL(DRIVER).fvs2d_fortran_driver(C.c_int(nsteps))
    finally:
        os.chdir(cwd)
    N = m.ncells
    r_f = np.fromfile(tmp_path / "resid.bin", "<f8").reshape(N, 4)
    cv_f = np.fromfile(tmp_path / "cvar.bin", "<f8").reshape(N, 4)
    dp_f = np.fromfile(tmp_path / "dp.bin", "<f8").reshape(2, N)
    assert np.array_equal(r_f, r_py)
    assert np.array_equal(cv_f, cv_py)
    assert np.array_equal(dp_f, dp_py)
    log = (tmp_path / "log_res.plt").read_text().splitlines()
    assert len(log) == 1 + nsteps
    assert log[0] == RES_HEADER
    norms = py_norms[0]
    for k, line in enumerate(log[1:]):
        f = line.split()
        assert int(f[0]) == k + 1
        assert np.allclose([float(v) for v in f[1:]], norms[k], rtol=1e-7, atol=0)
    # error_isentropic_vortex's files against the oracle's record
    orc = Oracle(m, **sch.oracle_kwargs())
    orc.cvar = cv0
    _, v_ref = orc.time_integration(0.0, nsteps, want_verr=True)
    vlog = (tmp_path / "log_vortex_err.plt").read_text().splitlines()
    assert vlog[0] == 'variables = "t", ' and vlog[5] == '"Q<sub>L2</sub>"'
    rows = np.array([[float(v) for v in ln.split()] for ln in vlog[6:]])
    assert rows.shape == (nsteps, 14)
    assert np.allclose(rows[:, 0], dt * np.arange(1, nsteps + 1), rtol=1e-7)
    assert np.allclose(rows[:, 1:], v_ref[:, :13], rtol=1e-7, atol=0)
    xy = (tmp_path / "log_vortex_err_xy.plt").read_text().splitlines()
    assert xy[0] == 'variables = "t", "x" "y"'
    xyr = np.array([[float(v) for v in ln.split()] for ln in xy[1:]])
    assert xyr.shape == (nsteps, 3)
    assert np.mean(np.all(np.isclose(xyr[:, 1:], v_ref[:, 13:15], rtol=1e-12, atol=1e-12), axis=1)) >= 0.9
    cput = [float(v) for v in (tmp_path / "cput.txt").read_text().split()]
    assert cput[0] > 0 and cput[1] == 0 and cput[2] > 0 and cput[3] > 0 and all(np.isfinite(cput))


RES_HEADER = ('variables = "iteration" "|<greek>\\r</greek>|<sub>2</sub>",  "|<greek>\\r</greek>u|<sub>2</sub>", '
              '"|<greek>\\r</greek>v|<sub>2</sub>", "|<greek>\\r</greek>E|<sub>2</sub>"')

REF_SRC = "/root/reference/src"
FLANG = "/opt/rocm/bin/amdflang"


def test_dropin_compiles_against_reference_modules(tmp_path):
    ref = [os.path.join(REF_SRC, n) for n in ("mainparam.f90", "data_grid.f90")]
    if not all(os.path.exists(p) for p in ref) or not os.path.exists(FLANG):
        pytest.skip("reference sources or amdflang not present")
    harness = open(os.path.join(ROOT, "tests", "fortran", "harness_modules.f90")).read()
    inp = re.search(r"^module input\b.*?^end module input", harness, re.S | re.M).group(0)
    (tmp_path / "input_harness.f90").write_text(inp + "\n")
    srcs = [ref[0], str(tmp_path / "input_harness.f90"), ref[1],
            os.path.join(ROOT, "fvs2d_amd", "fortran", "fvs2d_capi.f90"),
            os.path.join(ROOT, "fvs2d_amd", "fortran", "fvs2d_dropin.f90")]
    for s in srcs:
        cmd = [FLANG, "-O0", "-fPIC", "-fdefault-real-8", "-fdefault-double-8", "-module-dir", str(tmp_path), "-I",
               str(tmp_path), "-c", s, "-o", str(tmp_path / (os.path.basename(s) + ".o"))]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
