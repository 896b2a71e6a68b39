"""The Fortran drop-in modules (fvs2d_amd/fortran) driven like the reference
drives its own: data_solution_init, gradient_init, runge_kutta_init,
compute_residual, time_integration, gradient_cellcntr_1var.  The Fortran test program is a
shared library (tests/fortran, built by __graft_entry__.build) called in this
process; its results must equal the Python binding's bit for bit (same library,
same options), and the Python binding is checked against the oracle elsewhere."""
import ctypes as C
import os

import numpy as np
import pytest

from fvs2d_amd import meshgen
from fvs2d_amd.solver import Solver
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
    py.time_integration(0.0, nsteps)
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
