"""Persistent whole-step launches (fvs_set_persistent): one cooperative
launch per time-integration call, the whole-stage block body walked over the
launch blocks with a grid barrier after every stage and the step norms
reduced in the same launch.  Bit for bit the per-stage launches of the same
kernel (kernel_path 2): states and norms on a steady O-grid with the limiter,
on a ragged batch with masked instances, per-instance frozen limiters and
per-instance norms, and on SSP-RK3 (three stages per step, so the primitive
buffers end swapped after an odd step count)."""
import numpy as np
import pytest

from fvs2d_amd import _lib, cases, meshgen
from fvs2d_amd.solver import BatchSolver, Solver
from tests.common import vortex_scheme

pytestmark = pytest.mark.gpu


def test_persistent_needs_whole_stage_kernel():
    m, sch = cases.naca(32, 16)
    s = Solver(m, sch)  # kernel_path 0
    with pytest.raises(_lib.FvsError, match="kernel_path 2"):
        s.set_persistent(True)


def test_persistent_single_steady_bitwise():
    m, sch = cases.naca(64, 32, kernel_path=2)
    a, b = Solver(m, sch), Solver(m, sch)
    b.set_persistent(True)
    for s in (a, b):
        s.initialize_solution(1, 0.0)
    for n in (150, 73):
        na, nb = a.time_integration(0.0, n), b.time_integration(0.0, n)
        assert np.array_equal(na, nb)
        assert np.array_equal(a.cvar, b.cvar)
    for s in (a, b):
        s.freeze_limiter(True)
    na, nb = a.time_integration(0.0, 40), b.time_integration(0.0, 40)
    assert np.array_equal(na, nb) and np.array_equal(a.cvar, b.cvar)
    assert np.array_equal(a.wall_forces(), b.wall_forces())


def test_persistent_batch_masked_bitwise():
    conds = [((32, 16), (0.7, 1.5)), ((48, 24), (0.6, -1.0)), ((40, 20), (0.78, 2.5))]
    meshes = [cases.naca(n, l, kernel_path=2)[0] for (n, l), _ in conds]
    sch = cases.naca(32, 16, kernel_path=2)[1]
    a = BatchSolver(meshes, [ma for _, ma in conds], sch)
    b = BatchSolver(meshes, [ma for _, ma in conds], sch)
    b.set_persistent(True)
    for s in (a, b):
        s.initialize_freestream()
    na, nb = a.batch_time_integration(0.0, 60), b.batch_time_integration(0.0, 60)
    assert np.array_equal(na, nb) and np.array_equal(a.cvar, b.cvar)
    for s in (a, b):
        s.set_active(np.array([1, 0, 1], np.int32))
        s.freeze_limiter_batch(np.array([0, 0, 1], np.int32))
    na, nb = a.batch_time_integration(0.0, 45), b.batch_time_integration(0.0, 45)
    assert np.array_equal(na[:, [0, 2]], nb[:, [0, 2]]) and np.array_equal(a.cvar, b.cvar)
    assert np.array_equal(a.wall_forces_batch(), b.wall_forces_batch())


def test_persistent_ssp3_odd_steps_bitwise():
    m = meshgen.jittered_square(32, periodic=True, seed=1)
    sch = vortex_scheme(dt=0.01, reorder=1, kernel_path=2)
    a, b = Solver(m, sch), Solver(m, sch)
    b.set_persistent(True)
    for s in (a, b):
        s.initialize_solution(1, 0.0)
    for n in (7, 4, 9):
        na, nb = a.time_integration(0.0, n), b.time_integration(0.0, n)
        assert np.array_equal(na, nb)
        assert np.array_equal(a.cvar, b.cvar)
    assert np.array_equal(a.cvar2pvar(), b.cvar2pvar())
