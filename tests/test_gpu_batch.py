"""Batched ragged solves and graph replay on the GPU.

* batched == single: each instance of a batch (different resolutions, Mach
  numbers and angles of attack, one launch per stage for all of them) ends in
  exactly the state its own single-instance solve reaches, and has the same
  forces;
* masking: an instance taken out of the launches keeps its state bit for bit
  while the others continue unchanged;
* graph replay gives bitwise the same states as stream launches.
"""
import numpy as np
import pytest

from fvs2d_amd import cases
from fvs2d_amd.solver import BatchSolver, Solver
from tests.common import rel_err, vortex_scheme

pytestmark = pytest.mark.gpu

INST = [((48, 24), (0.75, 1.0)), ((64, 32), (0.8, 1.25)), ((32, 16), (0.6, -2.0))]


def single(k, steps, freeze_after=None):
    (n, l), (M, a) = INST[k]
    m, sch = cases.naca(n, l, mach=M, aoa=a, cfl=1.0)
    s = Solver(m, sch)
    s.initialize_solution(1, 0.0)
    if freeze_after is not None:
        s.time_integration(0.0, freeze_after)
        s.freeze_limiter(True)
        s.time_integration(0.0, steps - freeze_after)
    else:
        s.time_integration(0.0, steps)
    return s


def batch():
    meshes = [cases.naca(n, l, cfl=1.0)[0] for (n, l), _ in INST]
    _, sch = cases.naca(16, 8, cfl=1.0)
    b = BatchSolver(meshes, [ma for _, ma in INST], sch)
    b.initialize_freestream()
    return b


def test_batched_equals_single():
    b = batch()
    nb = b.batch_time_integration(0.0, 40)
    b.freeze_limiter(True)
    b.batch_time_integration(0.0, 20)
    forces = b.wall_forces_batch()
    for k in range(len(INST)):
        s = single(k, 60, freeze_after=40)
        assert np.array_equal(b.instance_state(k), s.cvar), k
        assert np.allclose(forces[k], s.wall_forces(), rtol=1e-12, atol=1e-15)
    # per-instance norms against the single solve's whole-mesh norms
    (n, l), (M, a) = INST[1]
    m, sch = cases.naca(n, l, mach=M, aoa=a, cfl=1.0)
    s = Solver(m, sch)
    s.initialize_solution(1, 0.0)
    ns = s.time_integration(0.0, 40)
    assert rel_err(nb[:, 1, :], ns) <= 1e-12


def test_masked_instance_is_frozen():
    b = batch()
    b.batch_time_integration(0.0, 10)
    before = b.cvar
    b.set_active([1, 0, 1])
    b.batch_time_integration(0.0, 15)
    after = b.cvar
    o = b.cell_offsets
    assert np.array_equal(after[o[1]:o[2]], before[o[1]:o[2]])
    for k in (0, 2):
        s = single(k, 25)
        assert np.array_equal(after[o[k]:o[k + 1]], s.cvar), k


@pytest.mark.parametrize("case", ["naca_ssp42", "vortex_periodic_ssp3"])
def test_graph_replay_bitwise(case):
    if case == "naca_ssp42":
        m, sch = cases.naca(64, 32, cfl=1.0)
    else:
        m, sch = cases.vortex(32, periodic=True, dt=0.01)
    out = []
    for graph in (False, True):
        s = Solver(m, sch)
        s.set_graph(graph)
        s.initialize_solution(1, 0.0)
        n = s.time_integration(0.0, 7)
        out.append((s.cvar, n))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_steady_solve_batch_converges():
    """Three coarse airfoils to a 1e-6 drop with the limiter frozen after 300
    steps; each converges and its iteration count matches its single solve."""
    b = batch()
    res = b.steady_solve_batch(20000, drop=1e-6, chunk=100, freeze_after=300)
    assert np.all(res["iterations"] > 0) and not np.any(res["diverged"])
    for k in range(len(INST)):
        (n, l), (M, a) = INST[k]
        m, sch = cases.naca(n, l, mach=M, aoa=a, cfl=1.0)
        s = Solver(m, sch)
        s.initialize_solution(1, 0.0)
        r = s.steady_solve(20000, drop=1e-6, chunk=100, freeze_after=300)
        assert r["iterations"] == res["iterations"][k], (k, r["iterations"], res["iterations"])
