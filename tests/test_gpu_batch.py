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
from fvs2d_amd._lib import RK_SSP3, RK_SSP42
from fvs2d_amd.solver import BatchSolver, Solver
from tests.common import rel_err, vortex_scheme

pytestmark = pytest.mark.gpu

INST = [((48, 24), (0.75, 1.0)), ((64, 32), (0.8, 1.25)), ((32, 16), (0.6, -2.0))]
# instance sizes 1600, 2704, 1296 cells: boundaries inside 256-cell blocks
INST_RAGGED = [((40, 20), (0.75, 1.0)), ((52, 26), (0.8, 1.25)), ((36, 18), (0.6, -2.0))]


def single(k, steps, freeze_after=None, inst=INST, rk=RK_SSP42, cfl=1.0):
    (n, l), (M, a) = inst[k]
    m, sch = cases.naca(n, l, mach=M, aoa=a, cfl=cfl, rk_scheme=rk)
    s = Solver(m, sch)
    s.initialize_solution(1, 0.0)
    if freeze_after is not None:
        s.time_integration(0.0, freeze_after)
        s.freeze_limiter(True)
        s.time_integration(0.0, steps - freeze_after)
    else:
        s.time_integration(0.0, steps)
    return s


def batch(inst=INST, kernel_path=0, rk=RK_SSP42, cfl=1.0):
    meshes = [cases.naca(n, l, cfl=cfl)[0] for (n, l), _ in inst]
    _, sch = cases.naca(16, 8, cfl=cfl, kernel_path=kernel_path, rk_scheme=rk)
    b = BatchSolver(meshes, [ma for _, ma in inst], sch)
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


@pytest.mark.parametrize("kernel_path", [0, 2])
@pytest.mark.parametrize("rk", [RK_SSP42, RK_SSP3], ids=["ssp42", "ssp3"])
def test_masked_instance_in_shared_blocks(rk, kernel_path):
    """Instance boundaries inside 256-cell launch blocks: a block that holds
    cells of a masked and of an active instance is still launched, and the
    masked cells keep their state bit for bit (SSP-RK3 runs its stages on
    conservative states, SSP-RK(4,2) on primitive ones; kernel_path 2 is the
    whole-stage kernel).  SSP-RK3 runs at CFL 0.3: the local time step is
    four times the common convention and SSP-RK3 diverges at 1.0."""
    cfl = 0.3 if rk == RK_SSP3 else 1.0
    b = batch(INST_RAGGED, kernel_path=kernel_path, rk=rk, cfl=cfl)
    b.batch_time_integration(0.0, 10)
    before = b.cvar
    assert np.all(np.isfinite(before))
    b.set_active([1, 0, 1])
    b.batch_time_integration(0.0, 15)
    after = b.cvar
    o = b.cell_offsets
    assert np.array_equal(after[o[1]:o[2]], before[o[1]:o[2]])
    for k in (0, 2):
        s = single(k, 25, inst=INST_RAGGED, rk=rk, cfl=cfl)
        assert np.array_equal(after[o[k]:o[k + 1]], s.cvar), k
    # unmasked again, the instance resumes from where it stopped
    b.set_active([1, 1, 1])
    b.batch_time_integration(0.0, 5)
    s = single(1, 15, inst=INST_RAGGED, rk=rk, cfl=cfl)
    assert np.array_equal(b.instance_state(1), s.cvar)


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


@pytest.mark.parametrize("kernel_path", [0, 2])
def test_per_instance_limiter_freeze(kernel_path):
    """Freezing the limiter of one instance (fvs_batch_freeze_limiter) leaves
    the others re-evaluating theirs: instance 1 frozen after 30 steps, 0 and 2
    never, each bit-identical to its own single solve.  Instance 1's cells
    share launch blocks with instances 0 and 2 (the block boundaries do not
    fall on instance boundaries), so the per-cell instance lookup is used."""
    b = batch(INST_RAGGED, kernel_path)
    o = b.cell_offsets
    assert o[1] % 256 != 0 and o[2] % 256 != 0
    b.batch_time_integration(0.0, 30)
    b.freeze_limiter_batch([0, 1, 0])
    b.batch_time_integration(0.0, 25)
    b.freeze_limiter(False)
    for k in range(len(INST_RAGGED)):
        s = single(k, 55, freeze_after=30 if k == 1 else None, inst=INST_RAGGED)
        assert np.array_equal(b.instance_state(k), s.cvar), k


def test_steady_batch_restart_on_divergence():
    """A diverging instance (forced by a threshold below its own residual)
    is reset through on_diverge and restarted once, then masked out; the
    others converge as before."""
    b = batch()
    b.initialize_freestream()
    r_ref = b.reference_norms()
    calls = []

    def reset(k):
        calls.append(k)
        cv = b.cvar
        cv[b.cell_offsets[k]:b.cell_offsets[k + 1]] = b.freestream_state(k)
        b.cvar = cv
        return True

    res = b.steady_solve_batch(3000, drop=1e-3, chunk=50, freeze_after=300, r_ref=r_ref, on_diverge=reset)
    assert not np.any(res["diverged"]) and np.all(res["iterations"] > 0) and calls == []
    # divergence at 1e3 reference norms; instance 2's reference is scaled by
    # 1e-6, so its first-step norm trips the test at once, it restarts from
    # the free stream and trips again
    b.initialize_freestream()
    res2 = b.steady_solve_batch(3000, drop=1e-3, chunk=50, freeze_after=300, r_ref=r_ref * np.array([1, 1, 1e-6]),
                                diverge=1e3, on_diverge=reset)
    assert calls == [2]
    assert res2["restarted"][2] and res2["diverged"][2] and res2["iterations"][2] == -1
    assert np.array_equal(res2["iterations"][:2], res["iterations"][:2])


def test_queued_batch_iterations_match_all_active():
    """Continuous batching (at most two of four instances stepping, the next
    starting as one converges): without the limiter every instance needs the
    same number of iterations as when all four step from the start, since
    masked instances are left untouched and each instance's count runs from
    its own start."""
    inst = [((32, 16), (0.6, -2.0)), ((40, 20), (0.5, 1.0)), ((36, 18), (0.55, 2.0)), ((32, 16), (0.65, 0.5))]

    def run(max_active):
        meshes = [cases.naca(n, l, cfl=1.0)[0] for (n, l), _ in inst]
        _, sch = cases.naca(16, 8, cfl=1.0, limiter=0)
        b = BatchSolver(meshes, [ma for _, ma in inst], sch)
        b.initialize_freestream()
        r_ref = b.reference_norms()
        res = b.steady_solve_batch(20000, drop=1e-6, chunk=100, r_ref=r_ref, max_active=max_active)
        return res["iterations"]

    all_on, queued = run(None), run(2)
    assert np.all(all_on > 0)
    assert np.array_equal(all_on, queued)


def test_batch_driver_queue_and_give_up():
    """The batch driver (python -m fvs2d_amd.cases batch) with the give-up
    rule and at most two instances stepping at once gives every instance the
    outcome and iteration count of the run with all instances stepping."""
    kw = dict(ninst=6, seed=1, max_steps=30000, drop=1e-6, cfl=1.25, freeze_after=30000,
              resolutions=((32, 16), (64, 32)), freeze_stall=(2000, 0.1), stall_scale=True, give_up=(8000, 0.5))
    a = cases.run_batch(**kw)
    q = cases.run_batch(**kw, max_active=2)
    assert a["instances"] == q["instances"] == 6 - len(a["rejected"])
    assert a["converged"] >= 4
    assert [p["iterations"] for p in a["per_instance"]] == [p["iterations"] for p in q["per_instance"]]
    assert a["gave_up"] == q["gave_up"]
