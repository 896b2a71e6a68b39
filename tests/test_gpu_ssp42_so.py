"""The reference's SSP-RK(4,2) in two-register form (RK_SSP42_SO) on the GPU.

The reference evaluates the stages as u_k = u_0 + h c_k (R_k + sum of the
earlier R) (src/runge_kutta.f90:225, 384; rk_coef 1, c = 1/3, 1/3, 1/3, 1/4,
src/runge_kutta.f90:56-60); the two-register form u_k = u_{k-1} + h/3 R
(k = 1..3), u_4 = u_0/4 + 3/4 (u_3 + h/3 R) is the same method, equal in exact
arithmetic, with its own rounding.  So it is held to the oracle (which runs the
reference's form) at the trajectory bars of the reference's scheme, and its own
kernel arrangements (conservative stages, primitive stages, the whole-stage
kernel, 128-cell blocks, batches) are held to each other bit for bit.
"""
import numpy as np
import pytest

from fvs2d_amd import _lib, cases, meshgen
from fvs2d_amd.solver import BatchSolver, Solver
from oracle.oracle import Oracle
from tests.common import rel_err, vortex_scheme

pytestmark = pytest.mark.gpu

SO = _lib.RK_SSP42_SO


def test_unsteady_vortex_matches_oracle():
    m = meshgen.jittered_square(64, periodic=True, seed=0)
    sch = vortex_scheme(rk_scheme=SO, dt=0.005, reorder=1)
    orc = Oracle(m, **sch.oracle_kwargs())
    orc.initialize(1, 0.0)
    n_ref = orc.time_integration(0.0, 100)
    gpu = Solver(m, sch)
    gpu.initialize_solution(1, 0.0)
    n_gpu = gpu.time_integration(0.0, 100)
    err = rel_err(gpu.cvar, orc.cvar)
    print(f"vortex, 100 steps: two-register form vs the oracle's reference form {err:.2e}")
    assert err <= 1e-10
    assert rel_err(n_gpu, n_ref) <= 1e-8


@pytest.mark.parametrize("limiter", [0, 1])
def test_steady_naca_matches_oracle(limiter):
    """60 pseudo-steps with local time steps on a coarse transonic O-grid;
    bar as in the reference form's test: max(1e-10, 10 x the oracle's own
    one-ulp spread)."""
    m, sch = cases.naca(64, 32, cfl=1.0, limiter=limiter, rk_scheme=SO)
    orc = Oracle(m, **sch.oracle_kwargs())
    orc.initialize(1, 0.0)
    orc.time_integration(0.0, 60)
    pert = Oracle(m, **sch.oracle_kwargs())
    pert.initialize(1, 0.0)
    cv = pert.cvar
    pert.cvar = cv * (1.0 + 2.2e-16 * np.random.default_rng(0).choice([-1.0, 1.0], cv.shape))
    pert.time_integration(0.0, 60)
    tol = max(1e-10, 10.0 * rel_err(pert.cvar, orc.cvar))
    gpu = Solver(m, sch)
    gpu.initialize_solution(1, 0.0)
    gpu.time_integration(0.0, 60)
    err = rel_err(gpu.cvar, orc.cvar)
    print(f"limiter {limiter}: two-register form vs the oracle {err:.2e}, bar {tol:.2e}")
    assert err <= tol


@pytest.mark.parametrize("limiter", [0, 1])
def test_arrangements_bitwise(limiter):
    """Conservative stage states (kernel_path 0), primitive stage states
    (kernel_path 3), the whole-stage kernel (kernel_path 2) and 128-cell
    blocks give the same states bit for bit, with the limiter frozen half
    way as the steady drivers do."""
    m, sch = cases.naca(128, 64, cfl=1.25, limiter=limiter, rk_scheme=SO)
    out = []
    for path, choices in ((0, {"gradient": 2}), (3, None), (2, None), (0, {"gradient": 6, "fused": 4})):
        s = Solver(m, sch, kernel_path=path)
        if choices:
            s.set_launch_choices(choices)
        s.initialize_solution(1, 0.0)
        s.time_integration(0.0, 20)
        s.freeze_limiter(True)
        s.time_integration(0.0, 20)
        out.append(s.cvar)
        s.close()
    for k in range(1, len(out)):
        assert np.array_equal(out[0], out[k]), k


def test_batch_equals_single_with_masking():
    """Every instance of a ragged batch ends bit-identical to its own single
    solve; an instance masked for a while keeps its state and then resumes."""
    conds = [(96, 48, 0.75, 1.0), (64, 32, 0.6, -1.5), (128, 64, 0.8, 2.0)]
    pairs = [cases.naca(na, nl, mach=ma, aoa=al, rk_scheme=SO) for na, nl, ma, al in conds]
    b = BatchSolver([p[0] for p in pairs], [(c[2], c[3]) for c in conds], pairs[0][1])
    b.initialize_freestream()
    b.batch_time_integration(0.0, 15)
    before = b.cvar
    b.set_active(np.array([1, 0, 1], np.int32))
    b.batch_time_integration(0.0, 10)
    o = b.cell_offsets
    assert np.array_equal(b.cvar[o[1]:o[2]], before[o[1]:o[2]])
    b.set_active(np.ones(3, np.int32))
    b.batch_time_integration(0.0, 10)
    steps = (35, 25, 35)
    for k, ((m, sch), n) in enumerate(zip(pairs, steps)):
        s = Solver(m, sch)
        s.initialize_solution(1, 0.0)
        s.time_integration(0.0, n)
        assert np.array_equal(b.instance_state(k), s.cvar), k
