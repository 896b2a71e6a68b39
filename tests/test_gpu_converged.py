"""Steady solves run to a 1e-8 density-residual drop on the GPU and on the
oracle, with the same stopping rule (norms read every 250 steps, the solve
stops at the first step whose density norm is at most 1e-8 of the first
step's; src/runge_kutta.f90:396-398 computes the norm).

Without a limiter the scheme has one discrete steady state, which both runs
approach; rounding differences along the way decay with the residual, so the
converged states, C_L and C_D agree far more closely than trajectories of the
same length do.  Subsonic (M 0.5, 2 deg) flow around a 12% 4-digit O-grid of
64 x 32 quads, LSQ-nn, SSP-RK(4,2) with local time steps at CFL 1.25.
"""
import os

import numpy as np
import pytest

from fvs2d_amd import cases
from fvs2d_amd.solver import Solver
from oracle.oracle import Oracle
from tests.common import rel_err

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
CHUNK = 250


def oracle_steady(orc, max_steps, drop):
    """The stopping rule of Solver.steady_solve on the oracle."""
    r1, done = None, 0
    while done < max_steps:
        n = orc.time_integration(0.0, CHUNK)
        if r1 is None:
            r1 = n[0, 0]
        hit = np.nonzero(n[:, 0] <= drop * r1)[0]
        done += CHUNK
        if hit.size:
            return done - CHUNK + int(hit[0]) + 1, done
    return None, done


def test_converged_steady_state_matches_oracle():
    m, sch = cases.naca(64, 32, mach=0.5, aoa=2.0, thickness=0.12, cfl=1.25, limiter=0)
    gpu = Solver(m, sch)
    gpu.initialize_solution(1, 0.0)
    r = gpu.steady_solve(60000, drop=1e-8, chunk=CHUNK)
    orc = Oracle(m, **sch.oracle_kwargs())
    orc.set_threads(THREADS)
    orc.initialize(1, 0.0)
    it, steps = oracle_steady(orc, 60000, 1e-8)
    print(f"iterations to a 1e-8 drop: GPU {r['iterations']}, oracle {it}")
    assert r["iterations"] is not None and it is not None
    assert abs(r["iterations"] - it) <= 2
    assert r["steps_run"] == steps
    err = rel_err(gpu.cvar, orc.cvar)
    orc.cvar2pvar()
    f_gpu, f_ref = gpu.wall_forces(), orc.wall_forces()
    print(f"converged state: GPU vs oracle {err:.2e}; C_L {f_gpu[0]:.10f} / {f_ref[0]:.10f}, "
          f"C_D {f_gpu[1]:.10f} / {f_ref[1]:.10f}")
    assert err <= 1e-9
    assert np.allclose(f_gpu, f_ref, rtol=1e-9, atol=1e-12)
