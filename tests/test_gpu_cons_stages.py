"""SSP-RK3 on the default path runs its stages on conservative states: the
gradient and flux kernels read the stage's conservative state and convert it
to primitive variables as they stage it, and no primitive state is written
between stages (U0 -> Ub -> Uc -> U0), with the staged-halo LSQ gradient and
with unlimited Green-Gauss gradients.  kernel_path 3 keeps the primitive-state
arrangement, so the two must agree bitwise: each face
flux, gradient and cell sum is the same arithmetic on the same bits, and the
conversion is compiled without multiply-add contraction wherever it runs.

After a time_integration call the library's primitive state must be the
converted final state (it is refreshed once at the end of the call): checked
against numpy's evaluation of cvar2pvar (src/data_solution.f90:72-86), which
has no contraction either, and through a gradient call that reads it.
"""
import numpy as np
import pytest

from fvs2d_amd import meshgen
from fvs2d_amd.solver import Solver
from oracle.oracle import Oracle
from tests.common import rel_err, vortex_scheme

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def square():
    return meshgen.jittered_square(40, periodic=True, seed=3)


def run(m, sch, steps, graph=False, **path):
    s = Solver(m, sch, reorder=1, **path)
    s.initialize_solution(1, 0.0)
    if graph:
        s.set_graph(True)
    s.time_integration(0.0, steps)
    return s


@pytest.mark.parametrize("limiter", [0, 1, 2])
def test_conservative_stages_match_primitive_path(square, limiter):
    sch = vortex_scheme(dt=0.01, limiter=limiter)
    a = run(square, sch, 25)
    b = run(square, sch, 25, kernel_path=3)
    ua, ub = a.cvar, b.cvar
    a.close()
    b.close()
    assert np.array_equal(ua, ub)
    orc = Oracle(square, **sch.oracle_kwargs())
    orc.initialize(1, 0.0)
    orc.time_integration(0.0, 25)
    # limited schemes amplify rounding (the oracle's own 1-ulp spread), so
    # the bar there is looser than the unlimited 1e-10
    assert rel_err(ua, orc.cvar) <= (1e-10 if limiter == 0 else 1e-8)


@pytest.mark.parametrize("gm", [1, 2], ids=["ggcb", "ggnb"])
def test_green_gauss_conservative_stages(square, gm):
    """Unlimited Green-Gauss gradients take the conservative states too
    (converted in their gathers); kernel_path 3 keeps primitive states."""
    sch = vortex_scheme(dt=0.01, grad_method=gm)
    a = run(square, sch, 25)
    b = run(square, sch, 25, kernel_path=3)
    ua, ub = a.cvar, b.cvar
    a.close()
    b.close()
    assert np.array_equal(ua, ub)
    orc = Oracle(square, **sch.oracle_kwargs())
    orc.initialize(1, 0.0)
    orc.time_integration(0.0, 25)
    assert rel_err(ua, orc.cvar) <= 1e-10


def test_graph_replay_of_conservative_stages(square):
    """Graph replay covers two SSP-RK3 steps per graph; 7 steps = 3 replays
    + 1 eager step, bitwise the eager run."""
    sch = vortex_scheme(dt=0.01)
    a = run(square, sch, 7)
    b = run(square, sch, 7, graph=True)
    assert np.array_equal(a.cvar, b.cvar)
    a.close()
    b.close()


def test_primitive_state_after_integration(square):
    sch = vortex_scheme(dt=0.01)
    s = run(square, sch, 9)
    u = s.cvar
    g_after = s.compute_gradient_cellcntr()  # reads the library's primitive state as left by the call
    w = s.cvar2pvar()
    s.close()
    # numpy: the same operations in the same order, no contraction
    ref = np.empty_like(u)
    ref[:, 0] = u[:, 0]
    ref[:, 1] = u[:, 1] / u[:, 0]
    ref[:, 2] = u[:, 2] / u[:, 0]
    ref[:, 3] = (1.4 - 1.0) * (u[:, 3] - 0.5 * ref[:, 0] * (ref[:, 1] * ref[:, 1] + ref[:, 2] * ref[:, 2]))
    assert np.array_equal(w, ref)
    fresh = Solver(square, sch, reorder=1)
    fresh.cvar = u
    fresh.cvar2pvar()
    g_fresh = fresh.compute_gradient_cellcntr()
    fresh.close()
    assert np.array_equal(g_after, g_fresh)


@pytest.mark.parametrize("gm", [1, 2], ids=["ggcb", "ggnb"])
@pytest.mark.parametrize("limiter", [1, 2], ids=["venk", "barth"])
def test_green_gauss_limiter_halo_bitwise(square, gm, limiter):
    """A limited Green-Gauss gradient takes its limiter bounds from the
    block's staged halo on the default path and from the stencil walk in HBM
    on kernel_path 3; min and max are exact, so the states agree bitwise."""
    sch = vortex_scheme(dt=0.01, grad_method=gm, limiter=limiter)
    a = run(square, sch, 20)
    b = run(square, sch, 20, kernel_path=3)
    ua, ub = a.cvar, b.cvar
    a.close()
    b.close()
    assert np.all(np.isfinite(ua))
    assert np.array_equal(ua, ub)
