"""The oracle's threaded evaluation (OpenMP over cells; face fluxes gathered
per cell in the serial scatter order) gives bitwise the serial results, for
every gradient, with and without a limiter, unsteady and steady."""
import numpy as np
import pytest

from fvs2d_amd import cases, meshgen
from oracle.oracle import Oracle
from tests.common import vortex_scheme

CASES = {
    "lsq": dict(),
    "lsq_venk": dict(limiter=1),
    "ggcb": dict(grad_method=1),
    "ggnb_venk": dict(grad_method=2, limiter=1),
    "rk4": dict(rk_scheme=0, rk_order=4),
    "ssp42_barth": dict(rk_scheme=1, rk_order=2, limiter=2),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("kind", ["periodic", "dirichlet"])
def test_threaded_oracle_is_bitwise_serial(kind, case):
    m = meshgen.jittered_square(24, periodic=(kind == "periodic"), seed=1)
    sch = vortex_scheme(dt=0.01, **CASES[case])
    out = []
    for n in (1, 3):
        o = Oracle(m, **sch.oracle_kwargs())
        o.set_threads(n)
        o.initialize(1, 0.0)
        norms = o.time_integration(0.0, 8)
        out.append((o.cvar, norms))
    assert np.all(np.isfinite(out[0][0]))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(np.asarray(out[0][1]), np.asarray(out[1][1]))


def test_threaded_oracle_steady_naca():
    m, sch = cases.naca(48, 24, cfl=1.0, limiter=1)
    res = []
    for n in (1, 4):
        o = Oracle(m, **sch.oracle_kwargs())
        o.set_threads(n)
        o.initialize(1, 0.0)
        o.time_integration(0.0, 15)
        res.append((o.cvar, o.wall_forces()))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
    with pytest.raises(ValueError):
        Oracle(m, **sch.oracle_kwargs()).set_threads(0)
