"""The pins of tests/pin_checks.py can fail: the oracle is compiled with one
planted mistake at a time (the MUT_* switches of oracle/fvs_oracle.c, into a
temporary directory) and the check that covers it must raise.  The oracle
library the other tests use is built without any of them."""
import pytest

from oracle import oracle
from tests import pin_checks as P

CASES = [
    (1, "GGCB weights swapped", [P.check_ggcb_graded_exact, P.check_ggcb_brute]),
    (2, "GGNB node weights 1/d^2", [P.check_ggnb_brute]),
    (3, "GGNB self incidence counted twice", [P.check_ggnb_brute]),
    (4, "limiter bounds over face neighbours", [lambda lp: P.check_limiter_brute(lp, 3, 1),
                                                lambda lp: P.check_limiter_brute(lp, 1, 1)]),
    (5, "limiter extrapolation with d/dx, d/dy exchanged", [lambda lp: P.check_limiter_brute(lp, 3, 2)]),
    (6, "limiter of the density alone", [lambda lp: P.check_limiter_brute(lp, 3, 1)]),
    (7, "slip-wall ghost with zero normal velocity", [P.check_slip_wall_closed_form]),
    (8, "order-2 table with b, c exchanged", [lambda lp: P.check_rk_polynomial(lp, 2)]),
    (9, "order-1 table with a, d exchanged", [lambda lp: P.check_rk_polynomial(lp, 1)]),
    (10, "third stage advanced by dt/2", [lambda lp: P.check_rk_polynomial(lp, 3)]),
]


@pytest.fixture(scope="module")
def variants(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    return {k: oracle.build_variant(k, str(d)) for k, _, _ in CASES}


@pytest.mark.parametrize("mutant,what,checks", CASES, ids=[c[1] for c in CASES])
def test_pin_catches_planted_mistake(variants, mutant, what, checks):
    for chk in checks:
        with pytest.raises(AssertionError):
            chk(variants[mutant])
