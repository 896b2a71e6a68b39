"""Host-side logic of the steady solvers (no device): the stall rule that
decides when an instance's limiter is frozen."""
from fvs2d_amd.solver import _stalled


def test_stall_rule():
    rule = (1000, 0.1)
    # less than a decade in 1000 steps -> stalled
    ends = [(250 * k, 1.0 * 0.9 ** k) for k in range(1, 6)]
    assert _stalled(ends, rule)
    # a decade per 500 steps -> not stalled
    ends = [(250 * k, 10.0 ** (-k / 2)) for k in range(1, 6)]
    assert not _stalled(ends, rule)
    # window not yet covered -> never stalled
    assert not _stalled([(250, 1.0), (500, 1.0), (750, 1.0)], rule)
    # the comparison point is the latest end at least window steps back
    ends = [(250, 1.0), (500, 0.5), (1250, 0.06)]
    assert not _stalled(ends, rule)          # 0.06 <= 0.1 * 1.0 (step 250)
    ends = [(250, 1.0), (500, 0.5), (1500, 0.06)]
    assert _stalled(ends, rule)              # 0.06 > 0.1 * 0.5 (step 500)


def test_stall_window_scales_with_grid():
    from fvs2d_amd.cases import stall_rule_for
    d64, d512 = dict(n_around=64), dict(n_around=512)
    assert stall_rule_for(d64, (2000, 0.1), True) == (2000, 0.1)
    assert stall_rule_for(d512, (2000, 0.1), True) == (16000, 0.1)
    assert stall_rule_for(d512, (2000, 0.1), False) == (2000, 0.1)
    assert stall_rule_for(d512, None, True) is None
