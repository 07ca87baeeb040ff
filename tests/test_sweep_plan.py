"""Host logic of the grouped sweep (no GPU): contiguous groups within a budget."""
import numpy as np

from pyalbsim_amd.sweep import plan_groups


def test_groups_are_contiguous_and_within_budget():
    rs = np.random.default_rng(0)
    b = rs.uniform(0.1, 3.0, 200)
    for budget in (0.5, 2.0, 7.5, 1e9):
        g = plan_groups(b, budget)
        assert g[0][0] == 0 and g[-1][1] == len(b)
        assert all(x[1] == y[0] for x, y in zip(g, g[1:]))
        for lo, hi in g:
            assert hi > lo
            assert hi - lo == 1 or b[lo:hi].sum() <= budget
            if hi < len(b):  # the next scene would not have fitted
                assert b[lo:hi + 1].sum() > budget
    assert plan_groups(b, 1e9) == [(0, len(b))]
    assert plan_groups([], 1.0) == []
