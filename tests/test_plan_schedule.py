"""The transform plans' scheduler (jgcm_lpt_schedule, host code, no GPU):
longest processing time first over persistent workgroups.  Checked against
what the rule fixes: every item dealt exactly once, each workgroup's items in
decreasing cost, the deterministic tie rules, Graham's bound (makespan at most
4/3 - 1/(3w) of the optimum, the optimum found by brute force on small
inputs), and the one-workgroup-per-item form."""
import itertools

import numpy as np
import pytest


@pytest.fixture(scope="module")
def lpt(jgcm):
    from jgcm_amd.plan import LPT_Schedule
    return LPT_Schedule


def loads(cost, order, starts):
    return np.array([cost[order[starts[w]:starts[w + 1]]].sum() for w in range(len(starts) - 1)])


def test_every_item_once_and_heaviest_first(lpt):
    rng = np.random.default_rng(0)
    for n, w in [(1, 1), (10, 3), (257, 16), (5000, 512), (40, 39)]:
        cost = rng.integers(1, 1000, n).astype(float)
        order, starts = lpt(cost, w)
        assert sorted(order.tolist()) == list(range(n))
        assert starts[0] == 0 and starts[-1] == n and len(starts) == min(w, n) + 1
        for q in range(len(starts) - 1):
            c = cost[order[starts[q]:starts[q + 1]]]
            assert np.all(np.diff(c) <= 0)
        # longest-first dealing: the last item placed on the most loaded
        # workgroup was no larger than the gap it closed, so loads differ by at
        # most the largest item
        L = loads(cost, order, starts)
        assert L.max() - L.min() <= cost.max()


def test_deterministic_ties(lpt):
    cost = np.array([5, 3, 8, 1, 1, 2], float)
    order, starts = lpt(cost, 2)
    # 8 -> wg0; 5 -> wg1; 3 -> wg1 (5 < 8); 2 -> wg0 (8 = 8, lower wg); 1, 1 -> wg1
    assert order.tolist() == [2, 5, 0, 1, 3, 4] and starts.tolist() == [0, 2, 6]


def test_one_workgroup_per_item(lpt):
    cost = np.array([1, 4, 4, 2, 9], float)
    for w in (0, 5, 9):
        order, starts = lpt(cost, w)
        assert order.tolist() == [4, 1, 2, 3, 0]       # decreasing, ties in index order
        assert starts.tolist() == list(range(6))


def test_graham_bound_against_brute_force(lpt):
    rng = np.random.default_rng(1)
    for trial in range(60):
        n, w = int(rng.integers(2, 9)), int(rng.integers(2, 4))
        cost = rng.integers(1, 20, n).astype(float)
        order, starts = lpt(cost, w)
        got = loads(cost, order, starts).max()
        best = min(max(np.bincount(np.array(a), weights=cost, minlength=w))
                   for a in itertools.product(range(w), repeat=n))
        assert best <= got <= (4 / 3 - 1 / (3 * w)) * best + 1e-9, (cost, w, got, best)
