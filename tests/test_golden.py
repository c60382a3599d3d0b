"""The committed fixtures (tests/golden/*.npz) are re-derived from the oracle
on the same seeded inputs: a change of the oracle's behaviour shows here."""
import os
import sys

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)

import make_golden  # noqa: E402


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", sorted(make_golden.BUILDERS))
def test_oracle_reproduces_fixture(name):
    stored = np.load(os.path.join(GOLDEN, name))
    fresh = make_golden.BUILDERS[name]()
    assert sorted(stored.files) == sorted(fresh)
    for k in stored.files:
        a, b = fresh[k], stored[k]
        assert a.shape == b.shape and a.dtype == b.dtype, k
        assert rel(a, b) <= 1e-12, (k, rel(a, b))


def test_fixture_inputs_are_the_seeded_inputs():
    from inputs import T21, seeded_grid, seeded_spec
    z = np.load(os.path.join(GOLDEN, "golden_t21.npz"))
    assert np.array_equal(z["grid_in"], seeded_grid(T21["nlev"], T21["nlat"], T21["seed"]))
    assert np.array_equal(z["spec_in"], seeded_spec(T21["N"], T21["nlev"], T21["seed"] + 1))
