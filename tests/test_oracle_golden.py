"""The oracle reproduces its committed golden vectors (tests/golden/oracle_vectors.npz,
written by tests/golden/make_golden.py): a regression pin on everything the
other oracle tests establish."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_golden as G  # noqa: E402

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_vectors.npz"))


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_oracle_reproduces_golden(name):
    got = G.compute(name)
    for k in ("U", "dU", "emap"):
        assert np.array_equal(got[k], GOLD[f"{name}/{k}"])
    for k in ("R", "JdU", "diag"):
        ref = GOLD[f"{name}/{k}"]
        assert np.abs(got[k] - ref).max() <= 1e-14 * np.abs(ref).max()
