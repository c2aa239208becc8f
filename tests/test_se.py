"""Standard errors on the host: the oracle's restatement of the reference's
getSE = 1 computation (per-site central differences, Hessian of Seo et al.
2004, HessianSKT2004 treesub.c:6979-7048; codeml.c:754-785) against what the
reference prints for the same data at the same parameters (golden fixtures
from tests/golden/make_se_golden.py)."""
import gzip
import json
import os

import numpy as np
import pytest

from oracle import oracle
from . import fixtures

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_se(name):
    with gzip.open(os.path.join(GOLDEN, f"se_{name}.json.gz"), "rt") as f:
        return json.load(f)


@pytest.mark.parametrize("hyp", ["H1", "H0"])
def test_oracle_standard_errors_match_reference(hyp):
    fx, tree, data = fixtures.load("toy_amb")
    g = load_se("toy_amb")[hyp]
    se = oracle.standard_errors(tree, data, g["x"], hyp == "H0")
    ref = np.array(g["se"])
    # printed at %8.6f; both sides difference per-site values at the same
    # step, whose last digits differ between the two eigen solvers
    assert np.all(np.abs(se - ref) <= 2e-4 * ref + 1e-6)
