"""The oracle against the reference CRM's own outputs (fixtures in tests/golden, written by
tests/golden/make_golden.py from the reference sources compiled with flang).  Bit-exact:
the oracle restates the reference's arithmetic operation by operation."""
import glob
import os

import numpy as np
import pytest

import oracle_lib
from parity import load_golden

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_reference_fixture(path):
    case, dt_gl, expected = load_golden(path)
    assert oracle_lib.oracle().run(case, dt_gl) == 0
    mismatched = []
    for key, want in expected.items():
        group, name = key.split(".", 1)
        got = getattr(case, group)[name]
        if not np.array_equal(got, want):
            mismatched.append((key, float(np.abs(got - want).max())))
    assert not mismatched, mismatched


def test_fixtures_present():
    names = {os.path.basename(p) for p in GOLDEN}
    assert {"single_32x28_gcmstep.npz", "convective_32x28_1sub.npz",
            "perturb_32x28_5sub.npz", "convective_64x58_4sub.npz"} <= names
