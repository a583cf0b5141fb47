"""The committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the FP64 oracle) against the oracle and the parameter packing of this checkout: a
change of the oracle's arithmetic, of the reference-layout initialisers or of the canonical
packing order shows up here."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402


@pytest.mark.parametrize("name", make_golden.CASES)
def test_golden_fixture_reproduces(name):
    stored = dict(np.load(make_golden.path(name)))
    fresh = make_golden.compute(name)
    assert set(stored) == set(fresh)
    for k in ("packed", "sizes", "r", "counts", "nbr", "sign"):
        np.testing.assert_array_equal(stored[k], fresh[k], err_msg=k)
    for k in ("h_out", "logpsi", "e_loc", "e_kin", "e_pot", "grad", "lap", "logpsi_after_move"):
        np.testing.assert_allclose(stored[k], fresh[k], rtol=1e-10, atol=1e-10, err_msg=k)
