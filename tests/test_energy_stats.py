"""Oracle of the local-energy statistics (update.py:45-69, 166-168) on hand-worked batches.

Each case is worked out by hand from the reference's definitions; a centre from the wrong
statistic, the standard deviation in place of the mean absolute deviation, a lower median in
place of the midpoint, or the sample standard deviation changes at least one of them."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import update as U


def test_outlier_within_bounds():
    # median 3; deviations 2, 1, 0, 1, 97 -> mean 20.2; bounds 3 -+ 101: nothing clipped
    s = U.energy_stats([1.0, 2.0, 3.0, 4.0, 100.0], 5.0, "median")
    assert s["center"] == 3.0 and abs(s["mad"] - 20.2) < 1e-12
    np.testing.assert_array_equal(s["clipped"], [1, 2, 3, 4, 100])
    np.testing.assert_allclose(s["diff"], np.array([1, 2, 3, 4, 100]) - 22.0, atol=1e-12)


def test_outlier_clipped():
    # median 0, mean abs deviation 2, clip 1: bounds [-2, 2]; clipped mean 0.4
    s = U.energy_stats([0.0, 0.0, 0.0, 0.0, 10.0], 1.0, "median")
    np.testing.assert_array_equal(s["clipped"], [0, 0, 0, 0, 2])
    np.testing.assert_allclose(s["diff"], [-0.4, -0.4, -0.4, -0.4, 1.6], atol=1e-12)
    assert s["E_mean"] == 2.0 and abs(s["E_std"] - 4.0) < 1e-12 and abs(s["E_clipped_mean"] - 0.4) < 1e-12


def test_even_count_midpoint_median():
    # sorted 1, 2, 3, 4: median 2.5; deviations 1.5, 1.5, .5, .5 -> 1; bounds [1.5, 3.5]
    c, center, mad = U.clip_local_energies([4.0, 1.0, 3.0, 2.0], 1.0, "median")
    assert center == 2.5 and mad == 1.0
    np.testing.assert_array_equal(c, [3.5, 1.5, 3.0, 2.0])


def test_mean_statistic():
    # mean 3; deviations 3, 3, 3, 9 -> 4.5; bounds [-1.5, 7.5]
    c, center, mad = U.clip_local_energies([0.0, 0.0, 0.0, 12.0], 1.0, "mean")
    assert center == 3.0 and mad == 4.5
    np.testing.assert_array_equal(c, [0, 0, 0, 7.5])


def test_no_clipping_and_population_std():
    s = U.energy_stats([1.0, 3.0], 0.0)
    np.testing.assert_array_equal(s["clipped"], [1.0, 3.0])
    assert math.isnan(s["center"]) and math.isnan(s["mad"])
    assert s["E_std"] == 1.0  # population standard deviation, not sqrt(2)
    with pytest.raises(ValueError):
        U.clip_local_energies([1.0], 1.0, "mode")
