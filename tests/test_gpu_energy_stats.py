"""fire_energy_stats against the oracle (update.py:45-69, 166-168): the median and every
clipped value bit-exact, sums within FP64 rounding (1e-12 relative)."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from oracle import update as U
from sparse_wf_amd import update as G

pytestmark = pytest.mark.gpu


def _batches():
    rng = np.random.default_rng(0)
    yield "one", np.array([-3.25])
    yield "two", np.array([1.0, -1.0])
    yield "three", np.array([5.0, -2.0, 7.0])
    yield "ties", np.repeat([-1.5, 2.0, 2.0, 7.0], 300)
    yield "gauss_1024", -600.0 + rng.normal(size=1024)
    yield "heavy_tail_4097", -128.0 + rng.standard_cauchy(size=4097)
    yield "signs_mixed", rng.normal(size=2048) * 1e-3
    yield "large_100k", -40.0 + rng.standard_t(3, size=100_000)
    yield "zeros_and_negzero", np.array([0.0, -0.0, 0.0, -0.0, 1.0])


@pytest.mark.parametrize("clip,stat", [(5.0, "median"), (1.0, "median"), (5.0, "mean"), (0.0, "median")])
def test_matches_oracle(gpu_device, clip, stat):
    for name, e in _batches():
        ref = U.energy_stats(e, clip, stat)
        got = G.energy_statistics(torch.as_tensor(e, device="cuda"), clip, stat)
        print(f"  {name} clip={clip} {stat}: centre {got['center']!r} vs {ref['center']!r}", flush=True)
        if clip > 0:
            if stat == "median":
                assert got["center"] == ref["center"], name
            else:
                assert abs(got["center"] - ref["center"]) <= 1e-12 * max(1.0, abs(ref["center"])), name
            assert abs(got["mad"] - ref["mad"]) <= 1e-12 * max(1e-300, abs(ref["mad"])), name
        else:
            assert math.isnan(got["center"]) and math.isnan(got["mad"])
        for k in ("E_mean", "E_clipped_mean"):
            assert abs(got[k] - ref[k]) <= 1e-12 * max(1.0, abs(ref[k])), (name, k)
        assert abs(got["E_std"] - ref["E_std"]) <= 1e-11 * max(1e-300, ref["E_std"]), name
        c = got["clipped"].cpu().numpy()
        # energies strictly inside the bounds pass unchanged (bit-exact); clipped ones take a bound,
        # which carries the rounding of the deviation sum
        if clip > 0:
            lo, hi = ref["center"] - clip * ref["mad"], ref["center"] + clip * ref["mad"]
            inside = (e > lo + 1e-9 * max(1.0, abs(lo))) & (e < hi - 1e-9 * max(1.0, abs(hi)))
        else:
            inside = np.ones_like(e, bool)
        np.testing.assert_array_equal(c[inside], e[inside])
        np.testing.assert_allclose(c, ref["clipped"], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(got["diff"].cpu().numpy(), ref["diff"], rtol=1e-9, atol=1e-9 * max(1.0, np.abs(e).max()))


def test_reference_names_and_nan(gpu_device):
    e = torch.tensor([0.0, 0.0, 0.0, 0.0, 10.0], device="cuda", dtype=torch.float64)
    np.testing.assert_array_equal(G.clip_local_energies(e, 1.0).cpu().numpy(), [0, 0, 0, 0, 2])
    np.testing.assert_allclose(G.local_energy_diff(e, 1.0).cpu().numpy(), [-0.4] * 4 + [1.6], atol=1e-12)
    s = G.energy_statistics(torch.tensor([1.0, float("nan"), 3.0], device="cuda"), 5.0)
    assert math.isnan(s["center"]) and math.isnan(s["E_mean"])
    with pytest.raises(ValueError):
        G.energy_statistics(torch.zeros(0, device="cuda"))
