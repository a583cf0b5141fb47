"""HIP path against the committed golden fixtures (FP64 oracle outputs, tests/golden/),
uploading the fixture's packed parameters: nothing is recomputed on the CPU.  Tolerances as
in test_gpu_parity.py (FP32 and mixed precision)."""
from __future__ import annotations

import os

import numpy as np
import pytest
import torch

from sparse_wf_amd.wave_function import FireWaveFunction, PackedParams

from cases import small_cfg

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", params=["small", "small_restricted"])
def golden(request, gpu_device):
    d = dict(np.load(os.path.join(GOLDEN, f"{request.param}.npz")))
    cfg = small_cfg(request.param == "small_restricted")
    wf = FireWaveFunction(cfg, device=gpu_device)
    prm = PackedParams(d["packed"], d["sizes"])
    wf.upload(prm)
    return cfg, wf, prm, d


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_neighbours(golden):
    cfg, wf, prm, d = golden
    counts, nbr, _ = wf.neighbours(d["r"])
    np.testing.assert_array_equal(counts.cpu().numpy(), d["counts"])
    np.testing.assert_array_equal(nbr.cpu().numpy(), d["nbr"])


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_values_and_energies(golden, precision):
    cfg, wf, prm, d = golden
    _, ho = wf.embedding(prm, d["r"], precision=precision)
    assert _rel(ho.cpu().numpy(), d["h_out"]) < 2e-5
    out = wf.local_energy_terms(prm, d["r"], precision=precision, with_derivatives=True)
    lp = out["logpsi"].cpu().numpy()
    assert np.all(np.abs(lp - d["logpsi"]) < 1e-4 * np.maximum(1.0, np.abs(d["logpsi"])))
    np.testing.assert_array_equal(out["sign"].cpu().numpy(), d["sign"])
    np.testing.assert_allclose(out["e_pot"].cpu().numpy(), d["e_pot"], rtol=1e-9)
    tol = 1e-3 if precision == "fp32" else 2e-3
    ek = out["e_kin"].cpu().numpy()
    assert np.all(np.abs(ek - d["e_kin"]) < tol * np.maximum(1.0, np.abs(d["e_kin"])))
    assert _rel(out["grad"].cpu().numpy(), d["grad"]) < (1e-4 if precision == "fp32" else 5e-4)


def test_low_rank_move(golden):
    cfg, wf, prm, d = golden
    (_, _), st = wf.log_psi_with_state(prm, d["r"], precision="fp32")
    r_new = d["r"].copy()
    r_new[:, 3] += np.array([2.0, 0.0, 0.0], np.float32)
    (_, lp), _ = wf.log_psi_low_rank_update(prm, r_new, np.array([3]), state=st)
    ref = d["logpsi_after_move"]
    assert np.all(np.abs(lp.cpu().numpy() - ref) < 1e-4 * np.maximum(1.0, np.abs(ref)))
