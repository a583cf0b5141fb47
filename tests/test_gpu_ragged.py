"""Ragged batch: geometries of different sizes sampled concurrently from captured sweep
graphs give the same walkers as each geometry sampled eagerly on its own."""
from __future__ import annotations

import pytest
import torch

from sparse_wf_amd.ragged import random_geometries, run_ragged

pytestmark = pytest.mark.gpu


def test_ragged_graphs_match_eager(gpu_device):
    geoms = random_geometries(4, seed=3, n_el=(10, 40), walkers=32, feature_dim=64)
    print("\n[ragged] N:", [g.cfg.n_el for g in geoms], flush=True)
    a = run_ragged(geoms, sweeps=2, use_graphs=True)
    b = run_ragged(geoms, sweeps=2, use_graphs=False)
    for ra, rb in zip(a["walkers"], b["walkers"]):
        assert torch.equal(ra, rb)
    assert all(0.0 < p < 1.0 for p in a["pmove"])
    assert a["pmove"] == b["pmove"]
    print(f"  graphs {a['walker_moves_per_s']:.3g} moves/s, eager {b['walker_moves_per_s']:.3g} moves/s", flush=True)
