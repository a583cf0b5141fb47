"""CPU restatement of the local-energy statistics of a VMC step (oracle; test infrastructure only).

Follows clip_local_energies and local_energy_diff (src/sparse_wf/update.py:45-69) and the
energy mean / standard deviation of Trainer.step (update.py:166-168) on one device's batch
(the reference's pmean / pgather over devices are the identity on one GPU).  The median is
numpy's (midpoint of the two middle order statistics for an even count, as jnp.median), the
clip is min(max(x, lo), hi) as jnp.clip.  Pinned by hand-worked batches in
tests/test_energy_stats.py.
"""
from __future__ import annotations

import numpy as np


def clip_local_energies(e_loc, clip_local_energy: float, stat: str = "median"):
    """update.py:45-61.  Returns (clipped energies, centre, mean absolute deviation);
    centre and deviation are NaN when clip_local_energy <= 0 (nothing clipped)."""
    e = np.asarray(e_loc, np.float64)
    if clip_local_energy <= 0.0:
        return e.copy(), float("nan"), float("nan")
    if stat == "median":
        center = float(np.median(e))
    elif stat == "mean":
        center = float(np.mean(e))
    else:
        raise ValueError(f"Unknown statistic: {stat}")
    mad = float(np.mean(np.abs(e - center)))
    max_dev = clip_local_energy * mad
    lo, hi = center - max_dev, center + max_dev
    return np.minimum(np.maximum(e, lo), hi), center, mad


def local_energy_diff(e_loc, clip_local_energy: float, stat: str = "median"):
    """update.py:64-68: clipped energies minus their mean."""
    c, _, _ = clip_local_energies(e_loc, clip_local_energy, stat)
    return c - np.mean(c)


def energy_stats(e_loc, clip_local_energy: float, stat: str = "median") -> dict:
    """Everything Trainer.step derives from the batch's local energies (update.py:157-168)."""
    e = np.asarray(e_loc, np.float64)
    c, center, mad = clip_local_energies(e, clip_local_energy, stat)
    mean = float(np.mean(e))
    return {"E_mean": mean, "E_std": float(np.mean((e - mean) ** 2) ** 0.5), "center": center, "mad": mad,
            "E_clipped_mean": float(np.mean(c)), "clipped": c, "diff": c - np.mean(c)}
