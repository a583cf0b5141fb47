"""Local-energy statistics of a VMC step on the GPU (one kernel, ``fire_energy_stats``).

Mirrors clip_local_energies / local_energy_diff (src/sparse_wf/update.py:45-69) and the
energy mean and standard deviation that Trainer.step reports (update.py:166-168) for the batch
of one GPU; ``stat`` is ``"median"`` (the reference default, config/default.yaml:137-139) or
``"mean"``.  Inputs are device tensors (any float dtype, computed in FP64).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib

_STATS = {"median": 0, "mean": 1}


def _run(e_loc, clip_local_energy: float, stat: str, want_clipped=True, want_diff=True):
    if stat not in _STATS:
        raise ValueError(f"Unknown statistic: {stat}")
    e = torch.as_tensor(e_loc)
    if e.device.type != "cuda":
        raise ValueError("fire_energy_stats runs on the GPU: pass a device tensor")
    e = e.reshape(-1).to(torch.float64).contiguous()
    if e.numel() == 0:
        raise ValueError("empty batch")
    lib = _lib.load()
    clipped = torch.empty_like(e) if want_clipped else None
    diff = torch.empty_like(e) if want_diff else None
    stats = torch.empty(5, dtype=torch.float64, device=e.device)
    with torch.cuda.device(e.device):
        stream = ctypes.c_void_p(torch.cuda.current_stream(e.device).cuda_stream)
        _lib.check(lib.fire_energy_stats(None, stream, ctypes.c_void_p(e.data_ptr()), e.numel(),
                                         float(clip_local_energy), _STATS[stat],
                                         ctypes.c_void_p(clipped.data_ptr() if clipped is not None else 0),
                                         ctypes.c_void_p(diff.data_ptr() if diff is not None else 0),
                                         ctypes.c_void_p(stats.data_ptr())))
    return clipped, diff, stats


def clip_local_energies(e_loc, clip_local_energy: float, stat: str = "median") -> torch.Tensor:
    """update.py:45-61 on one GPU's batch."""
    return _run(e_loc, clip_local_energy, stat, want_diff=False)[0]


def local_energy_diff(e_loc, clip_local_energy: float, stat: str = "median") -> torch.Tensor:
    """update.py:64-68: clipped energies minus their mean (dloss/dlogpsi of the energy loss)."""
    return _run(e_loc, clip_local_energy, stat, want_clipped=False)[1]


def energy_statistics(e_loc, clip_local_energy: float = 5.0, stat: str = "median") -> dict:
    """All of the above in one launch: E_mean, E_std (the unclipped batch, as the
    ``opt/E`` and ``opt/E_std`` of update.py:166-168), the clip centre and mean absolute
    deviation, the clipped mean, and the clipped energies and their deviations."""
    clipped, diff, stats = _run(e_loc, clip_local_energy, stat)
    s = stats.cpu().tolist()
    return {"E_mean": s[0], "E_std": s[1], "center": s[2], "mad": s[3], "E_clipped_mean": s[4],
            "clipped": clipped, "diff": diff}
