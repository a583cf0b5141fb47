"""Oracle local energies of many walkers in a pool of CPU processes (test infrastructure only).

Each worker process runs the oracle (oracle/fire.py) with one torch thread on its share of the
walkers; the processes are started with ``spawn``, so none inherits a GPU context from the
caller.  Used by tests/ and the CPU-baseline / FP64-reference legs of bench.py.
"""
from __future__ import annotations

import os
from concurrent.futures import ProcessPoolExecutor
import multiprocessing as mp

import numpy as np

_STATE = {}


def _init(params, cfg, dtype_name):
    import torch

    from . import fire

    torch.set_num_threads(1)
    _STATE["p"] = fire.to_torch(params, getattr(torch, dtype_name))
    _STATE["cfg"] = cfg


def _work(r):
    from . import fire

    return [fire.local_energy(_STATE["p"], _STATE["cfg"], rw) for rw in r]


def default_procs() -> int:
    return max(1, min(16, len(os.sched_getaffinity(0))))


def local_energies(params, cfg, walkers, dtype: str = "float64", procs: int = 0):
    """(E_loc, E_kin, E_pot) [n] of walkers [n, N, 3] (float32 positions) with the oracle in
    ``dtype``, spread over ``procs`` processes (default: the CPU share, at most 16)."""
    r = np.asarray(walkers, np.float32)
    procs = procs or default_procs()
    procs = max(1, min(procs, len(r)))
    parts = [p for p in np.array_split(r, procs * 4) if len(p)]
    with ProcessPoolExecutor(max_workers=procs, mp_context=mp.get_context("spawn"), initializer=_init,
                             initargs=(params, cfg, dtype)) as ex:
        out = [x for part in ex.map(_work, parts) for x in part]
    arr = np.asarray(out, np.float64).reshape(len(r), 3)
    return arr[:, 0], arr[:, 1], arr[:, 2]
