"""Oracle local energies of many walkers in a pool of CPU processes (test infrastructure only).

Each worker process runs the oracle (oracle/fire.py) with one torch thread on its share of the
walkers; the processes are started with ``spawn``, so none inherits a GPU context from the
caller.  Used by tests/ and the CPU-baseline / FP64-reference legs of bench.py.
"""
from __future__ import annotations

import multiprocessing as mp
import os
from concurrent.futures import ProcessPoolExecutor

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


class OraclePool:
    """``procs`` worker processes holding the oracle parameters in ``dtype``; reusable, so a
    timed call does not include process start-up (``with OraclePool(...) as p: p(walkers)``)."""

    def __init__(self, params, cfg, dtype: str = "float64", procs: int = 0):
        self.procs = procs or default_procs()
        self._ex = ProcessPoolExecutor(max_workers=self.procs, mp_context=mp.get_context("spawn"), initializer=_init,
                                       initargs=(params, cfg, dtype))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self._ex.shutdown()

    def warm(self):
        """Start every worker (imports, parameter conversion) before a timed call."""
        list(self._ex.map(_work, [np.zeros((0, 1, 3), np.float32)] * (4 * self.procs)))

    def __call__(self, walkers):
        """(E_loc, E_kin, E_pot) [n] of walkers [n, N, 3] (float32 positions)."""
        r = np.asarray(walkers, np.float32)
        parts = [p for p in np.array_split(r, self.procs * 4) if len(p)]
        out = [x for part in self._ex.map(_work, parts) for x in part]
        arr = np.asarray(out, np.float64).reshape(len(r), 3)
        return arr[:, 0], arr[:, 1], arr[:, 2]


def local_energies(params, cfg, walkers, dtype: str = "float64", procs: int = 0):
    """(E_loc, E_kin, E_pot) [n] of walkers [n, N, 3] with the oracle in ``dtype``, spread over
    ``procs`` processes (default: the CPU share, at most 16)."""
    r = np.asarray(walkers, np.float32)
    procs = max(1, min(procs or default_procs(), len(r)))
    with OraclePool(params, cfg, dtype, procs) as p:
        return p(r)
