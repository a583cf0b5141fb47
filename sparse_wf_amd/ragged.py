"""Ragged batches: many geometries with different electron counts, neighbour structures and
parameters sampled together on one GPU.

Each geometry keeps its own handle (parameters, resident walker state, FP64 inverses) and its
own HIP stream.  One host thread captures every geometry's Metropolis sweep into a hipGraph
(``fire_mcmc_steps`` captures on first use) and then issues all replays without waiting, so
the sweeps of all geometries run concurrently on the device; the host synchronises once at
the end.  The per-step work inside a sweep is already ragged: affected electrons are
compacted into CSR rows whose count stays on the device.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import List

import numpy as np
import torch

from . import _lib
from .config import FireConfig
from .wave_function import FireWaveFunction


@dataclass
class Geometry:
    cfg: FireConfig
    params: dict
    walkers: np.ndarray  # [B, N, 3]
    seed: int


def random_geometries(n: int, seed: int, n_el=(10, 200), walkers: int = 256, feature_dim: int = 256,
                      elements=("H", "C", "N", "O")) -> List[Geometry]:
    """Seeded random H/C/N/O chains (atoms 1.8-2.9 bohr apart) with electron counts uniform in
    ``n_el``, random-init parameters (scales set on the first walker) and ``walkers`` walkers."""
    from . import systems
    from .params import exercise_jastrow

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        N = int(rng.integers(n_el[0], n_el[1] + 1))
        mol = systems.random_chain(rng, N, elements=elements)
        cfg = FireConfig(mol=mol, cutoff=5.0, feature_dim=feature_dim)
        r = systems.init_electrons(mol, walkers, stddev=1.0, seed=int(rng.integers(1 << 31)))
        prm = FireWaveFunction(cfg).init(int(rng.integers(1 << 31)), r[0])
        exercise_jastrow(prm)
        out.append(Geometry(cfg, prm, r, int(rng.integers(1 << 31))))
    return out


def run_ragged(geoms: List[Geometry], sweeps: int, width: float = 0.3, precision: str = "mixed",
               use_graphs: bool = True, local_energy: bool = False, threads: int = 2) -> dict:
    """Sample every geometry for ``sweeps`` sweeps (after one warm-up sweep that also captures
    the graphs) with all geometries in flight at once; returns throughput and per-geometry
    acceptance.  With ``use_graphs=False`` the same steps are issued eagerly, one geometry
    after the other on the default stream (the reference point).  With ``local_energy`` the
    final walkers of every geometry are then evaluated: ``threads`` host threads, each with its
    own evaluation handle and stream, take the geometries in turn and retarget their handle to
    each (``fire_reconfigure`` keeps the device workspace, so the local-energy workspace held
    is that of ``threads`` handles, not one per geometry), and the evaluations of different
    threads overlap on the device; the result carries ``e_loc`` per geometry and the
    evaluation throughput."""
    wfs, streams, nacc, states = [], [], [], []
    for g in geoms:
        wf = FireWaveFunction(g.cfg, precision=precision)
        wf.autotune("near")
        wf.upload(g.params)
        wf.mcmc_options(use_graphs)
        s = torch.cuda.Stream() if use_graphs else torch.cuda.current_stream()
        with torch.cuda.stream(s):
            r = torch.as_tensor(g.walkers, device="cuda")
            wf.log_psi_with_state(g.params, r, precision=precision)
        wfs.append(wf)
        streams.append(s)
        nacc.append(torch.zeros(len(g.walkers), dtype=torch.int32, device="cuda"))

    def issue(step0: int, n_sweeps: int):
        for wf, g, s, na in zip(wfs, geoms, streams, nacc):
            N = g.cfg.n_el
            _lib.check(wf._lib.fire_mcmc_steps(wf._h, ctypes.c_void_p(s.cuda_stream), int(n_sweeps * N),
                                               float(width), int(g.seed), int(step0 * N),
                                               ctypes.c_void_p(na.data_ptr())))

    issue(0, 1)  # warm-up sweep (captures each geometry's sweep graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    issue(1, sweeps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    moves = sum(len(g.walkers) * g.cfg.n_el * sweeps for g in geoms)
    pmove = [float(na.double().mean()) / (g.cfg.n_el * sweeps) for na, g in zip(nacc, geoms)]
    walkers_out = []
    for wf in wfs:
        r, _, _ = wf.state_values()  # also reports a capacity overflow as an error
        walkers_out.append(r)
    out = {"geometries": len(geoms), "sweeps": sweeps, "seconds": dt, "walker_moves_per_s": moves / dt,
           "electrons": [g.cfg.n_el for g in geoms], "pmove": pmove, "walkers": walkers_out,
           "graphs": use_graphs}
    if local_energy:
        import queue
        import threading

        todo = queue.Queue()
        for k in sorted(range(len(geoms)), key=lambda k: -geoms[k].cfg.n_el):
            todo.put(k)
        e_loc = [None] * len(geoms)
        errors = []

        def worker():
            s = torch.cuda.Stream()
            wf = None
            try:
                while True:
                    try:
                        k = todo.get_nowait()
                    except queue.Empty:
                        return
                    g = geoms[k]
                    if wf is None:
                        wf = FireWaveFunction(g.cfg, precision=precision)
                        wf.autotune("near")
                    else:
                        wf.reconfigure(g.cfg)
                    with torch.cuda.stream(s):
                        e_loc[k] = wf.local_energy_terms(g.params, walkers_out[k], precision=precision)["e_loc"]
                    s.synchronize()
            except Exception as e:  # reported after the pool has drained
                errors.append(e)
            finally:
                del wf

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pool = [threading.Thread(target=worker) for _ in range(max(1, threads))]
        for t in pool:
            t.start()
        for t in pool:
            t.join()
        dt_e = time.perf_counter() - t0
        if errors:
            raise errors[0]
        out["e_loc"] = e_loc
        out["local_energy_seconds"] = dt_e
        out["local_energy_walkers_per_s"] = sum(len(g.walkers) for g in geoms) / dt_e
    return out
