"""Continuous-batching evaluation server on one GPU.

A request carries a geometry, parameters, walkers and a number of Metropolis sweeps; its
result is the local energy of the walkers after the sweeps (the VMC evaluation the reference
runs between optimisation steps: ``make_mcmc`` then ``make_local_energy``, mcmc.py:278-346,
hamiltonian.py:49-81).  Requests arrive over time; the scheduler admits a queued request as
soon as a slot is free and its estimated HBM footprint fits next to the requests already on
the device.  Every admitted request runs on its own HIP stream (default priority), driven by
its own host thread, with its walker state and FP64 inverses resident in HBM until it
finishes, so the work of concurrent requests overlaps on the GPU.  A request's result does
not depend on what else runs: every kernel works per walker and the random numbers are keyed
by (seed, walker, step).
"""
from __future__ import annotations

import heapq
import threading
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from .config import FireConfig
from .wave_function import FireWaveFunction


@dataclass
class Request:
    rid: int
    cfg: FireConfig
    params: dict
    walkers: np.ndarray          # [B, N, 3] float32
    sweeps: int
    width: float = 0.3
    seed: int = 0
    precision: str = "mixed"
    arrival: float = 0.0         # seconds after the start of the stream


@dataclass
class Result:
    rid: int
    e_loc: np.ndarray            # [B] local energies after the sweeps
    e_mean: float
    e_stderr: float
    pmove: float
    walkers: np.ndarray          # [B, N, 3] final positions
    arrival: float
    admitted: float
    started: float
    finished: float
    footprint_bytes: int
    error: Optional[str] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def footprint_bytes(cfg: FireConfig, n_walkers: int) -> int:
    """Estimated peak device memory of one request: the resident walker state, the value-path
    workspace of the state initialisation (all walkers at once) and the forward-Laplacian
    workspace of one walker chunk (its widest buffers)."""
    N, F, D = cfg.n_el, cfg.feature_dim, cfg.n_determinants
    ND, W, HD = N * D, cfg.mlp_width, cfg.attention_heads * cfg.attention_dim
    M, E = cfg.n_nuc, cfg.n_envelopes
    G = n_walkers * N
    state = G * (5 * F + F + 2 * ND + 2 + HD + cfg.attention_heads + 3) * 4 + 2 * n_walkers * D * N * N * 8
    value_ws = G * (5 * F + 4 * F + 2 * ND + 2 * HD) * 4 + D * G * M * E * 4
    rows = 5 * N + 3 * N * min(N - 1, 40)
    width = max(F, ND, W, 2 * HD)
    chunk = max(8, min(256, int(12e9 // (rows * width * 4))))
    tangent_ws = chunk * rows * (3 * max(F, W) + 2 * width + 4 * F) * 4 + D * chunk * 5 * N * M * E * 4
    return int(1.15 * (state + value_ws + tangent_ws)) + (256 << 20)


class Server:
    """Admits requests by HBM footprint and slot count; runs each on its own stream."""

    def __init__(self, device: int = 0, hbm_budget: Optional[int] = None, max_active: int = 8):
        self.device = device
        total = torch.cuda.get_device_properties(device).total_memory
        self.hbm_budget = int(hbm_budget if hbm_budget is not None else 0.8 * total)
        self.max_active = max_active
        self._lock = threading.Condition()
        self._used = 0
        self._active = 0
        self.max_concurrent = 0

    def _admit(self, need: int) -> bool:
        return self._active < self.max_active and (self._active == 0 or self._used + need <= self.hbm_budget)

    def _run_one(self, req: Request, t0: float, admitted: float, need: int, out: List[Result]):
        stream = torch.cuda.Stream(device=self.device)  # default priority
        started = time.perf_counter() - t0
        err = None
        try:
            with torch.cuda.device(self.device), torch.cuda.stream(stream):
                wf = FireWaveFunction(req.cfg, device=self.device, precision=req.precision)
                wf._lib.fire_autotune(wf._h, 0, None, 0, None)  # fixed tiles: no timing syncs
                wf.mcmc_options(False)  # steps issued eagerly: no stream capture from worker threads
                wf.upload(req.params)
                r = torch.as_tensor(req.walkers, device=f"cuda:{self.device}")
                wf.log_psi_with_state(req.params, r, precision=req.precision)
                r, st = wf.mcmc_steps_low_rank(req.params, r, req.sweeps * req.cfg.n_el, width=req.width,
                                               seed=req.seed, keep_state=True)
                e = wf.local_energy_terms(req.params, r, precision=req.precision)["e_loc"]
                stream.synchronize()
                e_np = e.cpu().numpy()
                r_np = r.cpu().numpy()
                pmove = st["pmove"]
                del wf
        except Exception as ex:  # reported in the result, the server keeps running
            err = f"{type(ex).__name__}: {ex}"
            e_np = np.full(len(req.walkers), np.nan)
            r_np = req.walkers
            pmove = float("nan")
        finished = time.perf_counter() - t0
        res = Result(req.rid, e_np, float(np.mean(e_np)), float(np.std(e_np) / np.sqrt(len(e_np))), pmove, r_np,
                     req.arrival, admitted, started, finished, need, err)
        with self._lock:
            out.append(res)
            self._used -= need
            self._active -= 1
            self._lock.notify_all()

    def run(self, requests: List[Request]) -> List[Result]:
        """Serve ``requests`` (each with its arrival time) to completion; returns the results in
        completion order.  Arrivals are replayed in real time from the call."""
        pending = sorted(requests, key=lambda q: q.arrival)
        queue: list = []   # (arrival, rid, request) waiting for admission
        out: List[Result] = []
        threads = []
        t0 = time.perf_counter()
        i = 0
        with self._lock:
            while i < len(pending) or queue or self._active:
                now = time.perf_counter() - t0
                while i < len(pending) and pending[i].arrival <= now:
                    heapq.heappush(queue, (pending[i].arrival, pending[i].rid, pending[i]))
                    i += 1
                # admit in arrival order while slots and memory allow
                while queue:
                    req = queue[0][2]
                    need = footprint_bytes(req.cfg, len(req.walkers))
                    if not self._admit(need):
                        break
                    heapq.heappop(queue)
                    self._used += need
                    self._active += 1
                    self.max_concurrent = max(self.max_concurrent, self._active)
                    th = threading.Thread(target=self._run_one, args=(req, t0, now, need, out), daemon=True)
                    th.start()
                    threads.append(th)
                if i < len(pending):
                    timeout = max(0.0, pending[i].arrival - (time.perf_counter() - t0))
                    self._lock.wait(timeout=timeout)
                elif queue or self._active:
                    self._lock.wait(timeout=1.0)
        for th in threads:
            th.join()
        return out


def latency_summary(results: List[Result]) -> dict:
    lat = np.array([r.latency for r in results])
    walk = sum(len(r.e_loc) for r in results)
    span = max(r.finished for r in results) - min(r.arrival for r in results)
    return {"requests": len(results), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "walkers_per_s": walk / span, "span_s": span,
            "errors": [r.error for r in results if r.error]}


def poisson_requests(n: int, qps: float, seed: int, n_el=(20, 200), walkers=(128, 1024), sweeps=(1, 3),
                     precision: str = "mixed", feature_dim: int = 256, elements=("H", "C")) -> List[Request]:
    """Seeded Poisson stream of random-chain requests with random-init parameters.  The
    electron-count, walker and sweep ranges are arguments so the stream can be sized to a run.
    Parameters are prepared here (client side, before the stream starts): reference
    initialisers with the normalisation scales set on the request's first walker
    (``FireWaveFunction.init``) and the Jastrow exercised."""
    from . import systems
    from .params import exercise_jastrow

    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for rid in range(n):
        t += rng.exponential(1.0 / qps)
        N = int(rng.integers(n_el[0], n_el[1] + 1))
        mol = systems.random_chain(rng, N, elements=elements)
        cfg = FireConfig(mol=mol, cutoff=5.0, feature_dim=feature_dim)
        B = int(rng.integers(walkers[0], walkers[1] + 1))
        r = systems.init_electrons(mol, B, stddev=1.0, seed=int(rng.integers(1 << 31)))
        prm = FireWaveFunction(cfg).init(int(rng.integers(1 << 31)), r[0])
        exercise_jastrow(prm)
        reqs.append(Request(rid, cfg, prm, r, int(rng.integers(sweeps[0], sweeps[1] + 1)), 0.3,
                            int(rng.integers(1 << 31)), precision, t))
    return reqs
