"""Continuous-batching evaluation server on one GPU.

A request carries a geometry, parameters, walkers and a number of Metropolis sweeps; its
result is the local energy of the walkers after the sweeps (the VMC evaluation the reference
runs between optimisation steps: ``make_mcmc`` then ``make_local_energy``, mcmc.py:278-346,
hamiltonian.py:49-81).

Design (one host thread, no per-request allocation of the big buffers):

* a pool of ``max_active`` slots, each a GPU handle with its own HIP stream (default priority),
  created on first use and afterwards retargeted to the next request's molecule with
  ``fire_reconfigure``, which keeps the device workspace;
* admission: a queued request is admitted, in arrival order, as soon as a slot is free and its
  HBM footprint (the larger of its estimate and what the slot already holds) fits next to the
  admitted requests and the idle slots' workspace; idle slots give their workspace back when
  that is what a request needs;
* dispatch: the thread walks the admitted requests round robin and keeps up to two sweeps of
  each in flight on its stream -- each sweep a replay of the request's captured hipGraph -- then
  its local energy; completion is polled with events, so sweeps of different requests overlap
  on the device while walker states and FP64 inverses stay resident;
* the state of a request is rebuilt from its positions every ``refresh_sweeps`` sweeps (FP64
  low-rank updates drift on long chains).

A request's result does not depend on what else runs: every kernel works per walker, random
numbers are keyed by (seed, walker, step), and tile shapes are chosen per contraction shape
once per process.
"""
from __future__ import annotations

import ctypes
import heapq
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import torch

from . import _lib
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
    moves: int = 0               # walker-moves of the sweeps (walkers x electrons x sweeps)

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


class _Slot:
    """A reusable GPU handle and stream."""

    def __init__(self, device: int):
        self.device = device
        self.wf: Optional[FireWaveFunction] = None
        self.stream = torch.cuda.Stream(device=device)
        self.retained = 0  # device bytes the idle handle holds
        self.created = 0   # handles created (a retargeted handle is not a new one)

    def take(self, cfg: FireConfig, precision: str) -> FireWaveFunction:
        if self.wf is None:
            self.wf = FireWaveFunction(cfg, device=self.device, precision=precision)
            self.wf.autotune("near")  # every request brings new shapes
            self.created += 1
        else:
            self.wf.reconfigure(cfg)
            self.wf.precision = precision
        return self.wf

    def release_memory(self):
        self.wf = None  # the handle frees its workspace
        self.retained = 0


class _Job:
    """An admitted request and where it stands."""

    def __init__(self, req: Request, slot: _Slot, need: int, admitted: float):
        self.req, self.slot, self.need, self.admitted = req, slot, need, admitted
        self.started = admitted
        self.issued = 0            # sweeps issued
        self.inflight: list = []   # events of sweeps not yet known complete
        self.nacc_total = None
        self.nacc = None
        self.e_loc = None
        self.done_event = None
        self.r = None


class Server:
    """Admits requests by HBM footprint and slot count; runs them on a pool of reusable
    handles, one stream each, from one host thread."""

    def __init__(self, device: int = 0, hbm_budget: Optional[int] = None, max_active: int = 8,
                 refresh_sweeps: int = 10, inflight: int = 2):
        self.device = device
        total = torch.cuda.get_device_properties(device).total_memory
        self.hbm_budget = int(hbm_budget if hbm_budget is not None else 0.8 * total)
        self.max_active = max_active
        self.refresh_sweeps = refresh_sweeps
        self.inflight = inflight
        self.slots = [_Slot(device) for _ in range(max_active)]
        self.max_concurrent = 0

    @property
    def handles_created(self) -> int:
        return sum(s.created for s in self.slots)

    # admission ---------------------------------------------------------------------------
    def _cost(self, active: List[_Job], idle: List[_Slot], need: int, slot: _Slot) -> int:
        held = sum(max(j.need, j.slot.retained) for j in active) + sum(s.retained for s in idle if s is not slot)
        return held + max(need, slot.retained)

    def _admit(self, req: Request, active: List[_Job], now: float) -> Optional[_Job]:
        idle = [s for s in self.slots if all(j.slot is not s for j in active)]
        if not idle:
            return None
        need = footprint_bytes(req.cfg, len(req.walkers))
        # prefer the idle slot whose held workspace is closest above the need
        idle.sort(key=lambda s: (s.retained < need, abs(s.retained - need)))
        slot = idle[0]
        if active and self._cost(active, idle, need, slot) > self.hbm_budget:
            for s in idle:  # give back idle workspace, then try again
                if s is not slot:
                    s.release_memory()
            if self._cost(active, idle, need, slot) > self.hbm_budget:
                return None
        return _Job(req, slot, need, now)

    # per-request work --------------------------------------------------------------------
    def _start(self, job: _Job):
        req = job.req
        wf = job.slot.take(req.cfg, req.precision)
        with torch.cuda.device(self.device), torch.cuda.stream(job.slot.stream):
            wf.upload(req.params)
            job.r = torch.as_tensor(req.walkers, device=f"cuda:{self.device}")
            wf.log_psi_with_state(req.params, job.r, precision=req.precision)
            B = len(req.walkers)
            job.nacc = torch.zeros(B, dtype=torch.int32, device=f"cuda:{self.device}")
            job.nacc_total = torch.zeros(B, dtype=torch.int64, device=f"cuda:{self.device}")

    def _issue_sweep(self, job: _Job):
        req, wf, st = job.req, job.slot.wf, job.slot.stream
        N = req.cfg.n_el
        with torch.cuda.device(self.device), torch.cuda.stream(st):
            if job.issued > 0 and self.refresh_sweeps > 0 and job.issued % self.refresh_sweeps == 0:
                wf.log_psi_with_state(req.params, wf.state_values()[0], precision=req.precision)
            _lib.check(wf._lib.fire_mcmc_steps(wf._h, ctypes.c_void_p(st.cuda_stream), N, float(req.width),
                                               int(req.seed) & (2**64 - 1), job.issued * N,
                                               ctypes.c_void_p(job.nacc.data_ptr())))
            job.nacc_total += job.nacc.long()
            ev = torch.cuda.Event()
            ev.record(st)
        job.inflight.append(ev)
        job.issued += 1

    def _issue_energy(self, job: _Job):
        req, wf, st = job.req, job.slot.wf, job.slot.stream
        with torch.cuda.device(self.device), torch.cuda.stream(st):
            job.r, _, _ = wf.state_values()
            job.e_loc = wf.local_energy_terms(req.params, job.r, precision=req.precision)["e_loc"]
            job.done_event = torch.cuda.Event()
            job.done_event.record(st)

    def _finish(self, job: _Job, t0: float, err: Optional[str] = None) -> Result:
        req = job.req
        if err is None:
            e_np = job.e_loc.cpu().numpy()
            r_np = job.r.cpu().numpy()
            pmove = float(job.nacc_total.double().mean()) / max(1, req.sweeps * req.cfg.n_el)
            job.slot.retained = job.slot.wf.workspace_bytes()
        else:
            e_np = np.full(len(req.walkers), np.nan)
            r_np = req.walkers
            pmove = float("nan")
            job.slot.release_memory()
        return Result(req.rid, e_np, float(np.mean(e_np)), float(np.std(e_np) / np.sqrt(len(e_np))), pmove, r_np,
                      req.arrival, job.admitted, job.started, time.perf_counter() - t0, job.need, err,
                      len(req.walkers) * req.cfg.n_el * req.sweeps if err is None else 0)

    # main loop ---------------------------------------------------------------------------
    def run(self, requests: List[Request]) -> List[Result]:
        """Serve ``requests`` (each with its arrival time) to completion; returns the results in
        completion order.  Arrivals are replayed in real time from the call."""
        pending = sorted(requests, key=lambda q: q.arrival)
        queue: list = []  # (arrival, rid, request) waiting for admission
        active: List[_Job] = []
        out: List[Result] = []
        t0 = time.perf_counter()
        i = 0
        while i < len(pending) or queue or active:
            now = time.perf_counter() - t0
            while i < len(pending) and pending[i].arrival <= now:
                heapq.heappush(queue, (pending[i].arrival, pending[i].rid, pending[i]))
                i += 1
            while queue:
                job = self._admit(queue[0][2], active, now)
                if job is None:
                    break
                heapq.heappop(queue)
                try:
                    job.started = time.perf_counter() - t0
                    self._start(job)
                except Exception as ex:  # reported in the result, the server keeps running
                    out.append(self._finish(job, t0, f"{type(ex).__name__}: {ex}"))
                    continue
                active.append(job)
                self.max_concurrent = max(self.max_concurrent, len(active))
            progressed = False
            for job in list(active):
                try:
                    job.inflight = [e for e in job.inflight if not e.query()]
                    if job.issued < job.req.sweeps:
                        if len(job.inflight) < self.inflight:
                            self._issue_sweep(job)
                            progressed = True
                    elif job.done_event is None:
                        if not job.inflight:
                            self._issue_energy(job)
                            progressed = True
                    elif job.done_event.query():
                        out.append(self._finish(job, t0))
                        active.remove(job)
                        progressed = True
                except Exception as ex:
                    torch.cuda.synchronize(self.device)
                    out.append(self._finish(job, t0, f"{type(ex).__name__}: {ex}"))
                    active.remove(job)
                    progressed = True
            if not progressed:
                # nothing to issue: wait a little for the device or the next arrival
                wait = 2e-4
                if not active and i < len(pending):
                    wait = max(0.0, pending[i].arrival - (time.perf_counter() - t0))
                time.sleep(min(wait, 0.05))
        return out


def latency_summary(results: List[Result]) -> dict:
    lat = np.array([r.latency for r in results])
    walk = sum(len(r.e_loc) for r in results)
    span = max(r.finished for r in results) - min(r.arrival for r in results)
    return {"requests": len(results), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "walkers_per_s": walk / span,
            "walker_moves_per_s": sum(r.moves for r in results) / span, "span_s": span,
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
    wf = None
    for rid in range(n):
        t += rng.exponential(1.0 / qps)
        N = int(rng.integers(n_el[0], n_el[1] + 1))
        mol = systems.random_chain(rng, N, elements=elements)
        cfg = FireConfig(mol=mol, cutoff=5.0, feature_dim=feature_dim)
        B = int(rng.integers(walkers[0], walkers[1] + 1))
        r = systems.init_electrons(mol, B, stddev=1.0, seed=int(rng.integers(1 << 31)))
        if wf is None:
            wf = FireWaveFunction(cfg)
        else:
            wf.reconfigure(cfg)
        prm = wf.init(int(rng.integers(1 << 31)), r[0])
        exercise_jastrow(prm)
        reqs.append(Request(rid, cfg, prm, r, int(rng.integers(sweeps[0], sweeps[1] + 1)), 0.3,
                            int(rng.integers(1 << 31)), precision, t))
    return reqs
