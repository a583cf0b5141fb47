"""CPU restatement of the single-electron Metropolis step (oracle; test infrastructure only).

Follows mcmc_steps_low_rank and proposal_single_electron (src/sparse_wf/mcmc.py:100-111,
196-275): step s moves electron k = s mod n_el of every walker by width * N(0, 1) per
coordinate and accepts when 2 (log|psi'| - log|psi|) > log u, u ~ U(0, 1) (:233-238).

The reference draws its normal and uniform variates from JAX's threefry stream, which cannot
run here.  The HIP sampler draws them from Philox4x32-10 (Salmon, Moraes, Dror and Shaw,
"Parallel random numbers: as easy as 1, 2, 3", SC'11) keyed by (seed, walker) with the step
as counter, turned into normals by the Box-Muller transform; this module restates that
stream so that single proposals and accept decisions can be compared one to one.  The
generator is pinned against the published Philox4x32-10 known-answer vectors
(tests/test_mcmc_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np

_M0, _M1 = 0xD2511F53, 0xCD9E8D57
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = 0xFFFFFFFF


def philox4x32_10(counter, key):
    """Philox4x32 with 10 rounds on uint32 arrays: counter [..., 4], key [..., 2] -> [..., 4]."""
    c = np.asarray(counter, np.uint64) & _MASK
    k = np.asarray(key, np.uint64) & _MASK
    c0, c1, c2, c3 = (c[..., i].copy() for i in range(4))
    k0, k1 = k[..., 0].copy(), k[..., 1].copy()
    for _ in range(10):
        p0 = c0 * np.uint64(_M0)
        p1 = c2 * np.uint64(_M1)
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(_MASK)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(_MASK)
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
        k0 = (k0 + np.uint64(_W0)) & np.uint64(_MASK)
        k1 = (k1 + np.uint64(_W1)) & np.uint64(_MASK)
    return np.stack([c0, c1, c2, c3], -1).astype(np.uint32)


def _u01(x):
    """(x >> 8 + 1/2) / 2^24: a uniform variate strictly inside (0, 1)."""
    return ((np.asarray(x, np.uint64) >> np.uint64(8)).astype(np.float64) + 0.5) * (1.0 / 16777216.0)


def step_randoms(seed: int, walkers, step: int):
    """Three N(0, 1) variates and one U(0, 1) variate per walker for step ``step``: Philox
    blocks (step_lo, step_hi, 0, 0) and (step_lo, step_hi, 1, 0) under the key
    (seed_lo ^ seed_hi, walker); Box-Muller on the first block, u from the second."""
    w = np.asarray(walkers, np.uint64)
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    kx = (seed & _MASK) ^ (seed >> 32)
    key = np.stack([np.full_like(w, kx), w], -1)
    s_lo, s_hi = step & _MASK, (step >> 32) & _MASK
    ctr_a = np.stack([np.full_like(w, s_lo), np.full_like(w, s_hi), np.zeros_like(w), np.zeros_like(w)], -1)
    ctr_b = ctr_a.copy()
    ctr_b[..., 2] = 1
    a = philox4x32_10(ctr_a, key)
    b = philox4x32_10(ctr_b, key)
    r1 = np.sqrt(-2.0 * np.log(_u01(a[..., 0])))
    r2 = np.sqrt(-2.0 * np.log(_u01(a[..., 2])))
    t1 = 2.0 * math.pi * _u01(a[..., 1])
    t2 = 2.0 * math.pi * _u01(a[..., 3])
    n3 = np.stack([r1 * np.cos(t1), r1 * np.sin(t1), r2 * np.cos(t2)], -1)
    return n3, _u01(b[..., 0])


def propose(r, k: int, width: float, seed: int, step: int):
    """proposal_single_electron (mcmc.py:100-111): walkers r [B, N, 3] float32 with electron k
    moved by the step's normal variates times ``width`` (rounded to float32 before the add);
    returns (proposed walkers, u [B])."""
    r = np.asarray(r, np.float32)
    n3, u = step_randoms(seed, np.arange(r.shape[0]), step)
    out = r.copy()
    out[:, k] = r[:, k] + (n3 * np.float64(np.float32(width))).astype(np.float32)
    return out, u


def accept(logpsi_old, logpsi_new, u):
    """Metropolis decision of mcmc_steps_low_rank (mcmc.py:237-238) for |psi|^2:
    2 (log|psi'| - log|psi|) > log u."""
    return 2.0 * (np.asarray(logpsi_new, np.float64) - np.asarray(logpsi_old, np.float64)) > np.log(u)
