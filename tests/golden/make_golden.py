"""Writes the golden fixtures of this directory from the FP64 oracle.

The reference ships no golden vectors and cannot run here (no jax / flax / folx), so the
fixtures come from the oracle (oracle/fire.py, itself pinned by tests/test_oracle.py).  Each
fixture holds the packed float32 parameters (the canonical list of fire_upload_params), the
walkers and the expected outputs of one small configuration:

* neighbour counts and CSR neighbour indices (exact);
* embeddings h_out [B, N, F];
* log|psi|, sign;
* E_loc, E_kin, E_pot, the gradient and Laplacian of log|psi|;
* log|psi| after the reference test's move (electron 3 by +2 bohr in x,
  tests/test_low_rank.py:36-45) through the oracle's low-rank update.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

CASES = ("small", "small_restricted")


def case_config(name):
    from cases import small_cfg

    return small_cfg(name == "small_restricted")


def compute(name):
    """Inputs and FP64 oracle outputs of one fixture (used by the script and by the CPU test
    that checks the committed files)."""
    import torch

    from cases import make_params, walkers
    from oracle import fire
    from sparse_wf_amd.params import pack_params

    cfg = case_config(name)
    prm = make_params(cfg, seed=0)
    r = walkers(cfg, 4, seed=21)
    pt = fire.to_torch(prm)
    packed = pack_params(cfg, prm)
    out = {"packed": np.concatenate([np.asarray(t, np.float32).ravel() for t in packed]),
           "sizes": np.asarray([t.size for t in packed], np.int64), "r": r}
    counts, nbrs, h, lp, sg, el, ek, ep, gr, lap, lp_lr = [], [], [], [], [], [], [], [], [], [], []
    for w in range(len(r)):
        nl = fire.neighbour_lists(r[w], cfg.n_up, cfg.cutoff)
        counts.append(np.diff(nl.offsets).astype(np.int32))
        nbrs.append(np.asarray(nl.nbr, np.int32))
        h.append(fire.embedding_apply(pt, cfg, r[w]).numpy())
        s, l = fire.signed(pt, cfg, r[w])
        lp.append(float(l))
        sg.append(float(s))
        e, k, p = fire.local_energy(pt, cfg, r[w])
        el.append(float(e))
        ek.append(float(k))
        ep.append(float(p))
        sgn, f = fire.log_psi_fwd_lap(pt, cfg, r[w])
        gr.append(f.j.numpy().reshape(-1, 3))
        lap.append(float(f.l))
        (_, _), st = fire.log_psi_with_state(pt, cfg, r[w])
        r_new = r[w].copy()
        r_new[3] += np.array([2.0, 0.0, 0.0], np.float32)
        (_, l2), _, _ = fire.log_psi_low_rank_update(pt, cfg, r_new, 3, st)
        lp_lr.append(float(l2))
    out.update(counts=np.stack(counts), nbr=np.concatenate(nbrs), h_out=np.stack(h), logpsi=np.array(lp),
               sign=np.array(sg), e_loc=np.array(el), e_kin=np.array(ek), e_pot=np.array(ep), grad=np.stack(gr),
               lap=np.array(lap), logpsi_after_move=np.array(lp_lr))
    return out


def path(name):
    return os.path.join(HERE, f"{name}.npz")


if __name__ == "__main__":
    for name in CASES:
        d = compute(name)
        np.savez_compressed(path(name), **d)
        print(name, {k: v.shape for k, v in d.items()}, os.path.getsize(path(name)), "bytes")
