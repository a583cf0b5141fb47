"""Shared small test systems: configurations, parameters and walkers."""
from __future__ import annotations

import functools

import numpy as np

from sparse_wf_amd import systems
from sparse_wf_amd.config import FireConfig
from sparse_wf_amd import params as P


def small_cfg(spin_restricted=False, **kw):
    """The reference test fixture's system (10 He atoms at 1 bohr, tests/utils.py:14-63) with
    F = 64 (the HIP kernels need F % 64 == 0) and no e-e MLP Jastrow."""
    mol = systems.atom_chain(10, 2)
    base = dict(cutoff=3.0, cutoff_1el=20.0, feature_dim=64, pair_mlp_widths=(4, 4), pair_n_envelopes=4,
                n_determinants=2, spin_restricted=spin_restricted, n_envelopes=8, mlp_width=32, mlp_depth=2,
                attention_heads=4, attention_dim=4)
    base.update(kw)
    return FireConfig(mol=mol, **base)


def water_cfg(**kw):
    mol, over, _, _ = systems.workload("h2o_like")
    base = dict(over)
    base.update(kw)
    return FireConfig(mol=mol, **base)


def alkane_cfg(**kw):
    """C16H34 (130 electrons) with one determinant and small layers: N = 130 puts the
    orbital-count dimension one column tile plus a narrow remainder wide (the FP32 column-tail
    kernel), and ND = 130 is not a multiple of 4 (the scalar determinant-Laplacian loads)."""
    mol = systems.zigzag_alkane(16)
    base = dict(cutoff=3.0, feature_dim=64, pair_mlp_widths=(4, 4), pair_n_envelopes=4, n_determinants=1,
                n_envelopes=8, mlp_width=32, mlp_depth=2, attention_heads=4, attention_dim=4)
    base.update(kw)
    return FireConfig(mol=mol, **base)


def make_params(cfg, seed=0, r_init=None, embed_fn=None):
    """Random params with the Jastrow exercised and the normalisation scales set from the
    oracle embedding on ``r_init`` (one walker)."""
    from oracle import fire

    prm = P.init_params(cfg, seed)
    P.exercise_jastrow(prm)
    if r_init is None:
        r_init = systems.init_electrons(cfg.mol, 1, seed=seed + 100)[0]

    def stages(pp):
        a, b = fire.embedding_apply(fire.to_torch(pp), cfg, r_init, return_stages=True)
        return a.numpy(), b.numpy()

    return P.set_scales(prm, embed_fn or stages)


def walkers(cfg, B, seed=1, stddev=1.0):
    return systems.init_electrons(cfg.mol, B, stddev=stddev, seed=seed)
