"""Adjoint of the parameter packing (CPU): for a random cotangent c over the packed tensors and a
random direction d over the tree's leaves, sum(pack_params_vjp(c) * d) must equal c . (dpack/dθ) d,
the latter by central differences of pack_params in FP64.  Covers unrestricted and
spin-restricted orbitals, both Jastrows with attention, and the psiformer cusps."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from sparse_wf_amd import params as P
from tests.cases import small_cfg


def _flat(tree):
    return [(p, np.asarray(v, np.float64)) for p, v in P.leaves_with_path(tree)]


def _with(tree, delta, h):
    it = iter(delta)
    return P.tree_map(lambda x: np.asarray(x, np.float64) + h * next(it), tree)


@pytest.mark.parametrize("kw", [dict(spin_restricted=False), dict(spin_restricted=True),
                                dict(spin_restricted=False, e_e_cusps="none", use_attention=False),
                                dict(spin_restricted=True, use_log_jastrow=False, pair_mlp_widths=(4, 3))])
def test_pack_params_vjp_matches_directional_derivative(kw):
    cfg = small_cfg(**kw)
    prm = P.tree_map(lambda x: np.asarray(x, np.float64), P.init_params(cfg, 3))
    P.exercise_jastrow(prm)
    prm["embedding"]["scales"] = [np.float64(0.7), np.float64(1.3)]
    prm["embedding"]["cutoff_param"] = np.float64(0.4)
    rng = np.random.default_rng(5)
    leaves = _flat(prm)
    sizes = [t.size for t in P.pack_params(cfg, prm, np.float64)]
    B = 2
    c = rng.standard_normal((B, sum(sizes)))
    g = P.pack_params_vjp(cfg, prm, torch.as_tensor(c))
    gl = dict((p, v) for p, v in P.leaves_with_path(g))
    assert set(gl) == {p for p, _ in leaves}
    for trial in range(3):
        d = [rng.standard_normal(v.shape) for _, v in leaves]
        h = 1e-6
        plus = np.concatenate(P.pack_params(cfg, _with(prm, d, h), np.float64))
        minus = np.concatenate(P.pack_params(cfg, _with(prm, d, -h), np.float64))
        jd = (plus - minus) / (2 * h)
        for b in range(B):
            want = float(c[b] @ jd)
            got = sum(float((gl[p][b].numpy().reshape(v.shape) * dv).sum()) for (p, v), dv in zip(leaves, d))
            assert abs(got - want) <= 1e-6 * max(1.0, abs(want)), (trial, b, got, want)


def test_pack_params_vjp_rejects_wrong_width():
    cfg = small_cfg()
    prm = P.init_params(cfg, 0)
    with pytest.raises(ValueError):
        P.pack_params_vjp(cfg, prm, torch.zeros(1, 7))
