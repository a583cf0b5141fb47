"""The oracle's parameter gradient of log|psi| (torch autograd of the value path) against
central finite differences on a sample of entries of every kind of leaf (FP64)."""
from __future__ import annotations

import numpy as np

from oracle import fire
from sparse_wf_amd import params as P
from tests.cases import make_params, small_cfg, walkers


def _get(tree, path):
    for k in path.strip("/").split("/"):
        tree = tree[int(k)] if isinstance(tree, list) else tree[k]
    return tree


def _set(tree, path, value):
    keys = path.strip("/").split("/")
    for k in keys[:-1]:
        tree = tree[int(k)] if isinstance(tree, list) else tree[k]
    last = keys[-1]
    if isinstance(tree, list):
        tree[int(last)] = value
    else:
        tree[last] = value


def test_grad_matches_finite_differences():
    cfg = small_cfg(False)
    prm = P.tree_map(lambda x: np.asarray(x, np.float64), make_params(cfg, seed=1))
    r = walkers(cfg, 1, seed=4)[0]
    g = fire.grad_logpsi(prm, cfg, r)
    rng = np.random.default_rng(0)
    checked = 0
    for path in g:
        base = np.asarray(_get(prm, path), np.float64)
        for _ in range(2):
            idx = tuple(int(rng.integers(s)) for s in base.shape)
            h = 1e-5 * max(1.0, abs(float(base[idx])))
            vals = []
            for sgn in (1.0, -1.0):
                x = base.copy()
                x[idx] += sgn * h
                _set(prm, path, x if x.ndim else np.float64(x))
                vals.append(float(fire.signed(fire.to_torch(prm), cfg, r)[1]))
            _set(prm, path, base if base.ndim else np.float64(base))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - float(g[path][idx])) <= 1e-5 * max(1.0, abs(fd)), (path, idx, fd, float(g[path][idx]))
            checked += 1
    assert checked >= 2 * 50
