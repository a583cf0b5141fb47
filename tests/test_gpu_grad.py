"""Per-walker parameter gradient of log|psi| (fire_grad_logpsi + the adjoint of the packing)
against the oracle's FP64 autograd gradient on the same FP32 parameters and walkers.

Bound: for every leaf, |g_gpu - g_oracle| <= 5e-4 |g_oracle| + 1e-5 |g_oracle of the whole
tree| (Frobenius norms over the walker's leaf).  The forward pass is FP32 with FP64
determinants; the largest leaf errors measured on MI355X for these walkers are about 6e-5 of
the leaf's norm (edge filter kernels, the dense ElecInit weights)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fire
from sparse_wf_amd import params as P
from sparse_wf_amd.wave_function import FireWaveFunction

from cases import alkane_cfg, make_params, small_cfg, walkers, water_cfg

pytestmark = pytest.mark.gpu

CASES = {
    "small": lambda: small_cfg(False),
    "small_restricted": lambda: small_cfg(True),
    "small_nocusp_noatt": lambda: small_cfg(False, e_e_cusps="none", use_attention=False, pair_mlp_widths=(4, 3)),
    "water": lambda: water_cfg(),
    "alkane130": lambda: alkane_cfg(),
}


@pytest.mark.parametrize("name", list(CASES))
def test_grad_logpsi_matches_oracle(name, gpu_device):
    cfg = CASES[name]()
    prm = make_params(cfg, seed=0)
    wf = FireWaveFunction(cfg, device=gpu_device)
    r = walkers(cfg, 3, seed=7)
    print(f"\n[grad {name}] N={cfg.n_el} F={cfg.feature_dim} D={cfg.n_determinants}", flush=True)
    g, lp = wf.grad_logpsi(prm, torch.as_tensor(r, device="cuda"))
    _, lp_ref = wf.signed(prm, torch.as_tensor(r, device="cuda"), precision="fp32")
    torch.cuda.synchronize()
    assert np.allclose(lp.cpu().numpy(), lp_ref.cpu().numpy(), rtol=0, atol=1e-3)
    gl = {p: v.cpu().numpy().astype(np.float64) for p, v in P.leaves_with_path(g)}
    prm64 = P.tree_map(lambda x: np.asarray(x, np.float64), prm)
    worst = []
    for b in range(r.shape[0]):
        ref = fire.grad_logpsi(prm64, cfg, r[b])
        assert set(ref) == set(gl)
        total = np.sqrt(sum(float((v ** 2).sum()) for v in ref.values()))
        for path, want in ref.items():
            got = gl[path][b].reshape(want.shape)
            err = float(np.linalg.norm(got - want))
            scale = float(np.linalg.norm(want))
            worst.append((err / (scale + 1e-2 * total), path, b, err, scale))
            assert err <= 5e-4 * scale + 1e-5 * total, (path, b, err, scale, total)
    worst.sort(reverse=True)
    print("  worst:", worst[:3], flush=True)


def test_grad_logpsi_across_chunks_and_empty(gpu_device):
    """260 walkers run as chunks of 128, 128 and 4: walkers on either side of each chunk
    boundary match the oracle to the same bound; an empty batch returns empty rows."""
    cfg = small_cfg(False)
    prm = make_params(cfg, seed=0)
    wf = FireWaveFunction(cfg, device=gpu_device)
    r = walkers(cfg, 260, seed=11)
    print("\n[grad chunks] B=260", flush=True)
    g, lp = wf.grad_logpsi(prm, torch.as_tensor(r, device="cuda"))
    torch.cuda.synchronize()
    gl = {p: v.cpu().numpy().astype(np.float64) for p, v in P.leaves_with_path(g)}
    prm64 = P.tree_map(lambda x: np.asarray(x, np.float64), prm)
    for b in (0, 127, 128, 255, 256, 259):
        ref = fire.grad_logpsi(prm64, cfg, r[b])
        total = np.sqrt(sum(float((v ** 2).sum()) for v in ref.values()))
        for path, want in ref.items():
            err = float(np.linalg.norm(gl[path][b].reshape(want.shape) - want))
            assert err <= 5e-4 * float(np.linalg.norm(want)) + 1e-5 * total, (b, path, err)
    n = int(sum(t.size for t in P.pack_params(cfg, prm)))
    g0, lp0 = wf.grad_logpsi(prm, torch.empty((0, cfg.n_el, 3), device="cuda"), packed=True)
    assert tuple(g0.shape) == (0, n) and tuple(lp0.shape) == (0,)
