"""The benchmark configurations at their real model sizes, against the FP64 oracle.

* H64 chain (64 electrons, F = 256, 4 determinants, cutoff 5): local energy, log psi and
  gradient of a few walkers, FP32 kernels and mixed precision;
* the mixed-precision batch-mean energy bias against FP64 on 256 equilibrated H64 walkers
  (|bias| <= 1 mHa, the north-star gate), with its standard error;
* ragged batch (random H/C/N/O chains with 10-200 electrons, F = 256): after a sweep with
  every geometry's graph in flight, the local energy of sampled walkers;
* serving (random H/C chains with 20-400 electrons, F = 256, 10 sweeps): the local energy of
  a walker of every request.

Bounds are absolute (Hartree) and sit a few times above the errors measured on MI355X; the
per-walker energy bound of the ragged and serving checks grows with |Laplacian of log psi|
(near-node amplification, see tests/test_gpu_parity.py).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fire, pool
from sparse_wf_amd import params as P
from sparse_wf_amd import systems
from sparse_wf_amd.config import FireConfig
from sparse_wf_amd.wave_function import FireWaveFunction

pytestmark = pytest.mark.gpu

# absolute bounds per walker: E_loc (Ha), log psi, relative gradient norm
BOUNDS = {"fp32": (1e-3, 1e-4, 1e-4), "mixed": (1e-3, 1e-4, 1e-4)}


def _oracle_energy(pt, cfg, r):
    """FP64 oracle (E_loc, Laplacian of log psi) of one walker."""
    _, lp = fire.log_psi_fwd_lap(pt, cfg, r)
    ek = float(-0.5 * (lp.l + (lp.j * lp.j).sum()))
    return ek + fire.potential_energy(r, np.asarray(cfg.mol.R, np.float32), cfg.mol.Z), float(lp.l)


@pytest.fixture(scope="module")
def h64(gpu_device):
    mol, over, _, sigma = systems.workload("h64")
    cfg = FireConfig(mol=mol, **over)
    r0 = systems.init_electrons(mol, 256, stddev=sigma, seed=0)
    wf = FireWaveFunction(cfg, device=gpu_device)
    prm = wf.init(0, r0[0])
    P.exercise_jastrow(prm)
    r, _ = wf.mcmc_steps_low_rank(prm, r0, 5 * cfg.n_el, width=0.3, seed=1, precision="mixed")
    return cfg, prm, wf, r.cpu().numpy()


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_h64_against_fp64(h64, precision):
    cfg, prm, wf, r = h64
    pt = fire.to_torch(prm)
    out = wf.local_energy_terms(prm, r[:4], precision=precision, with_derivatives=True)
    tol_e, tol_lp, tol_g = BOUNDS[precision]
    for w in range(4):
        sign, lp = fire.log_psi_fwd_lap(pt, cfg, r[w])
        e_ref = float(-0.5 * (lp.l + (lp.j * lp.j).sum())) + fire.potential_energy(
            r[w], np.asarray(cfg.mol.R, np.float32), cfg.mol.Z)
        de = float(out["e_loc"][w]) - e_ref
        dl = float(out["logpsi"][w]) - float(lp.x)
        g = out["grad"][w].cpu().numpy().reshape(-1)
        dg = float(np.linalg.norm(g - lp.j.numpy()) / np.linalg.norm(lp.j.numpy()))
        print(f"  {precision} walker {w}: dE {de * 1e3:+.3f} mHa  dlogpsi {dl:+.2e}  grad rel {dg:.2e}", flush=True)
        assert float(out["sign"][w]) == float(sign)
        assert abs(de) < tol_e and abs(dl) < tol_lp and dg < tol_g


def test_h64_mixed_bias_against_fp64(h64):
    """Batch-mean local-energy bias of mixed precision against the FP64 oracle."""
    cfg, prm, wf, r = h64
    n = 256
    e_mixed = wf.local_energy_terms(prm, r[:n], precision="mixed")["e_loc"].cpu().numpy()
    e_fp32 = wf.local_energy_terms(prm, r[:n], precision="fp32")["e_loc"].cpu().numpy()
    e_ref, _, _ = pool.local_energies(prm, cfg, r[:n])
    for name, e in (("mixed", e_mixed), ("fp32", e_fp32)):
        d = (e - e_ref) * 1e3
        bias, se = float(d.mean()), float(d.std(ddof=1) / np.sqrt(n))
        print(f"  {name}: bias {bias:+.4f} +- {se:.4f} mHa, rms {np.sqrt((d ** 2).mean()):.4f}, "
              f"max {np.abs(d).max():.3f} mHa over {n} walkers", flush=True)
        assert abs(bias) <= 1.0
        assert abs(bias) <= 1.0 - 3 * se or se < 0.3  # the gate is resolved by the sample


def test_ragged_real_shapes(gpu_device):
    from sparse_wf_amd.ragged import random_geometries, run_ragged

    geoms = random_geometries(4, seed=17, n_el=(10, 200), walkers=64, feature_dim=256)
    geoms.sort(key=lambda g: g.cfg.n_el)
    print("\n[ragged real shapes] N:", [g.cfg.n_el for g in geoms], flush=True)
    res = run_ragged(geoms, sweeps=1, use_graphs=True, local_energy=True, threads=2)
    print(f"  local energy of all geometries: {res['local_energy_walkers_per_s']:.0f} walkers/s", flush=True)
    for g, rw, el in zip(geoms, res["walkers"], res["e_loc"]):
        rr = rw[:2].cpu().numpy()
        e = el[:2].cpu().numpy()
        pt = fire.to_torch(g.params)
        for w in range(2):
            e_ref, lap = _oracle_energy(pt, g.cfg, rr[w])
            print(f"  N={g.cfg.n_el} walker {w}: dE {(e[w] - e_ref) * 1e3:+.3f} mHa (lap {lap:.0f})", flush=True)
            assert abs(e[w] - e_ref) < 2e-3 + 4e-6 * abs(lap)


def test_serving_real_shapes(gpu_device):
    from sparse_wf_amd.serving import Server, poisson_requests

    reqs = poisson_requests(3, qps=100.0, seed=23, n_el=(20, 400), walkers=(128, 160), sweeps=(10, 10),
                            precision="mixed", feature_dim=256)
    reqs[-1] = poisson_requests(1, qps=100.0, seed=29, n_el=(390, 400), walkers=(128, 128), sweeps=(10, 10),
                                precision="mixed", feature_dim=256)[0]
    reqs[-1].rid = 2
    print("\n[serving real shapes] N:", [q.cfg.n_el for q in reqs], "B:", [len(q.walkers) for q in reqs], flush=True)
    res = {r.rid: r for r in Server(device=gpu_device, max_active=3).run(reqs)}
    for q in reqs:
        r = res[q.rid]
        assert r.error is None, r.error
        e_ref, lap = _oracle_energy(fire.to_torch(q.params), q.cfg, r.walkers[0])
        print(f"  N={q.cfg.n_el}: E_loc[0] {r.e_loc[0]:.6f} vs {e_ref:.6f} (dE {(r.e_loc[0] - e_ref) * 1e3:+.3f} mHa, "
              f"lap {lap:.0f}), pmove {r.pmove:.3f}", flush=True)
        assert abs(r.e_loc[0] - e_ref) < 2e-3 + 4e-6 * abs(lap)
        assert 0.0 < r.pmove < 1.0
