"""Parameter uploads, the resident walker state and its error paths on the GPU.

* ``init`` leaves the final parameters (with both normalisation scales) on the device;
* a parameter tree changed in place is uploaded again before the next evaluation;
* an upload invalidates the resident state: the C ABI refuses proposals until the next
  state initialisation, and ``keep_state`` rebuilds the state from its own positions;
* ``refresh_sweeps`` reports the accept counts of the whole call;
* a move that changes more electrons than the state's capacity raises instead of returning
  the truncated update.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import fire
from sparse_wf_amd import _lib
from sparse_wf_amd.wave_function import FireWaveFunction, _dptr, _stream_ptr

from cases import make_params, small_cfg, walkers

pytestmark = pytest.mark.gpu


def test_init_then_signed_without_upload(gpu_device):
    cfg = small_cfg(False)
    r = walkers(cfg, 4, seed=5)
    wf = FireWaveFunction(cfg, device=gpu_device)
    prm = wf.init(3, r[0])
    assert float(prm["embedding"]["scales"][1]) != 1.0
    sg, lp = wf.signed(prm, r, precision="fp32")
    pt = fire.to_torch(prm)
    for w in range(len(r)):
        s_ref, l_ref = fire.signed(pt, cfg, r[w])
        assert float(sg[w]) == float(s_ref)
        assert abs(float(lp[w]) - float(l_ref)) < 1e-4 * max(1.0, abs(float(l_ref)))
    wf2 = FireWaveFunction(cfg, device=gpu_device)
    wf2.upload(prm)
    assert torch.equal(wf2.signed(prm, r, precision="fp32")[1], lp)


def test_in_place_change_is_uploaded(gpu_device):
    cfg = small_cfg(False)
    prm = make_params(cfg, seed=0)
    wf = FireWaveFunction(cfg, device=gpu_device)
    r = walkers(cfg, 3, seed=6)
    lp0 = wf.signed(prm, r, precision="fp32")[1].clone()
    prm["jastrow"]["params"]["mlp_scale"] = prm["jastrow"]["params"]["mlp_scale"] * 2.0
    lp1 = wf.signed(prm, r, precision="fp32")[1]
    assert not torch.equal(lp0, lp1)
    pt = fire.to_torch(prm)
    for w in range(len(r)):
        l_ref = float(fire.signed(pt, cfg, r[w])[1])
        assert abs(float(lp1[w]) - l_ref) < 1e-4 * max(1.0, abs(l_ref))


def test_upload_invalidates_state(gpu_device):
    cfg = small_cfg(False)
    prm = make_params(cfg, seed=0)
    prm2 = make_params(cfg, seed=1)
    wf = FireWaveFunction(cfg, device=gpu_device)
    r = walkers(cfg, 4, seed=7)
    (_, _), state = wf.log_psi_with_state(prm, r, precision="fp32")
    wf.upload(prm2)
    rn = r.copy()
    rn[:, 2] += 0.5
    with pytest.raises(ValueError):
        wf.log_psi_low_rank_update(prm2, rn, np.array([2]), state=state)
    # the C ABI itself refuses a proposal on the invalidated state
    rk = torch.as_tensor(rn[:, 2].copy(), device=wf.device)
    status = wf._lib.fire_state_propose(wf._h, _stream_ptr(), 2, _dptr(rk), None, None)
    assert status != 0 and b"no walker state" in _lib.load().fire_last_error()
    # keep_state after a parameter change rebuilds the state from its own positions
    (_, _), state = wf.log_psi_with_state(prm, r, precision="fp32")
    r_end, st = wf.mcmc_steps_low_rank(prm2, r, cfg.n_el, width=0.3, seed=3, keep_state=True, precision="fp32")
    lp_fresh = wf.signed(prm2, r_end, precision="fp32")[1]
    np.testing.assert_allclose(st["logpsi"].cpu().numpy(), lp_fresh.cpu().numpy(), rtol=1e-8, atol=1e-8)


def test_refresh_sweeps_counts_whole_call(gpu_device, monkeypatch):
    monkeypatch.setenv("FIRE_LOG_CAPTURE", "1")
    cfg = small_cfg(False)
    prm = make_params(cfg, seed=0)
    wf = FireWaveFunction(cfg, device=gpu_device)
    r = walkers(cfg, 16, seed=8)
    N = cfg.n_el
    r_a, st_a = wf.mcmc_steps_low_rank(prm, r, 3 * N, width=0.4, seed=5, refresh_sweeps=1, precision="fp32")
    cur, total, n_prop = r, 0, 0
    for q in range(3):
        cur, st = wf.mcmc_steps_low_rank(prm, cur, N, width=0.4, seed=5, step0=q * N, precision="fp32")
        cur = cur.cpu().numpy()
        total = total + st["n_accept"].cpu().numpy()
        n_prop += st["n_proposals"]
        print(f"  call {q}: {st['n_proposals']} proposals", flush=True)
    print(f"  refresh call: {st_a['n_proposals']} proposals", flush=True)
    np.testing.assert_array_equal(r_a.cpu().numpy(), cur)
    np.testing.assert_array_equal(st_a["n_accept"].cpu().numpy(), total)
    assert st_a["n_proposals"] == n_prop == 3 * N * len(r)
    assert abs(st_a["pmove"] - total.mean() / (3 * N)) < 1e-12


def test_capacity_overflow_raises(gpu_device, monkeypatch):
    cfg = small_cfg(False)
    prm = make_params(cfg, seed=0)
    wf = FireWaveFunction(cfg, device=gpu_device)
    r = walkers(cfg, 2, seed=9)
    monkeypatch.setenv("FIRE_STATE_CAP", "2")
    (_, _), state = wf.log_psi_with_state(prm, r, precision="fp32")
    rn = r.copy()
    rn[:, 0] = r[:, 1] + np.array([0.05, 0.0, 0.0], np.float32)  # next to electron 1: more than two change
    aff = [len(fire.affected_electrons(r[w], rn[w], 0, cfg.cutoff)) for w in range(2)]
    assert max(aff) > 2
    with pytest.raises(_lib.FireError, match="static capacity"):
        wf.log_psi_low_rank_update(prm, rn, np.array([0]), state=state)
    assert wf.state_stats()["max_changed"] == max(aff)
    monkeypatch.delenv("FIRE_STATE_CAP")
    (_, lp), state = wf.log_psi_with_state(prm, r, precision="fp32")
    (_, lp2), _ = wf.log_psi_low_rank_update(prm, rn, np.array([0]), state=state)
    assert torch.isfinite(lp2).all()
