"""Resident walker state, single-electron low-rank updates and Metropolis steps on the GPU.

Mirrors the reference's tests/test_low_rank.py (electron 3 moved by +2 in x, twice, the
update compared with a fresh evaluation of the moved walkers) and adds partial commits and
the Metropolis sweep.  The updated embedding / orbital / Jastrow rows are computed by the same
kernels as a fresh evaluation, so the update and the fresh GPU evaluation differ only by the
FP64 Woodbury determinant update versus a fresh FP64 inverse: tolerance 1e-9 |log psi|
(FP32 contractions; 1e-7 for split-BF16, whose per-row results do not depend on the batch
either).  Against the FP64 oracle's log_psi_low_rank_update the FP32 tolerance of the value
path applies (1e-4 relative).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fire
from sparse_wf_amd.wave_function import FireWaveFunction

from cases import make_params, small_cfg, walkers, water_cfg

pytestmark = pytest.mark.gpu

CASES = {
    "small": lambda: small_cfg(False),
    "small_restricted": lambda: small_cfg(True),
    "water": lambda: water_cfg(),
}
UPDATE_TOL = {"fp32": 1e-9, "mixed": 1e-7}


@pytest.fixture(scope="module", params=list(CASES))
def setup(request, gpu_device):
    cfg = CASES[request.param]()
    prm = make_params(cfg, seed=0)
    wf = FireWaveFunction(cfg, device=gpu_device)
    wf.upload(prm)
    r = walkers(cfg, 4, seed=11)
    print(f"\n[low-rank case {request.param}] N={cfg.n_el} D={cfg.n_determinants}", flush=True)
    return request.param, cfg, prm, wf, r


def _close(a, b, tol):
    a = np.asarray(torch.as_tensor(a).cpu(), np.float64)
    b = np.asarray(torch.as_tensor(b).cpu(), np.float64)
    assert np.all(np.abs(a - b) <= tol * np.maximum(1.0, np.abs(b))), (a, b)


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_state_init_matches_signed(setup, precision):
    name, cfg, prm, wf, r = setup
    (sg, lp), st = wf.log_psi_with_state(prm, r, precision=precision)
    sg_f, lp_f = wf.signed(prm, r, precision=precision)
    _close(lp, lp_f, UPDATE_TOL[precision])
    assert torch.equal(sg.cpu(), sg_f.cpu())


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_low_rank_update_logpsi(setup, precision):
    """tests/test_low_rank.py:29-62: electrons_new = electrons + [2, 0, 0] on electron 3,
    evaluated twice from the updated state (the second update moves it to the same place)."""
    name, cfg, prm, wf, r = setup
    (_, lp_old), state = wf.log_psi_with_state(prm, r, precision=precision)
    for step in range(2):
        r_new = r.copy()
        r_new[:, 3] += np.array([2.0, 0.0, 0.0], np.float32)
        sg_full, lp_full = wf.signed(prm, r_new, precision=precision)
        (sg_upd, lp_upd), state = wf.log_psi_low_rank_update(prm, r_new, np.array([3]), state=state)
        print(f"  step {step}: |dlogpsi| max {float((lp_upd - lp_full).abs().max()):.3e}", flush=True)
        _close(lp_upd, lp_full, UPDATE_TOL[precision])
        assert torch.equal(sg_upd.cpu(), sg_full.cpu())


def test_low_rank_matches_oracle(setup):
    """The GPU update against the FP64 oracle's log_psi_low_rank_update on one walker."""
    name, cfg, prm, wf, r = setup
    pt = fire.to_torch(prm)
    (_, _), state = wf.log_psi_with_state(prm, r, precision="fp32")
    r_new = r.copy()
    r_new[:, 3] += np.array([2.0, 0.0, 0.0], np.float32)
    (sg_upd, lp_upd), state = wf.log_psi_low_rank_update(prm, r_new, np.array([3]), state=state)
    for w in range(2):
        (_, _), ost = fire.log_psi_with_state(pt, cfg, r[w])
        (s_ref, l_ref), _, aff = fire.log_psi_low_rank_update(pt, cfg, r_new[w], 3, ost)
        d = abs(float(lp_upd[w]) - float(l_ref))
        assert d < 1e-4 * max(1.0, abs(float(l_ref))), (w, d)
        assert float(sg_upd[w]) == float(np.sign(float(s_ref)))


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_successive_moves_of_both_spins(setup, precision):
    """A chain of moves of up- and down-spin electrons, every one taken, against fresh
    evaluations of the moved walkers."""
    name, cfg, prm, wf, r = setup
    rng = np.random.default_rng(7)
    (_, _), state = wf.log_psi_with_state(prm, r, precision=precision)
    cur = r.copy()
    moves = [0, cfg.n_up, cfg.n_el - 1, 3, 0]
    for k in moves:
        cur = cur.copy()
        cur[:, k] += rng.normal(scale=0.7, size=(len(cur), 3)).astype(np.float32)
        (sg_upd, lp_upd), state = wf.log_psi_low_rank_update(prm, cur, np.array([k]), state=state)
        sg_full, lp_full = wf.signed(prm, cur, precision=precision)
        _close(lp_upd, lp_full, 10 * UPDATE_TOL[precision])
        assert torch.equal(sg_upd.cpu(), sg_full.cpu())
    wf.commit(state)
    r_state, sg_state, lp_state = wf.state_values()
    np.testing.assert_array_equal(r_state.cpu().numpy(), cur)


def test_partial_commit(setup):
    """Only accepted walkers take the proposal; the rest keep positions and log psi, and a
    following update from the mixed state still matches a fresh evaluation."""
    name, cfg, prm, wf, r = setup
    (_, lp0), state = wf.log_psi_with_state(prm, r, precision="fp32")
    r1 = r.copy()
    r1[:, 5] += np.array([0.3, -1.1, 0.6], np.float32)
    (_, lp1), prop = wf.log_psi_low_rank_update(prm, r1, np.array([5]), state=state)
    acc = np.array([1, 0, 1, 0][: len(r)], np.int32)
    wf.commit(prop, acc)
    r_state, _, lp_state = wf.state_values()
    expect_r = np.where(acc[:, None, None] == 1, r1, r)
    np.testing.assert_array_equal(r_state.cpu().numpy(), expect_r)
    expect_lp = torch.where(torch.as_tensor(acc, device=lp1.device) == 1, lp1, lp0)
    assert torch.equal(lp_state, expect_lp)
    r2 = expect_r.copy()
    r2[:, 2] += np.array([-0.4, 0.2, 0.9], np.float32)
    (sg2, lp2), _ = wf.log_psi_low_rank_update(prm, r2, np.array([2]), state=prop)
    sg_f, lp_f = wf.signed(prm, r2, precision="fp32")
    _close(lp2, lp_f, UPDATE_TOL["fp32"])
    assert torch.equal(sg2.cpu(), sg_f.cpu())


def test_mcmc_steps(setup):
    """Metropolis sweep: reproducible for a seed, independent of how the steps are split over
    calls, acceptance strictly between 0 and 1, walkers without accepted moves unchanged,
    and the resident log psi equal to a fresh evaluation of the final walkers."""
    name, cfg, prm, wf, r = setup
    r = walkers(cfg, 16, seed=12)
    steps = 2 * cfg.n_el
    r_a, st_a = wf.mcmc_steps_low_rank(prm, r, steps, width=0.3, seed=1234, precision="fp32")
    r_b, st_b = wf.mcmc_steps_low_rank(prm, r, steps // 2, width=0.3, seed=1234, precision="fp32")
    r_b, st_b2 = wf.mcmc_steps_low_rank(prm, r_b, steps - steps // 2, width=0.3, seed=1234,
                                        step0=steps // 2, keep_state=True)
    assert torch.equal(r_a, r_b)
    assert 0.0 < st_a["pmove"] < 1.0, st_a["pmove"]
    print(f"  pmove {st_a['pmove']:.3f} mean changed {st_a['mean_changed']:.2f} max {st_a['max_changed']}",
          flush=True)
    moved = (r_a.cpu().numpy() != r).any(axis=(1, 2))
    assert np.array_equal(moved, st_a["n_accept"].cpu().numpy() > 0)
    sg_f, lp_f = wf.signed(prm, r_a, precision="fp32")
    _close(st_a["logpsi"], lp_f, 1e-8)
    assert torch.equal(st_a["sign"].cpu(), sg_f.cpu())
    r_c, _ = wf.mcmc_steps_low_rank(prm, r, steps, width=0.3, seed=99, precision="fp32")
    assert not torch.equal(r_a, r_c)


def test_mcmc_acceptance_vs_width(gpu_device):
    """Acceptance falls as the step width grows (Metropolis with symmetric proposals)."""
    cfg = small_cfg(False)
    prm = make_params(cfg, seed=0)
    wf = FireWaveFunction(cfg, device=gpu_device)
    wf.upload(prm)
    r = walkers(cfg, 32, seed=21)
    rates = []
    for width in (0.05, 0.5, 3.0):
        _, st = wf.mcmc_steps_low_rank(prm, r, cfg.n_el, width=width, seed=5, precision="fp32")
        rates.append(st["pmove"])
    assert rates[0] > rates[1] > rates[2], rates


def test_alkane_updates_and_sweep(gpu_device):
    """C16H34 (130 electrons, cutoff 5, F = 256, 4 determinants): tens of electrons change per
    move, so the Woodbury kernels run with more than 64 KB of LDS.  Updates against fresh
    evaluations, then one Metropolis sweep with the resident log psi checked at the end."""
    from sparse_wf_amd import systems
    from sparse_wf_amd.config import FireConfig

    mol, over, _, sigma = systems.workload("c16h34")
    cfg = FireConfig(mol=mol, **over)
    wf = FireWaveFunction(cfg, device=gpu_device)
    r = systems.init_electrons(mol, 8, stddev=sigma, seed=4)
    prm = wf.init(0, r[0])
    wf.upload(prm)
    (_, _), state = wf.log_psi_with_state(prm, r, precision="fp32")
    cur = r.copy()
    rng = np.random.default_rng(3)
    for k in (0, 64, 65, 129):
        cur = cur.copy()
        cur[:, k] += rng.normal(scale=0.3, size=(len(cur), 3)).astype(np.float32)
        (sg_u, lp_u), state = wf.log_psi_low_rank_update(prm, cur, np.array([k]), state=state)
        sg_f, lp_f = wf.signed(prm, cur, precision="fp32")
        _close(lp_u, lp_f, 1e-8)
        assert torch.equal(sg_u.cpu(), sg_f.cpu())
    st = wf.state_stats()
    print(f"  mean changed {st['mean_changed']:.1f} max {st['max_changed']}", flush=True)
    assert st["max_changed"] > 10
    r_end, stm = wf.mcmc_steps_low_rank(prm, r, cfg.n_el, width=0.3, seed=2, precision="fp32")
    sg_f, lp_f = wf.signed(prm, r_end, precision="fp32")
    _close(stm["logpsi"], lp_f, 1e-8)
    assert 0.0 < stm["pmove"] < 1.0


def test_many_changed_electrons(gpu_device):
    """64 hydrogen atoms on a 4 x 4 x 4 cube of side 4.8 bohr with an 8 bohr cutoff: a move
    changes most of the 64 electrons, which runs the four-wave Gauss-Jordan elimination and
    the 3- and 4-tile classes of the inverse update (the capacity is all 64 electrons, so
    nothing can overflow).  Updates against fresh evaluations, then a sweep."""
    import itertools

    from sparse_wf_amd import systems
    from sparse_wf_amd.config import FireConfig

    R = 1.6 * np.array(list(itertools.product(range(4), repeat=3)), np.float64)
    mol = systems.make_molecule(R, np.ones(len(R), np.int64), name="h64_cube")
    cfg = FireConfig(mol=mol, cutoff=8.0, feature_dim=64, n_determinants=2)
    wf = FireWaveFunction(cfg, device=gpu_device)
    r = systems.init_electrons(mol, 8, stddev=0.5, seed=6)
    prm = wf.init(0, r[0])
    wf.upload(prm)
    (_, _), state = wf.log_psi_with_state(prm, r, precision="fp32")
    cur = r.copy()
    rng = np.random.default_rng(8)
    for k in (0, 17, 30, 45, 59):
        cur = cur.copy()
        cur[:, k] += rng.normal(scale=0.3, size=(len(cur), 3)).astype(np.float32)
        (sg_u, lp_u), state = wf.log_psi_low_rank_update(prm, cur, np.array([k]), state=state)
        sg_f, lp_f = wf.signed(prm, cur, precision="fp32")
        _close(lp_u, lp_f, 1e-8)
        assert torch.equal(sg_u.cpu(), sg_f.cpu())
    st = wf.state_stats()
    print(f"  mean changed {st['mean_changed']:.1f} max {st['max_changed']}", flush=True)
    assert st["mean_changed"] > 32 and st["max_changed"] > 48
    r_end, stm = wf.mcmc_steps_low_rank(prm, r, cfg.n_el, width=0.3, seed=3, precision="fp32")
    sg_f, lp_f = wf.signed(prm, r_end, precision="fp32")
    _close(stm["logpsi"], lp_f, 1e-8)
    assert 0.0 < stm["pmove"] < 1.0


def test_large_n_generic_update(gpu_device):
    """200 electrons (hydrogen chain, 1.8 bohr): more than 12 16-wide column tiles, so the
    low-rank update takes the generic Woodbury kernel and the state's inverses the non-register
    Gauss-Jordan / LU forms.  Updates against fresh evaluations, then a sweep."""
    from sparse_wf_amd import systems
    from sparse_wf_amd.config import FireConfig

    mol = systems.h_chain(200, 1.8)
    cfg = FireConfig(mol=mol, cutoff=5.0, feature_dim=64, n_determinants=1)
    wf = FireWaveFunction(cfg, device=gpu_device)
    r = systems.init_electrons(mol, 4, stddev=0.5, seed=12)
    prm = wf.init(0, r[0])
    wf.upload(prm)
    (_, _), state = wf.log_psi_with_state(prm, r, precision="fp32")
    cur = r.copy()
    rng = np.random.default_rng(5)
    for k in (0, 99, 100, 199):
        cur = cur.copy()
        cur[:, k] += rng.normal(scale=0.3, size=(len(cur), 3)).astype(np.float32)
        (sg_u, lp_u), state = wf.log_psi_low_rank_update(prm, cur, np.array([k]), state=state)
        sg_f, lp_f = wf.signed(prm, cur, precision="fp32")
        _close(lp_u, lp_f, 1e-8)
        assert torch.equal(sg_u.cpu(), sg_f.cpu())
    r_end, stm = wf.mcmc_steps_low_rank(prm, r, cfg.n_el, width=0.3, seed=4, precision="fp32")
    sg_f, lp_f = wf.signed(prm, r_end, precision="fp32")
    _close(stm["logpsi"], lp_f, 1e-8)
    assert 0.0 < stm["pmove"] < 1.0


def test_sweep_graph_matches_eager(setup):
    """Whole sweeps replayed from the captured hipGraph give bit-identical walkers, log psi
    and acceptance counts to the same steps issued eagerly, over several replays with eager
    work in between, and a run split at a sweep boundary equals one call."""
    name, cfg, prm, wf, r = setup
    r = walkers(cfg, 16, seed=31)
    N = cfg.n_el
    results = {}
    for use_graphs in (False, True):
        wf.mcmc_options(use_graphs)
        out = []
        (_, _), st = wf.log_psi_with_state(prm, r, precision="mixed")
        for rep in range(3):
            rr, s = wf.mcmc_steps_low_rank(prm, r, 2 * N, width=0.3, seed=77, step0=2 * N * rep, keep_state=True)
            wf.signed(prm, rr, precision="fp32")  # eager work between replays
            out.append((rr.clone(), s["logpsi"].clone(), s["n_accept"].clone()))
        results[use_graphs] = out
    wf.mcmc_options(True)
    for (ra, la, na), (rb, lb, nb) in zip(results[False], results[True]):
        assert torch.equal(ra, rb)
        assert torch.equal(la, lb)
        assert torch.equal(na, nb)
    # split at a sweep boundary plus a partial sweep: same as one call
    wf.log_psi_with_state(prm, r, precision="mixed")
    r1, _ = wf.mcmc_steps_low_rank(prm, r, N + 3, width=0.3, seed=5, keep_state=True)
    wf.log_psi_with_state(prm, r, precision="mixed")
    r2, _ = wf.mcmc_steps_low_rank(prm, r, N, width=0.3, seed=5, keep_state=True)
    r2, _ = wf.mcmc_steps_low_rank(prm, r2, 3, width=0.3, seed=5, step0=N, keep_state=True)
    assert torch.equal(r1, r2)


def test_refresh_keeps_graph_and_results(setup):
    """Rebuilding the resident state every sweep (refresh_sweeps=1) reuses the captured sweep
    (the state buffers stay in place) and gives bit-identical results to the same steps issued
    eagerly with the same rebuilds; the rebuilt log psi equals a fresh evaluation."""
    name, cfg, prm, wf, r = setup
    r = walkers(cfg, 16, seed=41)
    N = cfg.n_el
    out = {}
    for use_graphs in (False, True):
        wf.mcmc_options(use_graphs)
        rr, s = wf.mcmc_steps_low_rank(prm, r, 3 * N, width=0.3, seed=9, precision="mixed", refresh_sweeps=1)
        out[use_graphs] = (rr.clone(), s["logpsi"].clone(), s["n_accept"].clone())
    wf.mcmc_options(True)
    for a, b in zip(out[False], out[True]):
        assert torch.equal(a, b)
    _, lp_f = wf.signed(prm, out[True][0], precision="mixed")
    _close(out[True][1], lp_f, 1e-8)
