"""Pins the oracle's Metropolis step (oracle/mcmc.py) and static-input counts
(oracle/fire.py static_input) on the CPU.

* Philox4x32-10 against the published known-answer vectors of Salmon et al. (SC'11; the
  Random123 distribution's kat_vectors);
* Box-Muller normals and the uniform variate: moments and a Kolmogorov-Smirnov test;
* proposal + accept rule: the chain's stationary radial distribution for the analytic
  |psi|^2 = e^{-2r} / pi (hydrogen 1s) by a chi-square test -- a wrong acceptance exponent
  (|psi| instead of |psi|^2) or a reversed comparison fails it;
* static_input (new_sparse_model.py:399-437) against the neighbour lists, the affected set of
  the low-rank update and brute-force pair loops.
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy import stats

from oracle import fire, mcmc
from cases import small_cfg, walkers

KAT = [  # (counter, key, result) for Philox4x32 with 10 rounds
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    got = mcmc.philox4x32_10(np.array(ctr, np.uint64), np.array(key, np.uint64))
    assert [int(v) for v in got] == out


def test_step_randoms_distribution():
    n3, u = mcmc.step_randoms(1234, np.arange(40000), 7)
    z = n3.reshape(-1)
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1.0) < 0.03
    assert stats.kstest(z, "norm").pvalue > 1e-4
    assert np.all((u > 0) & (u < 1)) and stats.kstest(u, "uniform").pvalue > 1e-4
    # streams differ between walkers, steps and seeds
    n3b, _ = mcmc.step_randoms(1234, np.arange(4), 8)
    n3c, _ = mcmc.step_randoms(99, np.arange(4), 7)
    assert not np.allclose(n3[:4], n3b) and not np.allclose(n3[:4], n3c)
    # a seed is folded to 32 bits as lo ^ hi
    a, _ = mcmc.step_randoms((5 << 32) | 3, np.arange(3), 1)
    b, _ = mcmc.step_randoms(5 ^ 3, np.arange(3), 1)
    np.testing.assert_array_equal(a, b)


def test_metropolis_stationary_hydrogen_1s():
    """Chains of the oracle's proposal and accept rule on log|psi| = -r sample |psi|^2: the
    radius follows Gamma(3, 1/2)."""
    B, width = 20000, 0.6
    r = np.random.default_rng(0).normal(size=(B, 1, 3)).astype(np.float32)
    lp = -np.linalg.norm(r[:, 0].astype(np.float64), axis=-1)
    for s in range(200):
        prop, u = mcmc.propose(r, 0, width, seed=11, step=s)
        lpn = -np.linalg.norm(prop[:, 0].astype(np.float64), axis=-1)
        acc = mcmc.accept(lp, lpn, u)
        r = np.where(acc[:, None, None], prop, r)
        lp = np.where(acc, lpn, lp)
    rad = np.linalg.norm(r[:, 0].astype(np.float64), axis=-1)
    edges = stats.gamma(3, scale=0.5).ppf(np.linspace(0, 1, 21))
    counts, _ = np.histogram(rad, edges)
    chi2 = ((counts - B / 20) ** 2 / (B / 20)).sum()
    assert stats.chi2(19).sf(chi2) > 1e-4, chi2
    # |psi| instead of |psi|^2 (a dropped factor 2) samples Gamma(3, 1): far off
    edges_wrong = stats.gamma(3, scale=1.0).ppf(np.linspace(0, 1, 21))
    cw, _ = np.histogram(rad, edges_wrong)
    assert stats.chi2(19).sf(((cw - B / 20) ** 2 / (B / 20)).sum()) < 1e-10


def test_static_input_counts():
    cfg = small_cfg(False)
    r = walkers(cfg, 3, seed=4)
    for w in range(3):
        st = fire.static_input(r[w], cfg.n_up, cfg.cutoff, cfg.mol.R, cfg.cutoff_1el)
        nl = fire.neighbour_lists(r[w], cfg.n_up, cfg.cutoff)
        assert st["n_pairs_same"] == int(nl.same.sum()) and st["n_pairs_diff"] == int((~nl.same).sum())
        deg = np.diff(nl.offsets)
        assert st["n_triplets"] == int((deg**2).sum()) == fire.n_triplets(r[w], cfg.cutoff)
        assert st["n_changed_hout"] == 0 and st["n_pp_elecs"] == cfg.n_el
        for k in (0, cfg.n_up, cfg.n_el - 1):
            rn = r[w].copy()
            rn[k] += np.array([1.7, -0.4, 0.9], np.float32)
            st2 = fire.static_input(r[w], cfg.n_up, cfg.cutoff, cfg.mol.R, cfg.cutoff_1el, rn, k)
            aff = fire.affected_electrons(r[w], rn, k, cfg.cutoff)
            assert st2["n_changed_hout"] == len(aff)
            nl2 = fire.neighbour_lists(rn, cfg.n_up, cfg.cutoff)
            same = diff = 0
            for i in aff:
                for j in nl2.nbr[nl2.offsets[i]:nl2.offsets[i + 1]]:
                    if (i < cfg.n_up) == (j < cfg.n_up):
                        same += 1
                    else:
                        diff += 1
            assert (st2["n_changed_pair_same"], st2["n_changed_pair_diff"]) == (same, diff)


def test_static_input_nuclei_within_cutoff_1el():
    cfg = small_cfg(False, cutoff_1el=3.0)
    r = walkers(cfg, 1, seed=9)[0]
    st = fire.static_input(r, cfg.n_up, cfg.cutoff, cfg.mol.R, cfg.cutoff_1el)
    R = np.asarray(cfg.mol.R, np.float64)
    cnt = [sum(np.linalg.norm(r[i].astype(np.float64) - R[J]) < 3.0 for J in range(len(R))) for i in range(len(r))]
    assert st["n_neighbours_en"] == max(cnt)
