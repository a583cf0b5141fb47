"""Pseudopotential tables and the oracle's energy terms, pinned by closed forms:

* the spherical quadratures are designs: the average of P_l(d . g) over the points vanishes for
  every direction d up to the design's order (tetrahedron 2, octahedron 3, icosahedron 5,
  24 points 7);
* eval_ecp against the closed form of its Gaussian sums;
* for a wavefunction spherically symmetric about the ECP nucleus every psi ratio is 1 and the
  non-local energy is V_0(r) exactly; for psi = z (a p orbital about the nucleus) the
  addition theorem leaves V_1(r) -- a wrong (2l+1) factor, Legendre polynomial, ratio
  direction, weight or rotation fails one of them.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

from oracle import pseudopotentials as O
from sparse_wf_amd import pseudopotentials as P
from tests.ecp_cases import CARBON_ECP


@pytest.mark.parametrize("n,order", [(4, 2), (6, 3), (12, 5), (24, 7)])
def test_quadratures_are_designs(n, order):
    g, w = P.make_spherical_grid(n)
    assert abs(w.sum() - 1) < 1e-7
    rng = np.random.default_rng(n)
    for _ in range(5):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        x = g.astype(np.float64) @ d
        for l in range(1, order + 1):
            c = np.zeros(l + 1)
            c[l] = 1
            assert abs(float((w * npleg.legval(x, c)).sum())) < 2e-6, (n, l)
        c = np.zeros(order + 2)
        c[order + 1] = 1
        if n in (4, 6, 12):  # one order above the design is not integrated
            assert max(abs(float((w * npleg.legval(g.astype(np.float64) @ e, c)).sum()))
                       for e in np.eye(3)) > 1e-3


def test_eval_ecp_closed_form():
    coeffs = [[(1.0, 2.0)], [(0.5, 3.0)], [(0.25, -1.0)]]
    r = np.array([0.5, 1.0, 2.0])
    ref = 2 * np.exp(-r ** 2) + 3 * r * np.exp(-0.5 * r ** 2) - r ** 2 * np.exp(-0.25 * r ** 2)
    np.testing.assert_allclose(P.eval_ecp(r, coeffs), ref, rtol=1e-14)


def test_tables_of_a_molecule():
    Z = np.array([6.0, 1.0, 1.0, 6.0])
    pp = P.make_pseudopotential(Z, {"C": CARBON_ECP})
    np.testing.assert_array_equal(pp.effective_charges, [4, 1, 1, 4])
    np.testing.assert_array_equal(pp.atoms, [0, 3])
    np.testing.assert_array_equal(pp.n_quad, [4, 4])  # config/default.yaml: C uses 4 points
    g = pp.grid
    assert g.size == 10001 and abs(g[0] - 1e-5) < 1e-18 and abs(g[-1] - 1e10) < 1e-2
    i = 5000
    loc = (4.0 * g[i] * np.exp(-14.43502 * g[i] ** 2) - 25.81955 * g[i] ** 2 * np.exp(-7.38188 * g[i] ** 2)
           + 57.74008 * g[i] ** 3 * np.exp(-8.39889 * g[i] ** 2)) / g[i] ** 2
    assert abs(pp.v_loc[0, i] - loc) <= 1e-12 * abs(loc)
    assert pp.v_nl.shape == (2, 2, 10001)
    # cutoff: the last grid radius where a non-local channel exceeds 1e-7 in magnitude
    big = (np.abs(pp.v_nl[0]) > 1e-7).any(0)
    assert pp.cutoffs[0] == g[np.nonzero(big)[0].max()]
    assert 1.2 < pp.cutoffs[0] < 2.5
    assert P.make_pseudopotential(Z, {}).n_ecp == 0


def _single_nucleus_tables(n_quad=12):
    grid = np.logspace(-5, 10, 10001)
    v = np.stack([(l + 1.0) * np.exp(-(0.5 + 0.3 * l) * grid ** 2) for l in range(4)])
    dirs, _ = P.make_spherical_grid(n_quad)
    return {"atoms": np.array([0]), "grid": grid, "v_loc": -np.exp(-grid)[None], "v_nl": v[None], "cutoffs": np.array([4.0]),
            "n_quad": np.array([n_quad]), "quad_dirs": dirs[None].astype(np.float64)}


def test_rotation_is_proper():
    for w in range(3):
        for i in range(3):
            m = O.rotation(7, w, i)
            np.testing.assert_allclose(m @ m.T, np.eye(3), atol=1e-14)
            assert abs(np.linalg.det(m) - 1) < 1e-14


def test_spherical_psi_gives_v0():
    tab = _single_nucleus_tables()
    R = np.zeros((1, 3), np.float32)
    r = np.array([[0.7, -0.4, 0.9]], np.float32)
    d = float(np.linalg.norm(r[0].astype(np.float64)))
    v, n = O.nonlocal_energy_pp(lambda x: (1.0, -float(np.linalg.norm(x[0].astype(np.float64)))), r, R, tab, 3, 0)
    assert n == 12
    assert abs(v - 1.0 * math.exp(-0.5 * d * d)) < 1e-6


def test_p_orbital_gives_v1():
    tab = _single_nucleus_tables()
    R = np.zeros((1, 3), np.float32)
    for seed in range(3):
        r = np.array([[0.3, 0.5, 0.8]], np.float32)
        d = float(np.linalg.norm(r[0].astype(np.float64)))
        sz = lambda x: (float(np.sign(x[0, 2])), math.log(abs(float(x[0, 2]))))
        v, _ = O.nonlocal_energy_pp(sz, r, R, tab, seed, 0)
        assert abs(v - 2.0 * math.exp(-0.8 * d * d)) < 2e-6, (seed, v)


def test_local_part_at_grid_radius():
    tab = _single_nucleus_tables()
    g = tab["grid"][3500]
    r = np.array([[np.float32(g), 0, 0]], np.float32)
    x = float(np.float32(g))
    ref = float(np.interp(x, tab["grid"], tab["v_loc"][0]))
    assert abs(O.local_energy_pp(r, np.zeros((1, 3), np.float32), tab) - ref) < 1e-15
    assert abs(ref + math.exp(-x)) < 1e-6 * math.exp(-x)
