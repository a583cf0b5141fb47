"""Effective-core pseudopotentials for the local energy (host-side tables; the energy terms run
on the GPU in ``fire_ecp_energy``).

Mirrors src/sparse_wf/pseudopotentials.py: the ECP of each element is given in pyscf's layout
``(n_core, [(l, [[(exponent, coefficient), ...] for r^0, r^1, r^2 ...]), ...])`` with l = -1 the
local channel (the reference loads it with ``pyscf.gto.basis.load_ecp("ccecp", symbol)``;
pyscf is not a dependency here, so the caller passes the coefficients).  The channels are
evaluated on the reference's radial grid (``logspace(-5, 10, 10001)``) in FP64,
``r^2 U_l(r) = sum_k A_lk r^n_lk exp(-B_lk r^2)`` divided by r^2 (eval_ecp / eval_ecp_on_grid
:136-202); nuclei with an ECP carry the effective charge Z - n_core in the Coulomb terms; the
non-local channels are neglected beyond the last grid radius where any exceeds
``cutoff_value`` in magnitude (:448-451); each element's spherical quadrature has
``n_quad_points[symbol]`` points (``default`` otherwise; config/default.yaml:86-99).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Mapping, Sequence

import numpy as np

SYMBOLS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar",
           "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"]

DEFAULT_QUAD_POINTS = {"default": 12, "C": 4, "N": 4, "O": 4, "F": 4, "Ne": 4, "Na": 6, "Mg": 6, "Al": 6, "Si": 6,
                       "P": 6, "S": 6, "Cl": 6, "Ar": 6}


def _sph_to_cart(theta_phi) -> np.ndarray:
    t, p = np.asarray(theta_phi, np.float64).T
    return np.stack([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)], -1).astype(np.float32)


def make_spherical_grid(n_points: int):
    """Quadrature directions and weights on the unit sphere (make_spherical_grid :40-130):
    poles, tetrahedron, octahedron, the 12-point icosahedral 5-design and the 24-point 7-design
    of Sloane's tables; equal weights.  Directions in float32 as the reference builds them."""
    if n_points == 1:
        r = _sph_to_cart([[0, 0]])
    elif n_points == 2:
        r = _sph_to_cart([[0, 0], [np.pi, 0]])
    elif n_points == 4:
        t = np.arccos(-1 / 3)
        r = _sph_to_cart([[0, 0], [t, 0], [t, 2 * np.pi / 3], [t, 4 * np.pi / 3]])
    elif n_points == 6:
        r = _sph_to_cart([[0, 0], [np.pi, 0], [np.pi / 2, 0], [np.pi / 2, np.pi / 2], [np.pi / 2, np.pi],
                          [np.pi / 2, 3 * np.pi / 2]])
    elif n_points == 12:
        x, y = 0.850650808352, 0.525731112119
        r = np.array([[x, 0, -y], [y, -x, 0], [0, -y, x], [x, 0, y], [-y, -x, 0], [0, y, -x], [-x, 0, -y],
                      [-y, x, 0], [0, y, x], [-x, 0, y], [y, x, 0], [0, -y, -x]], np.float32)
    elif n_points == 24:
        a, b, c = 0.8662468181078205913835980, 0.4225186537611115291185464, 0.2666354015167047203315344
        r = np.array([[a, b, c], [a, -b, -c], [a, c, -b], [a, -c, b], [-a, b, -c], [-a, -b, c], [-a, c, b],
                      [-a, -c, -b], [c, a, b], [-c, a, -b], [-b, a, c], [b, a, -c], [-c, -a, b], [c, -a, -b],
                      [b, -a, c], [-b, -a, -c], [b, c, a], [-b, -c, a], [c, -b, a], [-c, b, a], [b, -c, -a],
                      [-b, c, -a], [c, b, -a], [-c, -b, -a]], np.float32)
    else:
        raise ValueError(f"Unsupported number of points: {n_points}")
    assert np.allclose(np.linalg.norm(r, axis=1), 1)
    return r, np.full(n_points, 1.0 / n_points, np.float32)


def eval_ecp(radius, coeffs) -> np.ndarray:
    """r^2 U_l(r) = sum_k A_lk r^n_lk exp(-B_lk r^2) (eval_ecp :136-142), FP64."""
    radius = np.asarray(radius, np.float64)
    val = np.zeros_like(radius)
    for r_exponent, terms in enumerate(coeffs):
        for exp_coeff, linear_coeff in terms:
            val = val + linear_coeff * radius ** r_exponent * np.exp(-exp_coeff * radius ** 2)
    return val


def eval_ecp_on_grid(ecp_all: Mapping[int, tuple], r_grid=None, log_r0: float = -5, log_rn: float = 10,
                     n_grid: int = 10001):
    """eval_ecp_on_grid (:145-202): (n_cores, {Z: [max_channels, n_grid] values}, grid,
    max_channels); the local channel (l = -1) is the last row."""
    if r_grid is None:
        r_grid = np.logspace(log_r0, log_rn, n_grid)
    r_grid = np.asarray(r_grid, np.float64)
    max_channels = max(len(v[1]) for v in ecp_all.values())
    n_cores = {z: v[0] for z, v in ecp_all.items()}
    v_grid = {}
    for z, (_, channels) in ecp_all.items():
        g = np.zeros((max_channels, r_grid.size))
        for angular, coeffs in channels:
            g[angular] = eval_ecp(r_grid, coeffs) / r_grid ** 2
        v_grid[z] = g
    return n_cores, v_grid, r_grid, max_channels


@dataclass
class Pseudopotential:
    """Device-ready tables of the pseudopotentials of one molecule."""

    effective_charges: np.ndarray  # [n_nuc]
    atoms: np.ndarray              # [n_ecp] nucleus indices
    grid: np.ndarray               # [n_grid]
    v_loc: np.ndarray              # [n_ecp, n_grid]
    v_nl: np.ndarray               # [n_ecp, n_l, n_grid]
    cutoffs: np.ndarray            # [n_ecp]
    n_quad: np.ndarray             # [n_ecp]
    quad_dirs: np.ndarray          # [n_ecp, q_max, 3]

    @property
    def n_ecp(self) -> int:
        return int(self.atoms.size)


def make_pseudopotential(charges: Sequence[float], ecp_data: Mapping, n_quad_points: Mapping | None = None,
                         cutoff_value: float = 1e-7) -> Pseudopotential:
    """make_pseudopotential (:411-454) with the ECP coefficients given: ``ecp_data`` maps an
    element (symbol or atomic number) to its ECP in pyscf's layout."""
    nq = dict(DEFAULT_QUAD_POINTS if n_quad_points is None else n_quad_points)
    data = {}
    for k, v in ecp_data.items():
        z = SYMBOLS.index(k.capitalize()) if isinstance(k, str) else int(k)
        data[z] = v
    charges = np.asarray(charges, np.float64)
    if not data:
        return Pseudopotential(charges.astype(np.float32), np.zeros(0, np.int32), np.zeros(2), np.zeros((0, 2)),
                               np.zeros((0, 0, 2)), np.zeros(0), np.zeros(0, np.int32), np.zeros((0, 1, 3)))
    n_cores, v_grid, grid, max_channels = eval_ecp_on_grid(data)
    eff = np.asarray([z - n_cores.get(int(round(z)), 0) for z in charges], np.float64)
    mask = np.abs(charges - eff) > 2.0e-6
    atoms = np.nonzero(mask)[0].astype(np.int32)
    vals = np.stack([v_grid[int(round(charges[a]))] for a in atoms]) if atoms.size else np.zeros((0, max_channels, grid.size))
    v_nl, v_loc = vals[:, : max_channels - 1], vals[:, max_channels - 1]
    larger = np.abs(v_nl) > cutoff_value
    cutoff_idx = np.max(larger * np.arange(grid.size), axis=(1, 2)) if v_nl.shape[1] else np.zeros(atoms.size, int)
    cutoffs = grid[cutoff_idx]
    n_quad = np.asarray([nq.get(SYMBOLS[int(round(charges[a]))], nq["default"]) for a in atoms], np.int32)
    q_max = int(n_quad.max()) if atoms.size else 1
    dirs = np.zeros((atoms.size, q_max, 3))
    for e, n in enumerate(n_quad):
        dirs[e, :n] = make_spherical_grid(int(n))[0]
    return Pseudopotential(eff.astype(np.float32), atoms, grid, np.ascontiguousarray(v_loc),
                           np.ascontiguousarray(v_nl), cutoffs, n_quad, dirs)


def upload(lib, handle, pp: Pseudopotential):
    """``fire_ecp_upload`` of the tables (and of the effective charges)."""
    from . import _lib

    c = lambda a, t: np.ascontiguousarray(a, t)
    atoms, grid, vloc = c(pp.atoms, np.int32), c(pp.grid, np.float64), c(pp.v_loc, np.float64)
    vnl, cut, nq = c(pp.v_nl, np.float64), c(pp.cutoffs, np.float64), c(pp.n_quad, np.int32)
    dirs, zeff = c(pp.quad_dirs, np.float64), c(pp.effective_charges, np.float32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    _lib.check(lib.fire_ecp_upload(handle, pp.n_ecp, p(atoms), int(vnl.shape[1]) if vnl.ndim == 3 else 0,
                                   int(grid.size), p(grid), p(vloc), p(vnl), p(cut), p(nq), int(dirs.shape[1]),
                                   p(dirs), p(zeff)))
