"""Synthetic molecules and walker initialisation.

Geometry builders follow the reference's conventions: ``chain`` places atoms at
``i * distance`` along x in bohr (src/sparse_wf/system.py:7-11); electrons are drawn as
N(nucleus, stddev) with the spin-to-atom assignment of ``assign_spins_to_atoms``
(src/sparse_wf/mcmc.py:381-428) and ``init_electrons`` (mcmc.py:431-465).  The reference
draws with ``jax.random``; this module uses NumPy's PCG64, so the walkers are synthetic
inputs of the same distribution, not the same numbers.
"""
from __future__ import annotations

import math

import numpy as np

from .config import Molecule

ELEMENTS = {"H": 1, "He": 2, "Li": 3, "C": 6, "N": 7, "O": 8}


def make_molecule(R, Z, charge: int = 0, spin: int | None = None, name: str = "") -> Molecule:
    R = np.asarray(R, dtype=np.float32).reshape(-1, 3)
    Z = np.asarray(Z, dtype=np.int32).reshape(-1)
    n_el = int(Z.sum()) - int(charge)
    if spin is None:
        spin = n_el % 2
    n_up = (n_el + spin) // 2
    return Molecule(R=R, Z=Z, n_el=n_el, n_up=n_up, name=name)


def chain(element: str, distance: float, n: int) -> Molecule:
    """Linear chain along x (system.py:7-11)."""
    R = np.zeros((n, 3), np.float32)
    R[:, 0] = np.arange(n) * distance
    return make_molecule(R, [ELEMENTS[element]] * n, name=f"{element}{n}_chain")


def atom_chain(n_nuc: int, Z: int) -> Molecule:
    """The reference test fixture's chain: ``n_nuc`` atoms of charge Z at 1 bohr spacing
    (tests/utils.py:14-18)."""
    R = np.arange(n_nuc)[:, None] * np.array([1.0, 0.0, 0.0])
    return make_molecule(R, [Z] * n_nuc, name=f"Z{Z}x{n_nuc}_chain")


def water_like() -> Molecule:
    """Z=[8,1,1] at (0,0,0), (1.81,0,0), (-0.45,1.75,0) bohr; 10 electrons (5/5)."""
    R = [[0.0, 0.0, 0.0], [1.81, 0.0, 0.0], [-0.45, 1.75, 0.0]]
    return make_molecule(R, [8, 1, 1], name="H2O_like")


def h_chain(n: int = 64, spacing: float = 1.8) -> Molecule:
    return chain("H", spacing, n)


def zigzag_alkane(n_carbon: int = 16, d_cc: float = 2.9, d_ch: float = 2.06) -> Molecule:
    """Zigzag all-trans alkane C_nH_{2n+2} with tetrahedral angles."""
    theta = math.acos(-1.0 / 3.0)  # 109.47 deg
    half = theta / 2
    dx = d_cc * math.sin(half)
    dy = d_cc * math.cos(half)
    C = np.zeros((n_carbon, 3))
    C[:, 0] = np.arange(n_carbon) * dx
    C[:, 1] = np.where(np.arange(n_carbon) % 2 == 0, 0.0, dy)
    H = []
    ez = np.array([0.0, 0.0, 1.0])
    for k in range(n_carbon):
        nbrs = [C[j] - C[k] for j in (k - 1, k + 1) if 0 <= j < n_carbon]
        units = [v / np.linalg.norm(v) for v in nbrs]
        if len(units) == 2:
            u = -(units[0] + units[1])
            u /= np.linalg.norm(u)
            n = np.cross(units[0], units[1])
            n /= np.linalg.norm(n)
            a = math.acos(-1.0 / 3.0) / 2
            for s in (+1, -1):
                H.append(C[k] + d_ch * (u * math.cos(a) + s * n * math.sin(a)))
        else:
            a = units[0]
            # in-plane perpendicular pointing away from the zigzag partner side
            p = np.cross(ez, a)
            p /= np.linalg.norm(p)
            if k == 0 and n_carbon > 1:
                p = p if p[1] <= 0 else -p
            else:
                p = p if (C[k][1] - C[k - 1][1]) * p[1] >= 0 else -p
            for phi in (0.0, 2 * math.pi / 3, 4 * math.pi / 3):
                d = a * math.cos(theta) + math.sin(theta) * (p * math.cos(phi) + ez * math.sin(phi))
                H.append(C[k] + d_ch * d)
    R = np.concatenate([C, np.array(H)], axis=0)
    Z = [6] * n_carbon + [1] * len(H)
    return make_molecule(R, Z, name=f"C{n_carbon}H{len(H)}_zigzag")


def random_chain(rng: np.random.Generator, n_el_target: int, elements=("H", "C", "N", "O"),
                 spacing=(1.8, 2.9)) -> Molecule:
    """Seeded random H/C/N/O chain whose electron count is ``n_el_target``
    (the last atom is an H-to-O chosen to hit the count exactly)."""
    Zs = []
    remaining = n_el_target
    while remaining > 0:
        choices = [ELEMENTS[e] for e in elements if ELEMENTS[e] <= remaining]
        z = int(rng.choice(choices))
        Zs.append(z)
        remaining -= z
    n = len(Zs)
    R = np.zeros((n, 3))
    R[1:, 0] = np.cumsum(rng.uniform(spacing[0], spacing[1], size=n - 1))
    R[:, 1] = rng.normal(0.0, 0.3, size=n)
    return make_molecule(R, Zs, name=f"chain{n_el_target}")


def assign_spins_to_atoms(R: np.ndarray, Z: np.ndarray) -> np.ndarray:
    """Atom index of every electron: first all up spins, then all down spins.
    Restates mcmc.py:381-428 (greedy anti-parallel placement of open-shell spins)."""
    Z = np.asarray(Z, dtype=np.int64)
    n_up_per_atom = Z // 2
    n_el_remaining = int(Z.sum() - 2 * n_up_per_atom.sum())
    if n_el_remaining > 0:
        ind_open_shell = np.where(Z % 2)[0]
        R_open = np.asarray(R)[ind_open_shell]
        dist = np.linalg.norm(R_open[:, None, :] - R_open[None, :, :], axis=-1)
        kernel = np.exp(-dist * 0.5)
        spins = np.zeros(n_el_remaining)
        n_dn_left = n_el_remaining // 2
        n_up_left = n_el_remaining - n_dn_left
        for _ in range(n_el_remaining):
            is_free = spins == 0
            spin_per_site = kernel[is_free, :] @ spins
            ind_atom = np.arange(n_el_remaining)[is_free]
            loss_up = spin_per_site
            loss_dn = -spin_per_site
            if (n_up_left > 0) and (np.min(loss_up) < np.min(loss_dn)):
                ind = ind_atom[np.argmin(loss_up)]
                spins[ind] = 1
                n_up_left -= 1
            else:
                ind = ind_atom[np.argmin(loss_dn)]
                spins[ind] = -1
                n_dn_left -= 1
        n_up_per_atom[ind_open_shell] += (spins == 1).astype(np.int64)
    n_dn_per_atom = Z - n_up_per_atom
    ind_atom = []
    for i, n in enumerate(n_up_per_atom):
        ind_atom += [i] * int(n)
    for i, n in enumerate(n_dn_per_atom):
        ind_atom += [i] * int(n)
    return np.array(ind_atom, dtype=np.int64)


def init_electrons(mol: Molecule, batch_size: int, stddev: float = 1.0, seed: int = 0) -> np.ndarray:
    """Walkers [batch, n_el, 3] float32 drawn as N(assigned nucleus, stddev) (mcmc.py:431-465)."""
    rng = np.random.default_rng(seed)
    r = rng.normal(size=(batch_size, mol.n_el, 3)).astype(np.float32) * np.float32(stddev)
    if mol.n_nuc > 1:
        if int(mol.Z.sum()) == mol.n_el:
            ind_atom = assign_spins_to_atoms(mol.R, mol.Z)
            r = r + mol.R[ind_atom][None]
        else:
            p = mol.Z / mol.Z.sum()
            ind = rng.choice(mol.n_nuc, size=(batch_size, mol.n_el), p=p)
            r = r + mol.R[ind]
    else:
        r = r + mol.R[0]
    return r.astype(np.float32)


# Named workloads ------------------------------------------------------------------------
def workload(name: str):
    """(molecule, model_args overrides, walkers, electron stddev) for the named workloads."""
    if name == "h2o_like":
        return water_like(), dict(cutoff=5.0, feature_dim=128, n_determinants=1), 512, 1.0
    if name == "h64":
        return h_chain(64, 1.8), dict(cutoff=5.0, feature_dim=256, n_determinants=4), 2048, 0.5
    if name == "c16h34":
        return zigzag_alkane(16), dict(cutoff=5.0, feature_dim=256, n_determinants=4), 1024, 1.0
    raise KeyError(name)
