"""Pseudopotential test systems: a carbon ECP in pyscf's layout with ccECP-like magnitudes
(the reference loads the real coefficients with pyscf, which is not available here)."""
from __future__ import annotations

import numpy as np

from sparse_wf_amd.config import FireConfig, Molecule
from sparse_wf_amd import systems

# (n_core, [(l, [[(exponent, coefficient)] for r^0, r^1, r^2, ...]), ...]); l = -1 is local
CARBON_ECP = (2, [
    (-1, [[], [(14.43502, 4.0)], [(7.38188, -25.81955)], [(8.39889, 57.74008)]]),
    (0, [[], [], [(7.76079, 52.13345)]]),
    (1, [[], [], [(5.10224, 2.40501)]]),
])


def ethane_valence(**kw) -> FireConfig:
    """C2H6 with the carbon cores removed: 14 valence electrons (7 up)."""
    m = systems.zigzag_alkane(2)
    mol = Molecule(R=m.R, Z=m.Z, n_el=14, n_up=7, name="C2H6_valence")
    base = dict(cutoff=3.0, feature_dim=64, pair_mlp_widths=(4, 4), pair_n_envelopes=4, n_determinants=2,
                n_envelopes=8, mlp_width=32, mlp_depth=2, attention_heads=4, attention_dim=4)
    base.update(kw)
    return FireConfig(mol=mol, **base)


def valence_walkers(cfg: FireConfig, B: int, seed: int = 0) -> np.ndarray:
    """Electrons near the nuclei: carbons hold 4 valence electrons each, hydrogens 1."""
    rng = np.random.default_rng(seed)
    owners = []
    for J, z in enumerate(cfg.mol.Z):
        owners += [J] * (4 if int(z) == 6 else 1)
    owners = np.asarray(owners)
    up = np.concatenate([owners[0::2], owners[1::2]])[: cfg.n_el]
    r = cfg.mol.R[up][None] + 0.8 * rng.normal(size=(B, cfg.n_el, 3))
    return r.astype(np.float32)
