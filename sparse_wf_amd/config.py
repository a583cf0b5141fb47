"""Model configuration for the FiRE hot path.

The fields mirror the reference's configuration surface:

* ``model_args.embedding.new`` -> ``NewEmbeddingArgs`` (src/sparse_wf/api.py:426-434,
  defaults config/default.yaml:31-39),
* ``model_args.jastrow`` -> ``JastrowArgs`` (api.py:453-462, config/default.yaml:48-57),
* ``model_args.envelopes`` -> ``EnvelopeArgs`` (api.py:475-478, config/default.yaml:58-65),
* ``model_args.n_determinants`` / ``spin_restricted`` (config/default.yaml:27-28).

Only ``embedding == "new_sparse"`` with ``n_updates == 1`` exists in the reference's sparse
model (src/sparse_wf/model/new_sparse_model.py:326); the same restriction is enforced here.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import Any, Mapping

import numpy as np


DEFAULT_NEW_EMBEDDING = dict(
    cutoff=3.0,
    cutoff_1el=20.0,
    feature_dim=256,
    pair_mlp_widths=(16, 8),
    pair_n_envelopes=32,
    low_rank_buffer=2,
    n_updates=1,
    rematerialize_pairs=False,
)

DEFAULT_JASTROW = dict(
    e_e_cusps="psiformer",
    use_e_e_mlp=False,
    use_mlp_jastrow=True,
    use_log_jastrow=True,
    mlp_width=256,
    mlp_depth=2,
    use_attention=True,
    attention_heads=16,
    attention_dim=16,
)

DEFAULT_ENVELOPES = dict(envelope="isotropic", isotropic_args=dict(n_envelopes=8))


@dataclass(frozen=True)
class Molecule:
    """Nuclei and electron counts (the part of a pyscf ``Mole`` the hot path reads)."""

    R: np.ndarray  # [n_nuc, 3] float32, bohr
    Z: np.ndarray  # [n_nuc] int32
    n_el: int
    n_up: int
    name: str = ""

    @property
    def n_nuc(self) -> int:
        return int(self.R.shape[0])

    @property
    def n_dn(self) -> int:
        return self.n_el - self.n_up


@dataclass(frozen=True)
class FireConfig:
    """Static hyper-parameters of one FiRE wavefunction (one molecule)."""

    mol: Molecule
    # model_args.embedding.new
    cutoff: float = 3.0
    cutoff_1el: float = 20.0
    feature_dim: int = 256
    pair_mlp_widths: tuple = (16, 8)
    pair_n_envelopes: int = 32
    low_rank_buffer: int = 2
    n_updates: int = 1
    rematerialize_pairs: bool = False
    # model_args
    n_determinants: int = 4
    spin_restricted: bool = False
    # model_args.envelopes (isotropic only)
    n_envelopes: int = 8
    # model_args.jastrow
    e_e_cusps: str = "psiformer"
    use_e_e_mlp: bool = False
    use_mlp_jastrow: bool = True
    use_log_jastrow: bool = True
    mlp_width: int = 256
    mlp_depth: int = 2
    use_attention: bool = True
    attention_heads: int = 16
    attention_dim: int = 16

    def __post_init__(self):
        if self.n_updates != 1:
            raise ValueError("Only n_updates==1 implemented (as new_sparse_model.py:326)")
        if self.e_e_cusps not in ("none", "psiformer"):
            raise ValueError(f"Unsupported e_e_cusps: {self.e_e_cusps!r} (supported: none, psiformer)")
        if self.use_e_e_mlp:
            raise ValueError("use_e_e_mlp=True (dense all-pair e-e MLP Jastrow) is not supported")
        if len(self.pair_mlp_widths) != 2:
            raise ValueError("pair_mlp_widths must have two entries")
        if self.mlp_depth < 1:
            raise ValueError("mlp_depth must be >= 1")

    # convenience sizes ---------------------------------------------------------------
    @property
    def n_el(self) -> int:
        return self.mol.n_el

    @property
    def n_up(self) -> int:
        return self.mol.n_up

    @property
    def n_nuc(self) -> int:
        return self.mol.n_nuc

    @classmethod
    def from_model_args(cls, mol: Molecule, model_args: Mapping[str, Any]) -> "FireConfig":
        """Build from a reference-style ``model_args`` dict (config/default.yaml:26-65)."""
        emb = model_args["embedding"]
        kind = emb.get("embedding", "new_sparse").lower()
        if kind != "new_sparse":
            raise ValueError(f"Only embedding='new_sparse' is provided here, got {kind!r}")
        new = dict(DEFAULT_NEW_EMBEDDING)
        new.update({k: v for k, v in emb.get("new", {}).items() if k in DEFAULT_NEW_EMBEDDING})
        jas = dict(DEFAULT_JASTROW)
        jas.update(model_args.get("jastrow", {}))
        env = model_args.get("envelopes", DEFAULT_ENVELOPES)
        if env.get("envelope", "isotropic") != "isotropic":
            raise ValueError("Only isotropic envelopes are provided")
        return cls(
            mol=mol,
            cutoff=float(new["cutoff"]),
            cutoff_1el=float(new["cutoff_1el"]),
            feature_dim=int(new["feature_dim"]),
            pair_mlp_widths=tuple(int(w) for w in new["pair_mlp_widths"]),
            pair_n_envelopes=int(new["pair_n_envelopes"]),
            low_rank_buffer=int(new["low_rank_buffer"]),
            n_updates=int(new["n_updates"]),
            rematerialize_pairs=bool(new["rematerialize_pairs"]),
            n_determinants=int(model_args.get("n_determinants", 4)),
            spin_restricted=bool(model_args.get("spin_restricted", False)),
            n_envelopes=int(env.get("isotropic_args", {}).get("n_envelopes", 8)),
            e_e_cusps=str(jas["e_e_cusps"]),
            use_e_e_mlp=bool(jas["use_e_e_mlp"]),
            use_mlp_jastrow=bool(jas["use_mlp_jastrow"]),
            use_log_jastrow=bool(jas["use_log_jastrow"]),
            mlp_width=int(jas["mlp_width"]),
            mlp_depth=int(jas["mlp_depth"]),
            use_attention=bool(jas["use_attention"]),
            attention_heads=int(jas["attention_heads"]),
            attention_dim=int(jas["attention_dim"]),
        )

    def replace(self, **kw) -> "FireConfig":
        return dataclasses.replace(self, **kw)
