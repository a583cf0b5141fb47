"""Host-side mirror of the reference's wavefunction API, backed by the HIP library.

``FireWaveFunction`` keeps the method names and argument meaning of
``MoonLikeWaveFunction`` (src/sparse_wf/model/wave_function.py:67-287) and the
``ParameterizedWaveFunction`` protocol (src/sparse_wf/api.py:161-185): ``create``, ``init``,
``get_static_input``, ``signed``, ``__call__``, ``kinetic_energy``, ``log_psi_with_state``,
``log_psi_low_rank_update``.  Electrons may be a single walker ``[N, 3]`` (as the reference's
per-walker functions) or a batch ``[B, N, 3]`` (the reference's ``vmap``); results carry
the same leading shape.  Arrays live on the GPU as torch tensors; every computation runs
in ``libfire_amd.so``.
"""
from __future__ import annotations

import ctypes
import hashlib
from dataclasses import dataclass
from typing import Any, Mapping, NamedTuple, Optional

import numpy as np
import torch

try:
    import xxhash
except ImportError:  # hashlib is the fallback
    xxhash = None

from . import _lib
from .config import FireConfig, Molecule
from .params import init_params, pack_params, pack_params_vjp, padded_w1, set_scales, tree_map

# "mixed" (the default): 16-bit matrix cores on split operands -- two scaled FP16 planes on the
# tangent-row contractions of the local energy, three BF16 planes elsewhere (fire_precision_plan);
# "mixed_fast": two BF16 planes (~2^-17 per product); "bf16": one plane; "fp32": FP32 MFMA.
PRECISIONS = {"fp32": _lib.FIRE_FP32, "mixed": _lib.FIRE_MIXED, "bf16": _lib.FIRE_BF16,
              "mixed_fast": _lib.FIRE_MIXED_FAST}


class StaticArgs(NamedTuple):
    """Neighbour counts of ``StaticArgs`` (new_sparse_model.py:46-80).  The HIP path uses
    ragged CSR lists, so these are informative (capacity bookkeeping of the reference)."""

    n_pairs_same: int
    n_pairs_diff: int
    n_triplets: int
    n_neighbours_en: int
    n_changed_hout: int
    n_changed_pair_same: int
    n_changed_pair_diff: int
    n_pp_elecs: int


def _config_c(cfg: FireConfig):
    """(fire_config, R float32 [M, 3], Z float32 [M]) of a configuration."""
    c = _lib.FireConfigC(
        n_el=cfg.n_el, n_up=cfg.n_up, n_nuc=cfg.n_nuc, feature_dim=cfg.feature_dim,
        filter_w0=cfg.pair_mlp_widths[0], filter_w1=padded_w1(cfg.pair_mlp_widths[1]),
        n_env_en=cfg.pair_n_envelopes, n_det=cfg.n_determinants, n_env_iso=cfg.n_envelopes,
        e_e_cusp=1 if cfg.e_e_cusps == "psiformer" else 0,
        use_mlp_jastrow=int(cfg.use_mlp_jastrow), use_log_jastrow=int(cfg.use_log_jastrow),
        use_attention=int(cfg.use_attention), mlp_width=cfg.mlp_width, mlp_depth=cfg.mlp_depth,
        att_heads=cfg.attention_heads, att_dim=cfg.attention_dim,
        cutoff=cfg.cutoff, cutoff_1el=cfg.cutoff_1el,
    )
    return c, np.ascontiguousarray(cfg.mol.R, np.float32), np.ascontiguousarray(cfg.mol.Z, np.float32)


def _stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def _dptr(t: Optional[torch.Tensor]):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class FireWaveFunction:
    """FiRE (``embedding: new_sparse``) wavefunction on one MI355X."""

    def __init__(self, cfg: FireConfig, device: int = 0, precision: str = "mixed"):
        self.cfg = cfg
        self.device = torch.device("cuda", device)
        self.precision = precision
        lib = _lib.load()
        c, R, Z = _config_c(cfg)
        h = ctypes.c_void_p()
        _lib.check(lib.fire_create(ctypes.byref(c), R.ctypes.data, Z.ctypes.data, device, ctypes.byref(h)))
        self._h = h
        self._lib = lib
        self._uploaded = None
        self._uploaded_fp = None
        self._gen = 0
        self._committed_gen = self._pending_gen = None

    def reconfigure(self, cfg: FireConfig):
        """Switch this wave function (and its GPU handle) to another molecule / model, keeping the
        device workspace; parameters must be uploaded again and the resident state is gone."""
        c, R, Z = _config_c(cfg)
        _lib.check(self._lib.fire_reconfigure(self._h, ctypes.byref(c), R.ctypes.data, Z.ctypes.data))
        self.cfg = cfg
        self._uploaded = self._uploaded_fp = None
        self._committed_gen = self._pending_gen = None
        self._pp = None

    def workspace_bytes(self) -> int:
        """Device bytes held by the handle (workspace, walker state, parameters)."""
        n = ctypes.c_int64()
        _lib.check(self._lib.fire_workspace_bytes(self._h, ctypes.byref(n)))
        return n.value

    # reference-style constructor ---------------------------------------------------------
    @classmethod
    def create(cls, mol: Molecule, embedding: Mapping[str, Any], jastrow: Mapping[str, Any],
               envelopes: Mapping[str, Any], n_determinants: int, spin_restricted: bool, **kw):
        """MoonLikeWaveFunction.create (wave_function.py:102-150)."""
        cfg = FireConfig.from_model_args(mol, dict(embedding=embedding, jastrow=jastrow, envelopes=envelopes,
                                                   n_determinants=n_determinants, spin_restricted=spin_restricted))
        return cls(cfg, **kw)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._lib.fire_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def R(self):
        return self.cfg.mol.R

    @property
    def Z(self):
        return self.cfg.mol.Z

    @property
    def n_electrons(self):
        return self.cfg.n_el

    @property
    def n_up(self):
        return self.cfg.n_up

    # parameters ------------------------------------------------------------------------
    def upload(self, params, fingerprint=None):
        """Upload a parameter tree (reference layout) or an already packed ``PackedParams``.
        The resident walker state of ``log_psi_with_state`` belongs to the previous parameters
        and is dropped."""
        if isinstance(params, PackedParams):
            flat = np.ascontiguousarray(params.flat, np.float32)
            sizes = np.ascontiguousarray(params.sizes, np.int64)
        else:
            tensors = pack_params(self.cfg, params)
            sizes = np.asarray([t.size for t in tensors], np.int64)
            flat = np.ascontiguousarray(np.concatenate(tensors).astype(np.float32))
        _lib.check(self._lib.fire_upload_params(self._h, flat.ctypes.data, sizes.ctypes.data, len(sizes)))
        self._uploaded = params
        self._uploaded_fp = _fingerprint(params) if fingerprint is None else fingerprint
        self._committed_gen = self._pending_gen = None

    def _ensure(self, params) -> bool:
        """Upload ``params`` unless the device already holds them: the same object with the
        same contents (a parameter tree changed in place is uploaded again).  Returns whether
        an upload happened; an upload invalidates the resident walker state."""
        fp = _fingerprint(params)
        if params is self._uploaded and fp == self._uploaded_fp:
            return False
        self.upload(params, fp)
        return True

    def init(self, seed: int, electrons) -> dict:
        """MoonLikeWaveFunction.init (wave_function.py:89-99): random parameters with the
        normalisation scales set to 1/std of the embedding on ``electrons`` (one walker).
        The returned parameters are the ones left on the device."""
        params = init_params(self.cfg, seed)

        def stages(p):
            self.upload(p)
            hc, ho = self.embedding(p, electrons, unscaled_out=True)
            return hc.double().cpu().numpy(), ho.double().cpu().numpy()

        params = set_scales(params, stages)
        # set_scales writes s1 into the tree after its last upload (made with s1 = 1)
        self.upload(params)
        return params

    def load_checkpoint(self, src) -> dict:
        """Read a reference training checkpoint (``chkpt*.msgpk``, path or bytes; written by
        ``TrainingState.serialize`` api.py:294-303) and upload its parameters.  Returns
        ``{"params", "electrons", "step", "key", "width"}``."""
        from .checkpoint import load_training_state

        st = load_training_state(src, self.cfg)
        self.upload(st["params"])
        return st

    # inputs ----------------------------------------------------------------------------
    def _walkers(self, electrons):
        r = torch.as_tensor(electrons, dtype=torch.float32, device=self.device)
        single = r.dim() == 2
        if single:
            r = r.unsqueeze(0)
        if r.shape[1:] != (self.cfg.n_el, 3):
            raise ValueError(f"electrons must be [..., {self.cfg.n_el}, 3], got {tuple(r.shape)}")
        return r.contiguous(), single

    def _prec(self, precision):
        p = precision or self.precision
        if p not in PRECISIONS:
            raise ValueError(f"precision must be one of {list(PRECISIONS)}")
        return PRECISIONS[p]

    # API -------------------------------------------------------------------------------
    def get_static_input(self, electrons, electrons_new=None, idx_changed=None, laplacian=True) -> StaticArgs:
        """get_static_input (new_sparse_model.py:399-437), computed from the GPU neighbour lists
        and taken as the maximum over walkers: pair counts by spin, triplets, nuclei within
        cutoff_1el and, for a move of electron ``idx_changed`` from ``electrons`` to
        ``electrons_new``, the affected electrons (neighbours of the moved one before or after
        the move, plus itself) and the new pairs centred on them."""
        r_old, _ = self._walkers(electrons)
        r, _ = self._walkers(electrons if electrons_new is None else electrons_new)
        B, N, nu = r.shape[0], self.cfg.n_el, self.cfg.n_up
        dev = self.device

        def pairs(rr):
            counts, nbr, _ = self.neighbours(rr)
            ctr = torch.repeat_interleave(torch.arange(B * N, device=dev), counts.reshape(-1).long())
            return counts.reshape(B, N).long(), ctr // N, ctr % N, nbr.long()

        counts, w, i, j = pairs(r)
        spin_i, spin_j = i < nu, j < nu
        same = spin_i == spin_j

        def per_walker(mask):
            return torch.zeros(B, dtype=torch.long, device=dev).index_add_(0, w, mask.long())

        n_same = per_walker(same)
        n_diff = per_walker(~same)
        trip = int((counts ** 2).sum(1).max()) if laplacian else 0
        R = torch.as_tensor(np.asarray(self.cfg.mol.R, np.float32), device=dev)
        d = r[:, :, None, :] - R[None, None]
        den = torch.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
        n_en = int((den < np.float32(self.cfg.cutoff_1el)).sum(-1).max())
        n_hout = n_cs = n_cd = 0
        if electrons_new is not None:
            k = int(np.asarray(idx_changed).reshape(-1)[0])
            aff = torch.zeros((B, N), dtype=torch.bool, device=dev)
            aff[w[j == k], i[j == k]] = True
            _, wo, io, jo = pairs(r_old)
            aff[wo[jo == k], io[jo == k]] = True
            aff[:, k] = True
            req = aff[w, i]
            n_hout = int(aff.sum(1).max())
            n_cs = int(per_walker(req & same).max())
            n_cd = int(per_walker(req & ~same).max())
        return StaticArgs(int(n_same.max()), int(n_diff.max()), trip, n_en, n_hout, n_cs, n_cd, N)

    def neighbours(self, electrons):
        """CSR neighbour lists: counts [B, N], neighbour indices [P], row pointer [B*N+1]."""
        r, _ = self._walkers(electrons)
        B, N = r.shape[0], self.cfg.n_el
        counts = torch.empty((B, N), dtype=torch.int32, device=self.device)
        npairs = ctypes.c_int64()
        _lib.check(self._lib.fire_neighbours(self._h, _stream_ptr(), _dptr(r), B, _dptr(counts), None, 0,
                                             ctypes.byref(npairs)))
        nbr = torch.empty(max(npairs.value, 1), dtype=torch.int32, device=self.device)
        _lib.check(self._lib.fire_neighbours(self._h, _stream_ptr(), _dptr(r), B, _dptr(counts), _dptr(nbr),
                                             nbr.numel(), ctypes.byref(npairs)))
        ptr = torch.zeros(B * N + 1, dtype=torch.int64, device=self.device)
        ptr[1:] = torch.cumsum(counts.reshape(-1).long(), 0)
        return counts, nbr[: npairs.value], ptr

    def signed(self, params, electrons, static=None, precision=None):
        """(sign, log|psi|) -- MoonLikeWaveFunction.signed (wave_function.py:183-190)."""
        self._ensure(params)
        r, single = self._walkers(electrons)
        B = r.shape[0]
        lp = torch.empty(B, dtype=torch.float64, device=self.device)
        sg = torch.empty(B, dtype=torch.float64, device=self.device)
        _lib.check(self._lib.fire_signed(self._h, _stream_ptr(), _dptr(r), B, self._prec(precision), _dptr(lp),
                                         _dptr(sg)))
        return (sg[0], lp[0]) if single else (sg, lp)

    def __call__(self, params, electrons, static=None, precision=None):
        return self.signed(params, electrons, static, precision)[1]

    def grad_logpsi(self, params, electrons, packed: bool = False):
        """Per-walker d log|psi| / d theta -- the jax.grad(log_p) mapped over walkers by the
        natural-gradient preconditioners (preconditioner.py:191-194, 290-294).  Returns
        (grad, logpsi): grad a tree shaped like ``params`` with leaves [B, *shape] (float32,
        through the adjoint of the packing), or with ``packed`` (always for ``PackedParams``)
        the [B, n] rows over the packed tensors; logpsi [B] float64."""
        self._ensure(params)
        r, single = self._walkers(electrons)
        B = r.shape[0]
        n = ctypes.c_int64()
        _lib.check(self._lib.fire_param_count(self._h, ctypes.byref(n)))
        g = torch.empty((B, n.value), dtype=torch.float32, device=self.device)
        lp = torch.empty(B, dtype=torch.float64, device=self.device)
        _lib.check(self._lib.fire_grad_logpsi(self._h, _stream_ptr(), _dptr(r), B, _dptr(g), _dptr(lp)))
        if not (packed or isinstance(params, PackedParams)):
            g = pack_params_vjp(self.cfg, params, g)
        if single:
            g = tree_map(lambda t: t[0], g) if isinstance(g, dict) else g[0]
            lp = lp[0]
        return g, lp

    def local_energy_terms(self, params, electrons, static=None, precision=None, with_derivatives=False):
        """dict with e_loc, e_kin, e_pot, logpsi, sign (and grad [.., N, 3], lap of log|psi|)."""
        self._ensure(params)
        r, single = self._walkers(electrons)
        B, N = r.shape[0], self.cfg.n_el
        out = {k: torch.empty(B, dtype=torch.float64, device=self.device)
               for k in ("e_loc", "e_kin", "e_pot", "logpsi", "sign")}
        if with_derivatives:
            out["grad"] = torch.empty((B, N, 3), dtype=torch.float64, device=self.device)
            out["lap"] = torch.empty(B, dtype=torch.float64, device=self.device)
        _lib.check(self._lib.fire_local_energy(
            self._h, _stream_ptr(), _dptr(r), B, self._prec(precision), _dptr(out["e_loc"]), _dptr(out["e_kin"]),
            _dptr(out["e_pot"]), _dptr(out["logpsi"]), _dptr(out["sign"]), _dptr(out.get("grad")),
            _dptr(out.get("lap"))))
        if single:
            out = {k: v[0] for k, v in out.items()}
        return out

    # pseudopotentials -------------------------------------------------------------------
    def set_pseudopotential(self, pp):
        """Pseudopotential tables (``pseudopotentials.make_pseudopotential``) for
        ``pseudopotential_energy``; from now on the Coulomb terms of the local energy use the
        effective charges (make_local_energy passes them to potential_energy,
        hamiltonian.py:66-75).  ``None`` removes them and restores the nuclear charges."""
        from .pseudopotentials import make_pseudopotential, upload

        if pp is None:
            pp = make_pseudopotential(self.cfg.mol.Z, {})
        upload(self._lib, self._h, pp)
        self._pp = pp if pp.n_ecp else None

    def pseudopotential_energy(self, params, electrons, seed: int = 0, precision=None) -> dict:
        """Local and non-local pseudopotential energy per walker (pp_local / pp_nonloc of
        make_local_energy, hamiltonian.py:72-75): the walkers become the resident state
        (log_psi_with_state), the psi ratios of the quadrature points come from its low-rank
        proposals.  Returns ``{"v_loc", "v_nl"}`` (double [B]) and ``n_points``, the active
        integration points (the reference's n_pp_elecs statistic)."""
        if getattr(self, "_pp", None) is None:
            raise ValueError("no pseudopotential set (set_pseudopotential)")
        r, single = self._walkers(electrons)
        self.log_psi_with_state(params, r, precision=precision)
        B = r.shape[0]
        vl = torch.empty(B, dtype=torch.float64, device=self.device)
        vn = torch.empty(B, dtype=torch.float64, device=self.device)
        n = ctypes.c_int64()
        _lib.check(self._lib.fire_ecp_energy(self._h, _stream_ptr(), ctypes.c_uint64(int(seed) & (2 ** 64 - 1)),
                                             _dptr(vl), _dptr(vn), ctypes.byref(n)))
        return {"v_loc": vl[0] if single else vl, "v_nl": vn[0] if single else vn, "n_points": n.value}

    def kinetic_energy(self, params, electrons, static=None, precision=None):
        """-1/2 (Laplacian + |grad|^2) of log|psi| -- wave_function.py:215-221."""
        return self.local_energy_terms(params, electrons, static, precision)["e_kin"]

    def embedding(self, params, electrons, precision="fp32", unscaled_out=False):
        """(h after message passing, h_out) [B, N, F] -- NewSparseEmbedding.apply
        (new_sparse_model.py:483-551).  ``unscaled_out`` returns the update-MLP output before
        the scale s1 (what init's second normalize() sees, :537)."""
        self._ensure(params)
        r, single = self._walkers(electrons)
        B, N, F = r.shape[0], self.cfg.n_el, self.cfg.feature_dim
        hc = torch.empty((B, N, F), dtype=torch.float32, device=self.device)
        ho = torch.empty((B, N, F), dtype=torch.float32, device=self.device)
        _lib.check(self._lib.fire_embedding(self._h, _stream_ptr(), _dptr(r), B, self._prec(precision), _dptr(hc),
                                            _dptr(ho)))
        if unscaled_out:
            ho = ho / float(params["embedding"]["scales"][1])
        return (hc[0], ho[0]) if single else (hc, ho)

    # resident state and low-rank updates ------------------------------------------------
    def log_psi_with_state(self, params, electrons, static=None, precision=None):
        """((sign, log|psi|), state) -- MoonLikeWaveFunction.log_psi_with_state
        (wave_function.py:226-249).  The state stays resident on the GPU inside this wave
        function (one state per instance); the returned ``WalkerState`` names it."""
        self._ensure(params)
        r, single = self._walkers(electrons)
        B = r.shape[0]
        lp = torch.empty(B, dtype=torch.float64, device=self.device)
        sg = torch.empty(B, dtype=torch.float64, device=self.device)
        _lib.check(self._lib.fire_state_init(self._h, _stream_ptr(), _dptr(r), B, self._prec(precision), _dptr(lp),
                                             _dptr(sg)))
        self._gen += 1
        self._committed_gen, self._pending_gen = self._gen, None
        st = WalkerState(self, self._gen, B, single, pending=False)
        return ((sg[0], lp[0]) if single else (sg, lp)), st

    def _use_state(self, params, state: "WalkerState"):
        if params is not self._uploaded or _fingerprint(params) != self._uploaded_fp:
            raise ValueError("parameters changed since log_psi_with_state: the resident state belongs to the old ones")
        if state.wf is not self:
            raise ValueError("state belongs to another wave function")
        if self._committed_gen is None:
            raise ValueError("no resident state: parameters were uploaded after log_psi_with_state")
        if state.pending:
            if state.gen != self._pending_gen:
                raise ValueError("stale proposal state")
            self.commit(state)  # the proposal's state is taken as the new state
        elif state.gen != self._committed_gen:
            raise ValueError("stale state: a later update was committed")

    def log_psi_low_rank_update(self, params, electrons, changed_electrons, static=None, state=None):
        """((sign, log|psi|), proposal state) after moving electron ``changed_electrons`` of every
        walker to its position in ``electrons`` -- MoonLikeWaveFunction.log_psi_low_rank_update
        (wave_function.py:251-287).  The resident state is not modified until the returned
        proposal is committed: by ``commit(proposal, accept)``, or in full when the proposal is
        passed as ``state`` to the next update (the reference's functional semantics)."""
        if state is None:
            raise ValueError("state from log_psi_with_state is required")
        self._use_state(params, state)
        r, single = self._walkers(electrons)
        idx = np.asarray(changed_electrons).reshape(-1)
        if idx.size != 1:
            raise ValueError("one changed electron per update (single-electron moves)")
        k = int(idx[0])
        if r.shape[0] != state.n_walkers:
            raise ValueError("walker count differs from the state's")
        rk = r[:, k, :].contiguous()
        B = r.shape[0]
        lp = torch.empty(B, dtype=torch.float64, device=self.device)
        sg = torch.empty(B, dtype=torch.float64, device=self.device)
        _lib.check(self._lib.fire_state_propose(self._h, _stream_ptr(), k, _dptr(rk), _dptr(lp), _dptr(sg)))
        self._gen += 1
        self._pending_gen = self._gen
        st = WalkerState(self, self._gen, B, state.single, pending=True)
        return ((sg[0], lp[0]) if single else (sg, lp)), st

    def commit(self, proposal: "WalkerState", accept=None):
        """Apply a proposal to the walkers where ``accept`` is true (all when None) -- the
        jnp.where(accept, new, old) state selection of mcmc_steps_low_rank (mcmc.py:240-244)."""
        if not proposal.pending or proposal.gen != self._pending_gen:
            raise ValueError("not the pending proposal")
        B = proposal.n_walkers
        if accept is None:
            acc = torch.ones(B, dtype=torch.int32, device=self.device)
        else:
            acc = torch.as_tensor(accept, device=self.device).reshape(-1).to(torch.int32).contiguous()
            if acc.numel() != B:
                raise ValueError("accept must have one entry per walker")
        _lib.check(self._lib.fire_state_commit(self._h, _stream_ptr(), _dptr(acc)))
        self._committed_gen, self._pending_gen = proposal.gen, None
        proposal.pending = False

    def state_values(self):
        """(electrons [B, N, 3], sign [B], log|psi| [B]) of the committed resident state."""
        B = self._state_walkers()
        r = torch.empty((B, self.cfg.n_el, 3), dtype=torch.float32, device=self.device)
        lp = torch.empty(B, dtype=torch.float64, device=self.device)
        sg = torch.empty(B, dtype=torch.float64, device=self.device)
        _lib.check(self._lib.fire_state_get(self._h, _stream_ptr(), _dptr(r), _dptr(lp), _dptr(sg)))
        self._pending_gen = None
        return r, sg, lp

    def _state_walkers(self) -> int:
        nw = ctypes.c_int32()
        _lib.check(self._lib.fire_state_stats(self._h, ctypes.byref(nw), None, None, None))
        if nw.value == 0:
            raise ValueError("no resident state (call log_psi_with_state first)")
        return nw.value

    def autotune(self, mode="on"):
        """Contraction tile autotuner of this process and device (``fire_autotune``): "on"
        (time each new shape class), "off", "reset" (on, earlier choices forgotten) or "near"
        (on, and shapes up to two row buckets from a tuned one reuse its choice: ragged and
        serving batches, whose geometries each bring new shapes).  An int v runs every
        contraction of this handle on tile form v where that form applies (a diagnostic)."""
        codes = {"off": 0, "on": 1, "reset": 2, "near": 3}
        if isinstance(mode, int) and not isinstance(mode, bool):
            code = 100 + mode
        elif mode in codes:
            code = codes[mode]
        else:
            raise ValueError(f"autotune mode {mode!r} (one of {sorted(codes)} or a tile form number)")
        n = ctypes.c_int32()
        _lib.check(self._lib.fire_autotune(self._h, code, None, 0, ctypes.byref(n)))

    def tile_choices(self):
        """This handle's tile form per contraction shape, as (split, N, K, log2 M code, form)."""
        cap = 512
        table = (ctypes.c_int32 * (7 * cap))()
        n = ctypes.c_int32()
        _lib.check(self._lib.fire_autotune(self._h, -1, table, cap, ctypes.byref(n)))
        return [tuple(table[7 * i:7 * i + 5]) for i in range(min(n.value, cap))]

    @staticmethod
    def tile_forms() -> int:
        """Number of tile forms of the split contractions (valid arguments of autotune(int))."""
        return int(_lib.load().fire_tile_forms())

    def precision_plan(self, planes=None):
        """Operand planes per contraction class of the "mixed" precision (ElecInit MLP, update MLP,
        orbitals, MLP Jastrow, attention Q/V, attention readout; 1-3 BF16 planes, 4 two scaled
        FP16 planes): sets the leading classes from ``planes`` (if given) and returns the plan."""
        planes = [] if planes is None else [int(p) for p in planes]
        arr = (ctypes.c_int32 * 6)(*(planes + [0] * (6 - len(planes))))
        cur = (ctypes.c_int32 * 6)()
        _lib.check(self._lib.fire_precision_plan(self._h, arr, len(planes), cur))
        return list(cur)

    def det_options(self, fp64_laplacian: bool):
        """Determinant-Laplacian products in FP64 (True) or on an FP32 copy of the inverse."""
        _lib.check(self._lib.fire_det_options(self._h, int(bool(fp64_laplacian))))

    def mcmc_options(self, use_graphs: bool = True):
        """Replay whole sweeps from a captured hipGraph (default) or issue every step eagerly;
        both give identical results."""
        _lib.check(self._lib.fire_mcmc_options(self._h, int(bool(use_graphs))))

    def state_stats(self) -> dict:
        """Counters of the state's single-electron updates: proposals and the mean / maximum
        number of electrons whose embedding changed per move."""
        nw, npr, s, mx = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self._lib.fire_state_stats(self._h, ctypes.byref(nw), ctypes.byref(npr), ctypes.byref(s),
                                              ctypes.byref(mx)))
        return dict(n_walkers=nw.value, n_proposals=npr.value, sum_changed=s.value,
                    mean_changed=(s.value / npr.value) if npr.value else 0.0, max_changed=mx.value)

    def mcmc_steps_low_rank(self, params, electrons, steps: int, width: float, seed: int = 0, step0: int = 0,
                            precision=None, keep_state: bool = False, refresh_sweeps: int = 0):
        """mcmc_steps_low_rank with proposal_single_electron (mcmc.py:100-111, 196-275):
        ``steps`` Metropolis steps, step s moving electron s mod n_el of every walker.
        Returns (electrons, stats) with stats["pmove"] the acceptance rate.  With
        ``keep_state`` the resident state of an earlier call (same walkers) is continued
        instead of being rebuilt from ``electrons`` (the reference rebuilds it at the start
        of every call); when the parameters changed since, the state is rebuilt from its own
        positions.  ``refresh_sweeps`` > 0 rebuilds the state from the current positions
        every that many sweeps within the call: the FP64 low-rank updates accumulate rounding
        in the inverse, which long chains on ill-conditioned walkers make visible.  Accept
        counts and the changed-electron statistics cover the whole call."""
        if keep_state and self._committed_gen is not None:
            if params is not self._uploaded or _fingerprint(params) != self._uploaded_fp:
                r_res = self.state_values()[0]
                self._ensure(params)
                self.log_psi_with_state(params, r_res, precision=precision)
            elif self._pending_gen is not None:
                self._pending_gen = None  # a pending proposal is dropped by the next step
        else:
            self._ensure(params)
            self.log_psi_with_state(params, electrons, precision=precision)
        B = self._state_walkers()
        nacc = torch.zeros(B, dtype=torch.int32, device=self.device)
        nacc_total = torch.zeros(B, dtype=torch.int64, device=self.device)
        block = int(steps) if refresh_sweeps <= 0 else int(refresh_sweeps) * self.cfg.n_el
        done = 0
        n_prop = sum_changed = max_changed = 0
        while done < steps:
            if done > 0:
                st = self.state_stats()
                n_prop += st["n_proposals"]
                sum_changed += st["sum_changed"]
                max_changed = max(max_changed, st["max_changed"])
                self.log_psi_with_state(params, self.state_values()[0], precision=precision)
            n = min(block, int(steps) - done)
            _lib.check(self._lib.fire_mcmc_steps(self._h, _stream_ptr(), n, float(width), int(seed) & (2**64 - 1),
                                                 int(step0) + done, _dptr(nacc)))
            nacc_total += nacc.long()
            done += n
        r, sg, lp = self.state_values()
        st = self.state_stats()
        n_prop += st["n_proposals"]
        sum_changed += st["sum_changed"]
        max_changed = max(max_changed, st["max_changed"])
        pmove = float(nacc_total.double().mean() / steps) if steps else 0.0
        return r, dict(pmove=pmove, n_accept=nacc_total, logpsi=lp, sign=sg, n_walkers=B, n_proposals=n_prop,
                       mean_changed=(sum_changed / n_prop) if n_prop else 0.0, max_changed=max_changed)


def _fingerprint(params) -> int:
    """Content hash of a parameter tree (or ``PackedParams``): detects in-place changes of
    parameters already on the device."""
    h = xxhash.xxh3_64() if xxhash is not None else hashlib.blake2b(digest_size=8)

    def visit(x):
        if isinstance(x, dict):
            for k in x:
                h.update(str(k).encode())
                visit(x[k])
        elif isinstance(x, (list, tuple)):
            for v in x:
                visit(v)
        elif isinstance(x, PackedParams):
            visit(x.flat)
            visit(x.sizes)
        else:
            if isinstance(x, torch.Tensor):
                x = x.detach().cpu().numpy()
            a = np.ascontiguousarray(np.asarray(x))
            h.update(str(a.dtype).encode() + str(a.shape).encode())
            h.update(a.view(np.uint8).reshape(-1).data if a.size else b"")

    visit(params)
    return int.from_bytes(h.digest()[:8], "little")


@dataclass(eq=False)
class PackedParams:
    """Parameters already in the canonical packed form of ``fire_upload_params``: float32
    tensors back to back (``flat``) and their element counts (``sizes``)."""

    flat: np.ndarray
    sizes: np.ndarray


@dataclass
class WalkerState:
    """Names the GPU-resident state of a ``FireWaveFunction`` (committed, or a pending
    proposal) -- the reference's model state pytree (wave_function.py:226-287)."""

    wf: FireWaveFunction
    gen: int
    n_walkers: int
    single: bool
    pending: bool


def make_local_energy(wf: FireWaveFunction, energy_operator: str = "sparse", precision: Optional[str] = None,
                      pseudopotential=None):
    """make_local_energy (hamiltonian.py:49-81): returns
    ``local_energy(key, params, electrons, static) -> (E_loc, static)``.  With a
    ``pseudopotential`` (pseudopotentials.make_pseudopotential) E_loc adds its local and
    non-local parts, the Coulomb terms use the effective charges, and ``key`` (an integer)
    seeds the quadrature rotations."""
    if energy_operator.lower() != "sparse":
        raise ValueError("only the sparse forward-Laplacian energy operator is provided")
    if pseudopotential is not None:
        wf.set_pseudopotential(pseudopotential)

    def local_energy(key, params, electrons, static=None):
        terms = wf.local_energy_terms(params, electrons, static, precision)
        e = terms["e_loc"]
        if pseudopotential is not None and pseudopotential.n_ecp:
            seed = int(np.asarray(key).reshape(-1).astype(np.uint64).sum()) if key is not None else 0
            pp = wf.pseudopotential_energy(params, electrons, seed, precision)
            e = e + pp["v_loc"] + pp["v_nl"]
        return e, static

    return local_energy
