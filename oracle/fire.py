"""CPU restatement of FiRE's per-walker hot path (oracle; test infrastructure only).

Every function follows the reference file:line given in its docstring.  The arithmetic runs
in the dtype of the parameters handed in (float64 for the FP64 oracle, float32 for the
plain-FP32 restatement).  Neighbour lists are always built from float32 positions with the
float32 expression ``sqrt((dx*dx + dy*dy) + dz*dz) < cutoff`` so that every precision uses
the same graph.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import fwdlap as fl
from .fwdlap import FL

NO_NEIGHBOUR = 1_000_000  # model/graph_utils.py:14


# -----------------------------------------------------------------------------------------
# parameters
# -----------------------------------------------------------------------------------------
def to_torch(params, dtype=torch.float64):
    if isinstance(params, dict):
        return {k: to_torch(v, dtype) for k, v in params.items()}
    if isinstance(params, (list, tuple)):
        return [to_torch(v, dtype) for v in params]
    return torch.as_tensor(np.asarray(params, dtype=np.float64)).to(dtype)


def softplus(x):
    return torch.nn.functional.softplus(x)


# -----------------------------------------------------------------------------------------
# a1: neighbour lists
# -----------------------------------------------------------------------------------------
def pair_distances_f32(r: np.ndarray) -> np.ndarray:
    """float32 e-e distances, diagonal = +inf (is_neighbour, graph_utils.py:179-184;
    get_pair_indices, new_sparse_model.py:238-245)."""
    r = np.asarray(r, np.float32)
    d = r[:, None, :] - r[None, :, :]
    sq = d * d
    d2 = (sq[..., 0] + sq[..., 1]) + sq[..., 2]
    dist = np.sqrt(d2)
    np.fill_diagonal(dist, np.inf)
    return dist


@dataclass
class NeighbourLists:
    offsets: np.ndarray  # [N+1] CSR row pointer over centres
    nbr: np.ndarray  # [P] neighbour index, row-major (centre asc, neighbour asc)
    ctr: np.ndarray  # [P] centre index of each pair
    same: np.ndarray  # [P] bool: same spin
    jac_idx: np.ndarray  # [P] jacobian row of pair p (= N + p)
    en_offsets: np.ndarray  # [N+1] CSR over electrons of neighbouring nuclei
    en_idx: np.ndarray  # nuclei within cutoff_1el, ascending per electron

    @property
    def n_pairs(self):
        return len(self.nbr)


def neighbour_lists(r, n_up, cutoff, R=None, cutoff_1el=None) -> NeighbourLists:
    """Pair lists of get_pair_indices (new_sparse_model.py:238-263): strict ``d < cutoff``
    on the *static* cutoff, pairs in ``jnp.where`` row-major order, joint list = jacobian
    rows N..N+P-1; spin split at n_up.  e-n list of get_neighbouring_nuclei (:476-481)."""
    r = np.asarray(r, np.float32)
    N = r.shape[0]
    dist = pair_distances_f32(r)
    inc = dist < np.float32(cutoff)
    ctr, nbr = np.nonzero(inc)
    counts = inc.sum(1)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    spin = np.arange(N) < n_up
    same = spin[ctr] == spin[nbr]
    jac_idx = N + np.arange(len(ctr))
    if R is not None:
        Rf = np.asarray(R, np.float32)
        dd = r[:, None, :] - Rf[None, :, :]
        sq = dd * dd
        den = np.sqrt((sq[..., 0] + sq[..., 1]) + sq[..., 2])
        inen = den < np.float32(cutoff_1el)
        e_c, e_j = np.nonzero(inen)
        en_offsets = np.concatenate([[0], np.cumsum(inen.sum(1))]).astype(np.int64)
        en_idx = e_j
    else:
        en_offsets = np.zeros(N + 1, np.int64)
        en_idx = np.zeros(0, np.int64)
    return NeighbourLists(offsets, nbr.astype(np.int64), ctr.astype(np.int64), same, jac_idx, en_offsets, en_idx)


def reference_pair_lists(r, n_up, cutoff):
    """The padded-list outputs of get_pair_indices exactly as the reference returns them
    (same list, diff list, jac_idx via searchsorted, joint jac indices), unpadded."""
    nl = neighbour_lists(r, n_up, cutoff)
    N = len(r)
    key = nl.ctr * N + nl.nbr
    s = nl.same
    same = (nl.ctr[s], nl.nbr[s], np.searchsorted(key, nl.ctr[s] * N + nl.nbr[s]) + N)
    diff = (nl.ctr[~s], nl.nbr[~s], np.searchsorted(key, nl.ctr[~s] * N + nl.nbr[~s]) + N)
    jac = (np.concatenate([np.arange(N), nl.ctr]), np.concatenate([np.arange(N), nl.nbr]))
    return same, diff, jac


def static_input(r_old, n_up, cutoff, R, cutoff_1el, r_new=None, idx_changed=None, laplacian=True):
    """get_static_input (new_sparse_model.py:399-437) of one walker, as a dict of the
    StaticArgs fields: pair counts by spin, triplets sum_{a,b,c} in(a,c) in(b,c), the largest
    number of nuclei within cutoff_1el of an electron, and for a move of electron
    ``idx_changed`` from r_old to r_new the affected electrons (neighbours of the moved one
    before or after, plus itself) and the pairs of the new graph centred on them."""
    N = len(r_old)
    r = np.asarray(r_old if r_new is None else r_new, np.float32)
    spin = np.arange(N) < n_up
    same = spin[:, None] == spin[None, :]
    inc = pair_distances_f32(r) < np.float32(cutoff)
    dd = r[:, None, :] - np.asarray(R, np.float32)[None]
    sq = dd * dd
    den = np.sqrt((sq[..., 0] + sq[..., 1]) + sq[..., 2])
    out = dict(n_pairs_same=int((inc & same).sum()), n_pairs_diff=int((inc & ~same).sum()),
               n_triplets=int((inc[:, None, :] & inc[None, :, :]).sum()) if laplacian else 0,
               n_neighbours_en=int((den < np.float32(cutoff_1el)).sum(1).max()),
               n_changed_hout=0, n_changed_pair_same=0, n_changed_pair_diff=0, n_pp_elecs=N)
    if r_new is not None:
        inc_old = pair_distances_f32(np.asarray(r_old, np.float32)) < np.float32(cutoff)
        aff = inc[:, idx_changed] | inc_old[:, idx_changed]
        aff[idx_changed] = True
        req = inc & aff[:, None]
        out.update(n_changed_hout=int(aff.sum()), n_changed_pair_same=int((req & same).sum()),
                   n_changed_pair_diff=int((req & ~same).sum()))
    return out


def n_triplets(r, cutoff):
    """get_static_input n_triplets (new_sparse_model.py:412)."""
    inc = pair_distances_f32(r) < np.float32(cutoff)
    return int((inc[:, None, :] & inc[None, :, :]).sum())


# -----------------------------------------------------------------------------------------
# a2: ElecInit (h0 + 4 message vectors); generic over Tensor / FL
# -----------------------------------------------------------------------------------------
def _diff_features(diff):
    """get_diff_features (utils.py:436-449): [|d|, d]."""
    d = fl.norm_last(diff)
    if isinstance(diff, FL):
        return fl.concat([d.unsqueeze(-1), diff], -1 + diff.x.dim()), d
    return torch.cat([d.unsqueeze(-1), diff], -1), d


def _log_scaled(feat, d):
    """features / features[..., :1] * log1p(features[..., :1]) (new_sparse_model.py:156,180)."""
    dd = d.unsqueeze(-1)
    return feat / dd * fl.log1p(dd)


def layer_norm(h, scale, bias, eps=1e-6):
    """flax ``nn.LayerNorm()`` over the feature axis (biased variance, eps = 1e-6 inside the
    square root), as used by ElecInit (new_sparse_model.py:209)."""
    F = h.shape[-1]
    mu = h.sum(-1) / F
    c = h - mu.unsqueeze(-1)
    var = fl.square(c).sum(-1) / F
    inv = fl.reciprocal(fl.sqrt(var + eps))
    return c * inv.unsqueeze(-1) * scale + bias


def glu(h, kernel):
    """GatedLinearUnit (utils.py:55-62): Dense(2F, no bias), split into (x, gate), x * silu(gate)."""
    F = kernel.shape[-1] // 2
    g2 = fl.linear(h, kernel)
    return g2[:, :F] * fl.silu(g2[:, F:])


def en_edge_sum(p, cfg, r, R, en_mask):
    """sum_J Gamma_J * e_J of ElecInit (new_sparse_model.py:204-206) over the nuclei selected
    by ``en_mask`` [N,M], with ElecNucEdge (:139-158) and the dynamic e-n PairwiseFilter
    (utils.py:393-433).  ``r``: [N,3] (Tensor, or FL whose tangents are per-electron)."""
    ep = p["embedding"]
    dyn = ep["dynamic_params_en"]
    edge = ep["elec_init"]["params"]["ElecNucEdge_0"]
    pf = edge["PairwiseFilter_0"]
    diff = r.unsqueeze(1) - R.unsqueeze(0) if isinstance(r, FL) else r[:, None, :] - R[None, :, :]
    feat, d = _diff_features(diff)  # [N,M,4], [N,M]
    # directional MLP with nucleus-dependent kernel/bias
    pre = (feat.unsqueeze(-1) * dyn["filter"]["kernel"]).sum(-2) + dyn["filter"]["bias"]
    dirf = fl.linear(fl.silu(pre), pf["Dense_0"]["kernel"], pf["Dense_0"]["bias"])
    scales = softplus(dyn["filter"]["scales"])  # [M,E]
    x = d.unsqueeze(-1) / scales
    env = fl.exp(-fl.square(x)) * fl.cutoff_poly(d / cfg.cutoff_1el).unsqueeze(-1)
    env = fl.linear(env, pf["Dense_1"]["kernel"])
    beta = dirf * env
    gamma = fl.linear(beta, edge["Dense_0"]["kernel"])
    e = fl.linear(_log_scaled(feat, d), edge["Dense_1"]["kernel"]) + dyn["nuc_embedding"]
    m = en_mask.to(gamma.dtype).unsqueeze(-1)
    return (gamma * m * e).sum(1)  # [N,F]


def elec_init_mlp(ei, h):
    """The dense stage of ElecInit (new_sparse_model.py:207-215) on the e-n sum h [N,F]:
    LayerNorm, GatedLinearUnit, silu(Dense_F), Dense_5F split into
    (h0, msg_ct_same, msg_ct_diff, msg_nb_same, msg_nb_diff) -- the order of :213-215 as
    unpacked by NewSparseEmbedding.apply (:516)."""
    F = h.shape[-1]
    h = layer_norm(h, ei["LayerNorm_0"]["scale"], ei["LayerNorm_0"]["bias"])
    h = glu(h, ei["GatedLinearUnit_0"]["Dense_0"]["kernel"])
    h = fl.silu(fl.linear(h, ei["Dense_0"]["kernel"], ei["Dense_0"]["bias"]))
    out = fl.linear(h, ei["Dense_1"]["kernel"], ei["Dense_1"]["bias"])
    return [out[:, k * F:(k + 1) * F] for k in range(5)]


def elec_init(p, cfg, r, R, en_mask, learned=None):
    """ElecInit.__call__ (new_sparse_model.py:190-215): the e-n edge sum followed by the
    dense stage."""
    h = en_edge_sum(p, cfg, r, R, en_mask)
    return elec_init_mlp(p["embedding"]["elec_init"]["params"], h)


# -----------------------------------------------------------------------------------------
# a3: e-e edges
# -----------------------------------------------------------------------------------------
def learned_cutoff(p, cfg):
    """cutoff * sigmoid(cutoff_param) (new_sparse_model.py:506,593,644)."""
    return cfg.cutoff * torch.sigmoid(p["embedding"]["cutoff_param"])


def ee_edge(ep, cut, diff):
    """ElecElecEdges.__call__ (new_sparse_model.py:161-187) with the static PairwiseFilter
    ``beta_ee`` (utils.py:393-433).  ``diff`` = r_ctr - r_nb, [P,3]."""
    q = ep["params"]
    bf = q["beta_ee"]
    feat, d = _diff_features(diff)
    pre = fl.linear(feat, bf["kernel"], bf["bias"])
    dirf = fl.linear(fl.silu(pre), bf["Dense_0"]["kernel"], bf["Dense_0"]["bias"])
    x = d.unsqueeze(-1) / softplus(bf["scale"])
    env = fl.exp(-fl.square(x)) * fl.cutoff_poly(d / cut).unsqueeze(-1)
    env = fl.linear(env, bf["Dense_1"]["kernel"])
    beta = dirf * env
    gamma = fl.linear(beta, q["Dense_0"]["kernel"])
    e = fl.linear(_log_scaled(feat, d), q["Dense_1"]["kernel"], q["Dense_1"]["bias"])
    return gamma, e


# -----------------------------------------------------------------------------------------
# value path: embedding, orbitals, determinants, Jastrow
# -----------------------------------------------------------------------------------------
def _en_mask(r32, R, cutoff_1el):
    rr = np.asarray(r32, np.float32)
    dd = rr[:, None, :] - np.asarray(R, np.float32)[None]
    sq = dd * dd
    den = np.sqrt((sq[..., 0] + sq[..., 1]) + sq[..., 2])
    return torch.as_tensor(den < np.float32(cutoff_1el))


def embedding_apply(p, cfg, r32, return_stages=False, return_state=False):
    """NewSparseEmbedding.apply (new_sparse_model.py:483-551)."""
    dtype = p["embedding"]["cutoff_param"].dtype
    r = torch.as_tensor(np.asarray(r32, np.float64)).to(dtype)
    R = torch.as_tensor(np.asarray(cfg.mol.R, np.float64)).to(dtype)
    nl = neighbour_lists(r32, cfg.n_up, cfg.cutoff)
    h0, mcs, mcd, mns, mnd = elec_init(p, cfg, r, R, _en_mask(r32, cfg.mol.R, cfg.cutoff_1el))
    cut = learned_cutoff(p, cfg)
    h = h0.clone()
    ep = p["embedding"]
    for same, edge_p, mct, mnb in ((True, ep["edge_same"], mcs, mns), (False, ep["edge_diff"], mcd, mnd)):
        sel = nl.same == same
        ci, ni = torch.as_tensor(nl.ctr[sel]), torch.as_tensor(nl.nbr[sel])
        if len(ci) == 0:
            continue
        gamma, e = ee_edge(edge_p, cut, r[ci] - r[ni])
        msg = gamma * fl.silu(e + mct[ci] + mnb[ni])
        h = h.index_add(0, ci, msg)
    h_contr = h
    s0, s1 = ep["scales"]
    up = ep["updates"][0]["params"]
    x = h * s0
    x = fl.silu(fl.linear(x, up["Linear_0"]["kernel"], up["Linear_0"]["bias"]))
    x = fl.silu(fl.linear(x, up["Linear_1"]["kernel"], up["Linear_1"]["bias"]))
    h_mlp = x
    hout = x * s1
    if return_stages:
        return h_contr, h_mlp
    if return_state:
        return hout, dict(electrons=r, h0=h0, msg_ct_same=mcs, msg_ct_diff=mcd, msg_nb_same=mns,
                          msg_nb_diff=mnd, h_out=hout)
    return hout


def _swap_updn(phi, N, n_up):
    """_swap_spin_updn (orbitals.py:21-26) on the last axis laid out as (det, orb)."""
    sh = phi.shape
    phi = phi.reshape(*sh[:-1], -1, N)
    phi = torch.cat([phi[..., n_up:], phi[..., :n_up]], -1)
    return phi.reshape(sh)


def envelopes_iso(ep, cfg, r, R):
    """EfficientIsotropicEnvelopes (envelopes.py:55-82): sum_n sum_e pi exp(-|r-R_n| softplus(sigma))
    -> [..., D*N].  ``r``: Tensor [n,3] or FL with per-electron tangents."""
    diff = r.unsqueeze(1) - R.unsqueeze(0) if isinstance(r, FL) else r[:, None, :] - R[None, :, :]
    d = fl.norm_last(diff)  # [n,M]
    sig = softplus(ep["sigma"])  # [M,D,E]
    x = d.unsqueeze(-1).unsqueeze(-1) * sig  # [n,M,D,E]
    env = fl.exp(-x)
    pi = ep["pi"]  # [M,D,E,O]
    M, D, E, O = pi.shape
    if isinstance(env, FL):
        e2 = env.reshape(env.shape[0], M, D * E)  # n, M, (D E)
        # out[n,d,o] = sum_{m,e} env[n,m,d,e] pi[m,d,e,o]
        outs = []
        for dd in range(D):
            part = env[:, :, dd, :].reshape(env.shape[0], M * E)  # [n, M*E]
            w = pi[:, dd].reshape(M * E, O)
            outs.append(part @ w)
        out = fl.concat(outs, 1)
        del e2
        return out
    out = torch.einsum("nmde,mdeo->ndo", env, pi)
    return out.reshape(out.shape[0], -1)


def orbitals_apply(p, cfg, r32, h, return_state=False):
    """Orbitals.__call__ (orbitals.py:109-119): Phi[d,i,o] = (W h_i + b)[d,o] * env_i[d,o]."""
    dtype = h.dtype
    r = torch.as_tensor(np.asarray(r32, np.float64)).to(dtype)
    R = torch.as_tensor(np.asarray(cfg.mol.R, np.float64)).to(dtype)
    op = p["to_orbitals"]["params"]
    N, n_up, D = cfg.n_el, cfg.n_up, cfg.n_determinants
    up = op["to_orbitals_up"]
    orb_up = h[:n_up] @ up["kernel"] + up["bias"]
    env_up = envelopes_iso(op["env_up"], cfg, r[:n_up], R)
    if cfg.spin_restricted:
        orb_dn = _swap_updn(h[n_up:] @ up["kernel"] + up["bias"], N, n_up)
        env_dn = _swap_updn(envelopes_iso(op["env_up"], cfg, r[n_up:], R), N, n_up)
    else:
        dn = op["to_orbitals_dn"]
        orb_dn = h[n_up:] @ dn["kernel"] + dn["bias"]
        env_dn = envelopes_iso(op["env_dn"], cfg, r[n_up:], R)
    orb_raw = torch.cat([orb_up, orb_dn], 0)
    env = torch.cat([env_up, env_dn], 0)
    phi = (orb_raw * env).reshape(N, D, N).permute(1, 0, 2).contiguous()
    if return_state:
        return phi, dict(envelopes=env, orbitals=orb_raw)
    return phi


def slog_and_inverse(A):
    """slog_and_inverse (utils.py:164-175): LU-based sign, log|det| and inverse."""
    sign, logdet = torch.linalg.slogdet(A)
    return (sign, logdet), torch.linalg.inv(A)


def logsumexp_signed(logdets, signs):
    """jax.nn.logsumexp(logdet, b=sign, return_sign=True) (utils.py:210)."""
    m = logdets.max()
    s = (signs * torch.exp(logdets - m)).sum()
    return torch.sign(s), m + torch.log(torch.abs(s))


def _all_pairs(N, n_up):
    """get_dist_same_diff pair sets (utils.py:309-322): same-spin i<j, and up x down."""
    iu = np.triu_indices(n_up, 1)
    idn = np.triu_indices(N - n_up, 1)
    same_i = np.concatenate([iu[0], idn[0] + n_up])
    same_j = np.concatenate([iu[1], idn[1] + n_up])
    gi, gj = np.meshgrid(np.arange(n_up), np.arange(n_up, N), indexing="ij")
    return (same_i, same_j), (gi.reshape(-1), gj.reshape(-1))


def cusp_terms(jp, cfg, diff_same, diff_diff):
    """ElElCusp (jastrow.py:48-64) on pair differences (Tensor or FL)."""
    c = jp["pairwise_cusps"]
    out = 0.0
    for diff, a, f in ((diff_same, c["alpha_same"], c["factor_same"]), (diff_diff, c["alpha_diff"], c["factor_diff"])):
        if diff is None:
            continue
        d = fl.norm_last(diff)
        term = fl.reciprocal(d + a) * (a * a)
        out = out + term.sum(0) * f
    return out


def jastrow_mlp_layers(jp, cfg):
    mlp = jp["mlp"]
    return [(mlp[f"Linear_{i}"]["kernel"], mlp[f"Linear_{i}"]["bias"]) for i in range(cfg.mlp_depth + 1)]


def att_out_layers(jp, cfg):
    out = jp["att"]["out"]
    return [(out[f"Dense_{i}"]["kernel"], out[f"Dense_{i}"]["bias"]) for i in range(cfg.mlp_depth + 1)]


def att_readout(jp, cfg, norm, values):
    """GlobalAttentionJastrow.readout (jastrow.py:93-100): MLP([W]*depth + [2]) on
    values/normaliser, silu on hidden layers only, then * scale + [0, bias]."""
    x = (values / norm.unsqueeze(-1)).reshape(-1)
    layers = att_out_layers(jp, cfg)
    for k, (W, b) in enumerate(layers):
        x = fl.linear(x, W, b)
        if k < len(layers) - 1:
            x = fl.silu(x)
    a = jp["att"]
    zb = torch.cat([torch.zeros(1, dtype=a["bias"].dtype), a["bias"]])
    return x * a["scale"] + zb


def attention_norm_values(jp, cfg, h):
    """GlobalAttentionJastrow.attention_and_values and contract (jastrow.py:84-90, 102-152),
    value path: per electron attention exp(-q.key / sqrt(Dk)) [N,H] and attention * values
    [N,H,Dk], and their per-spin sums (norm [2,H], values [2,H,Dk]; up first)."""
    a = jp["att"]
    N, n_up = h.shape[0], cfg.n_up
    H, Dk = cfg.attention_heads, cfg.attention_dim
    q = (h @ a["Q_W"]["kernel"] + a["Q_W"]["bias"]).reshape(N, H, Dk)
    v = (h @ a["V_W"]["kernel"] + a["V_W"]["bias"]).reshape(N, H, Dk)
    att = torch.exp(-(q * a["register_keys"]).sum(-1) / math.sqrt(Dk))
    vals = att.unsqueeze(-1) * v
    norm = torch.stack([att[:n_up].sum(0), att[n_up:].sum(0)])
    values = torch.stack([vals[:n_up].sum(0), vals[n_up:].sum(0)])
    return att, vals, norm, values


def mlp_to_logpsi(cfg, J):
    """Jastrow._mlp_to_logpsi (jastrow.py:372-382)."""
    logpsi = 0.0
    sign = 1.0
    if cfg.use_mlp_jastrow:
        logpsi = logpsi + J[0]
    if cfg.use_log_jastrow:
        Jx = J.x if isinstance(J, FL) else J
        sign = torch.sign(Jx[1])
        logpsi = logpsi + fl.log_abs(J[1])
    return sign, logpsi


def jastrow_apply(p, cfg, r, h, return_state=False):
    """Jastrow.__call__ (jastrow.py:261-307)."""
    jp = p["jastrow"]["params"]
    dtype = h.dtype
    N, n_up = cfg.n_el, cfg.n_up
    logpsi = torch.zeros((), dtype=dtype)
    if cfg.e_e_cusps == "psiformer":
        (si, sj), (di, dj) = _all_pairs(N, n_up)
        logpsi = logpsi + cusp_terms(jp, cfg, r[si] - r[sj], r[di] - r[dj])
    J = torch.zeros(2, dtype=dtype)
    att = vals = one_el = None
    if cfg.use_attention:
        att, vals, norm, values = attention_norm_values(jp, cfg, h)
        J = J + att_readout(jp, cfg, norm, values)
    if cfg.use_mlp_jastrow or cfg.use_log_jastrow:
        x = h
        for W, b in jastrow_mlp_layers(jp, cfg):
            x = fl.silu(x @ W + b)  # SparseMLP: silu after every layer (sparse_fwd_lap.py:219-224)
        zb = torch.stack([torch.zeros((), dtype=dtype), jp["log_bias"] if cfg.use_log_jastrow else torch.zeros((), dtype=dtype)])
        one_el = x * jp["mlp_scale"] + zb
        J = J + one_el.sum(0)
    sign, lj = mlp_to_logpsi(cfg, J)
    if return_state:
        return (sign, logpsi + lj), dict(one_el=one_el, attention=att, values=vals)
    return sign, logpsi + lj


def signed(p, cfg, r32):
    """MoonLikeWaveFunction.signed (wave_function.py:183-190)."""
    h = embedding_apply(p, cfg, r32)
    phi = orbitals_apply(p, cfg, r32, h)
    sl, ld = torch.linalg.slogdet(phi)
    sgn, lp = logsumexp_signed(ld, sl)
    r = torch.as_tensor(np.asarray(r32, np.float64)).to(h.dtype)
    sJ, lJ = jastrow_apply(p, cfg, r, h)
    return sgn * sJ, lp + lJ


def log_psi_value_torch(p, cfg, r, nl_r32):
    """logpsi as a differentiable torch function of ``r`` (float64) with the neighbour graph
    frozen at the float32 positions ``nl_r32`` -- used for the dense Hessian-trace check."""
    return _signed_torch(p, cfg, r, nl_r32)[1]


def _signed_torch_embedding(p, cfg, r, r32):
    """Embedding h_out as a differentiable function of ``r`` (graph frozen at ``r32``)."""
    return _embed_torch(p, cfg, r, r32)


def _orbitals_torch(p, cfg, r, r32):
    """Slater matrices [D, N, N] as a differentiable function of ``r``."""
    h = _embed_torch(p, cfg, r, r32)
    return _phi_torch(p, cfg, r, h)


def _embed_torch(p, cfg, r, r32):
    dtype = r.dtype
    R = torch.as_tensor(np.asarray(cfg.mol.R, np.float64)).to(dtype)
    nl = neighbour_lists(r32, cfg.n_up, cfg.cutoff)
    h0, mcs, mcd, mns, mnd = elec_init(p, cfg, r, R, _en_mask(r32, cfg.mol.R, cfg.cutoff_1el))
    cut = learned_cutoff(p, cfg)
    h = h0
    ep = p["embedding"]
    for same, edge_p, mct, mnb in ((True, ep["edge_same"], mcs, mns), (False, ep["edge_diff"], mcd, mnd)):
        sel = nl.same == same
        ci, ni = torch.as_tensor(nl.ctr[sel]), torch.as_tensor(nl.nbr[sel])
        if len(ci) == 0:
            continue
        gamma, e = ee_edge(edge_p, cut, r[ci] - r[ni])
        h = h.index_add(0, ci, gamma * fl.silu(e + mct[ci] + mnb[ni]))
    s0, s1 = ep["scales"]
    up = ep["updates"][0]["params"]
    x = fl.silu(fl.linear(h * s0, up["Linear_0"]["kernel"], up["Linear_0"]["bias"]))
    x = fl.silu(fl.linear(x, up["Linear_1"]["kernel"], up["Linear_1"]["bias"]))
    return x * s1


def _phi_torch(p, cfg, r, h):
    dtype = r.dtype
    R = torch.as_tensor(np.asarray(cfg.mol.R, np.float64)).to(dtype)
    op = p["to_orbitals"]["params"]
    N, n_up, D = cfg.n_el, cfg.n_up, cfg.n_determinants
    upw = op["to_orbitals_up"]
    orb_up = h[:n_up] @ upw["kernel"] + upw["bias"]
    env_up = envelopes_iso(op["env_up"], cfg, r[:n_up], R)
    if cfg.spin_restricted:
        orb_dn = _swap_updn(h[n_up:] @ upw["kernel"] + upw["bias"], N, n_up)
        env_dn = _swap_updn(envelopes_iso(op["env_up"], cfg, r[n_up:], R), N, n_up)
    else:
        dn = op["to_orbitals_dn"]
        orb_dn = h[n_up:] @ dn["kernel"] + dn["bias"]
        env_dn = envelopes_iso(op["env_dn"], cfg, r[n_up:], R)
    return (torch.cat([orb_up, orb_dn]) * torch.cat([env_up, env_dn])).reshape(N, D, N).permute(1, 0, 2)


def _signed_torch(p, cfg, r, r32):
    dtype = r.dtype
    R = torch.as_tensor(np.asarray(cfg.mol.R, np.float64)).to(dtype)
    nl = neighbour_lists(r32, cfg.n_up, cfg.cutoff)
    h0, mcs, mcd, mns, mnd = elec_init(p, cfg, r, R, _en_mask(r32, cfg.mol.R, cfg.cutoff_1el))
    cut = learned_cutoff(p, cfg)
    h = h0
    ep = p["embedding"]
    for same, edge_p, mct, mnb in ((True, ep["edge_same"], mcs, mns), (False, ep["edge_diff"], mcd, mnd)):
        sel = nl.same == same
        ci, ni = torch.as_tensor(nl.ctr[sel]), torch.as_tensor(nl.nbr[sel])
        if len(ci) == 0:
            continue
        gamma, e = ee_edge(edge_p, cut, r[ci] - r[ni])
        h = h.index_add(0, ci, gamma * fl.silu(e + mct[ci] + mnb[ni]))
    s0, s1 = ep["scales"]
    up = ep["updates"][0]["params"]
    x = fl.silu(fl.linear(h * s0, up["Linear_0"]["kernel"], up["Linear_0"]["bias"]))
    x = fl.silu(fl.linear(x, up["Linear_1"]["kernel"], up["Linear_1"]["bias"]))
    h = x * s1
    op = p["to_orbitals"]["params"]
    N, n_up, D = cfg.n_el, cfg.n_up, cfg.n_determinants
    upw = op["to_orbitals_up"]
    orb_up = h[:n_up] @ upw["kernel"] + upw["bias"]
    env_up = envelopes_iso(op["env_up"], cfg, r[:n_up], R)
    if cfg.spin_restricted:
        orb_dn = _swap_updn(h[n_up:] @ upw["kernel"] + upw["bias"], N, n_up)
        env_dn = _swap_updn(envelopes_iso(op["env_up"], cfg, r[n_up:], R), N, n_up)
    else:
        dn = op["to_orbitals_dn"]
        orb_dn = h[n_up:] @ dn["kernel"] + dn["bias"]
        env_dn = envelopes_iso(op["env_dn"], cfg, r[n_up:], R)
    phi = (torch.cat([orb_up, orb_dn]) * torch.cat([env_up, env_dn])).reshape(N, D, N).permute(1, 0, 2)
    sl, ld = torch.linalg.slogdet(phi)
    sgn, lp = logsumexp_signed(ld, sl)
    sJ, lJ = jastrow_apply(p, cfg, r, h)
    return sgn * sJ, lp + lJ


def potential_energy(r, R, Z):
    """potential_energy (hamiltonian.py:15-26): e-e, e-n and n-n Coulomb terms, float64."""
    r = np.asarray(r, np.float64)
    R = np.asarray(R, np.float64)
    Z = np.asarray(Z, np.float64)
    N = len(r)
    dee = np.linalg.norm(r[:, None] - r[None], axis=-1)
    iu = np.triu_indices(N, 1)
    e_ee = np.sum(1.0 / dee[iu])
    den = np.linalg.norm(r[:, None] - R[None], axis=-1)
    e_en = -np.sum(Z[None] / den)
    M = len(R)
    inn = np.triu_indices(M, 1)
    dnn = np.linalg.norm(R[:, None] - R[None], axis=-1)
    e_nn = np.sum((Z[:, None] * Z[None])[inn] / dnn[inn]) if M > 1 else 0.0
    return e_ee + e_en + e_nn


# -----------------------------------------------------------------------------------------
# sparse forward-Laplacian path (a4-a9)
# -----------------------------------------------------------------------------------------
class Node:
    """NodeWithFwdLap (sparse_fwd_lap.py:14-126): x [N,...], jac [(N+P),3,...], lap [N,...];
    jacobian row k is d x[idx_ctr[k]] / d r[idx_dep[k]]; rows 0..N-1 are self rows."""

    def __init__(self, x, jac, lap, idx_ctr, idx_dep):
        self.x, self.jac, self.lap, self.idx_ctr, self.idx_dep = x, jac, lap, idx_ctr, idx_dep

    def _b(self, t):  # broadcast a per-centre tensor to jacobian rows
        return t[self.idx_ctr].unsqueeze(1)

    def jac_sqr(self):
        """get_jac_sqr (:24-28)."""
        out = torch.zeros_like(self.x)
        return out.index_add(0, self.idx_ctr, (self.jac ** 2).sum(1))

    def linear(self, W, b=None):
        """Linear on NodeWithFwdLap (:160-181): bias on the value only."""
        y = self.x @ W
        if b is not None:
            y = y + b
        return Node(y, self.jac @ W, self.lap @ W, self.idx_ctr, self.idx_dep)

    def elementwise(self, f, df, d2f):
        """ElementWise.apply_with_fwd_lap (:184-206)."""
        d1, d2 = df(self.x), d2f(self.x)
        return Node(f(self.x), self.jac * self._b(d1), self.lap * d1 + self.jac_sqr() * d2,
                    self.idx_ctr, self.idx_dep)

    def silu(self):
        s = torch.sigmoid
        return self.elementwise(lambda x: x * s(x), lambda x: s(x) * (1 + x * (1 - s(x))),
                                lambda x: s(x) * (1 - s(x)) * (2 + x * (1 - 2 * s(x))))

    def scale(self, c):
        return Node(self.x * c, self.jac * c, self.lap * c, self.idx_ctr, self.idx_dep)

    def add_const(self, c):
        return Node(self.x + c, self.jac, self.lap, self.idx_ctr, self.idx_dep)

    def mul(self, o: "Node"):
        """NodeWithFwdLap.__mul__ with another node of identical dependencies (:87-96)."""
        y = self.x * o.x
        jac = self.jac * self._b(o.x) + o.jac * self._b(self.x)
        lap = self.lap * o.x + self.x * o.lap
        lap = lap.index_add(0, self.idx_ctr, 2 * (self.jac * o.jac).sum(1))
        return Node(y, jac, lap, self.idx_ctr, self.idx_dep)

    def reshape_feat(self, *fshape):
        n, nr = self.x.shape[0], self.jac.shape[0]
        return Node(self.x.reshape(n, *fshape), self.jac.reshape(nr, 3, *fshape), self.lap.reshape(n, *fshape),
                    self.idx_ctr, self.idx_dep)

    def sum_feat_last(self):
        return Node(self.x.sum(-1), self.jac.sum(-1), self.lap.sum(-1), self.idx_ctr, self.idx_dep)

    def sum_from_to(self, lo, hi, n_el) -> FL:
        """sum_from_to (:47-56): sum over centres lo..hi-1, dense jacobian over all 3N coords."""
        x = self.x[lo:hi].sum(0)
        lap = self.lap[lo:hi].sum(0)
        mask = ((self.idx_ctr >= lo) & (self.idx_ctr < hi)).to(self.jac.dtype)
        mask = mask.reshape(-1, *([1] * (self.jac.dim() - 1)))
        jac = torch.zeros((n_el, 3) + tuple(self.x.shape[1:]), dtype=self.x.dtype)
        jac = jac.index_add(0, self.idx_dep, self.jac * mask)
        return FL(x, jac.reshape(n_el * 3, *self.x.shape[1:]), lap)

    def sum_over_nodes(self, n_el) -> FL:
        return self.sum_from_to(0, n_el, n_el)


def _per_electron_fl(r):
    """FL over electrons [N,3] whose tangent t of electron i is d/dr_i[t]."""
    N = r.shape[0]
    j = torch.zeros((3, N, 3), dtype=r.dtype)
    for t in range(3):
        j[t, :, t] = 1.0
    return FL(r, j, torch.zeros_like(r))


def _pair_fl(diff):
    """FL over pairs of r_ctr - r_nb with 6 tangents (3 centre, 3 neighbour)."""
    P = diff.shape[0]
    j = torch.zeros((6, P, 3), dtype=diff.dtype)
    for t in range(3):
        j[t, :, t] = 1.0
        j[3 + t, :, t] = -1.0
    return FL(diff, j, torch.zeros_like(diff))


def embedding_fwd_lap(p, cfg, r32, nl=None):
    """apply_with_fwd_lap (new_sparse_model.py:631-695) with contract_with_fwd_lap (:95-131)."""
    dtype = p["embedding"]["cutoff_param"].dtype
    r = torch.as_tensor(np.asarray(r32, np.float64)).to(dtype)
    R = torch.as_tensor(np.asarray(cfg.mol.R, np.float64)).to(dtype)
    N, F = cfg.n_el, cfg.feature_dim
    if nl is None:
        nl = neighbour_lists(r32, cfg.n_up, cfg.cutoff)
    P = nl.n_pairs
    h0, mcs, mcd, mns, mnd = elec_init(p, cfg, _per_electron_fl(r), R, _en_mask(r32, cfg.mol.R, cfg.cutoff_1el))
    idx_ctr = torch.as_tensor(np.concatenate([np.arange(N), nl.ctr]))
    idx_dep = torch.as_tensor(np.concatenate([np.arange(N), nl.nbr]))
    jac = torch.zeros((N + P, 3, F), dtype=dtype)
    jac[:N] = h0.j.permute(1, 0, 2)
    x, lap = h0.x.clone(), h0.l.clone()
    cut = learned_cutoff(p, cfg)
    ep = p["embedding"]
    for same, edge_p, mct, mnb in ((True, ep["edge_same"], mcs, mns), (False, ep["edge_diff"], mcd, mnd)):
        sel = nl.same == same
        if not sel.any():
            continue
        ci, ni = torch.as_tensor(nl.ctr[sel]), torch.as_tensor(nl.nbr[sel])
        jidx = torch.as_tensor(nl.jac_idx[sel])
        gamma, e = ee_edge(edge_p, cut, _pair_fl(r[ci] - r[ni]))
        u = e + mct[ci].pad_tangents(6, 0) + mnb[ni].pad_tangents(6, 3)
        msg = gamma * fl.silu(u)
        x = x.index_add(0, ci, msg.x)
        lap = lap.index_add(0, ci, msg.l)
        jac = jac.index_add(0, ci, msg.j[:3].permute(1, 0, 2))
        jac = jac.index_add(0, jidx, msg.j[3:].permute(1, 0, 2))
    h = Node(x, jac, lap, idx_ctr, idx_dep)
    s0, s1 = ep["scales"]
    up = ep["updates"][0]["params"]
    h = h.scale(s0)
    h = h.linear(up["Linear_0"]["kernel"], up["Linear_0"]["bias"]).silu()
    h = h.linear(up["Linear_1"]["kernel"], up["Linear_1"]["bias"]).silu()
    return h.scale(s1), nl


def _swap_node(node: Node, N, n_up):
    """_swap_spin_updn_with_fwd_lap (orbitals.py:29-40)."""
    x = _swap_updn(node.x, N, n_up)
    lap = _swap_updn(node.lap, N, n_up)
    js = _swap_updn(node.jac, N, n_up)
    keep = (node.idx_ctr < n_up).reshape(-1, 1, 1)
    jac = torch.where(keep, node.jac, js)
    return Node(x, jac, lap, node.idx_ctr, node.idx_dep)


def _swap_fl(a: FL, N, n_up):
    return FL(_swap_updn(a.x, N, n_up), _swap_updn(a.j, N, n_up), _swap_updn(a.l, N, n_up))


def orbitals_fwd_lap(p, cfg, r32, h: Node):
    """Orbitals.fwd_lap, sparse branch (orbitals.py:148-174): merge_up_down
    (sparse_fwd_lap.py:129-137) and multiply_with_1el_fn (:322-335).
    Returns x [D,N,N], jac [D,N+P,3,N], lap [D,N,N]."""
    dtype = h.x.dtype
    r = torch.as_tensor(np.asarray(r32, np.float64)).to(dtype)
    R = torch.as_tensor(np.asarray(cfg.mol.R, np.float64)).to(dtype)
    op = p["to_orbitals"]["params"]
    N, n_up, D = cfg.n_el, cfg.n_up, cfg.n_determinants
    up = op["to_orbitals_up"]
    rfl = _per_electron_fl(r)
    env_up = envelopes_iso(op["env_up"], cfg, rfl[:n_up], R)
    orb_up = h.linear(up["kernel"], up["bias"])
    if cfg.spin_restricted:
        env_dn = _swap_fl(envelopes_iso(op["env_up"], cfg, rfl[n_up:], R), N, n_up)
        orb_dn = _swap_node(orb_up, N, n_up)
    else:
        dn = op["to_orbitals_dn"]
        env_dn = envelopes_iso(op["env_dn"], cfg, rfl[n_up:], R)
        orb_dn = h.linear(dn["kernel"], dn["bias"])
    env = fl.concat([env_up, env_dn], 0)  # x [N, DN], j [3, N, DN]
    # merge_up_down
    xo = torch.cat([orb_up.x[:n_up], orb_dn.x[n_up:]], 0)
    lo = torch.cat([orb_up.lap[:n_up], orb_dn.lap[n_up:]], 0)
    jo = torch.where((h.idx_ctr < n_up).reshape(-1, 1, 1), orb_up.jac, orb_dn.jac)
    # multiply_with_1el_fn
    jb = env.j.permute(1, 0, 2)  # [N,3,DN]
    out = xo * env.x
    jac = torch.zeros_like(jo)
    jac[:N] = xo[:, None, :] * jb + env.x[:, None, :] * jo[:N]
    ctrp = h.idx_ctr[N:]
    jac[N:] = env.x[ctrp][:, None, :] * jo[N:]
    lap = lo * env.x + env.l * xo + 2 * (jo[:N] * jb).sum(1)
    phi = out.reshape(N, D, N).permute(1, 0, 2)
    lapd = lap.reshape(N, D, N).permute(1, 0, 2)
    jacd = jac.reshape(-1, 3, D, N).permute(2, 0, 1, 3)
    return phi, jacd, lapd


def sparse_slogdet_fwd_lap(x, jac, lap, idx_ctr, idx_dep, N):
    """sparse_slogdet_with_fwd_lap (sparse_fwd_lap.py:294-319): sign, logdet, gradient [3N],
    Laplacian, with the triplet sum over (i,k,n), i,k in {n} U N(n)."""
    (sign, logdet), Ainv = slog_and_inverse(x)
    M = jac @ Ainv  # [(N+P),3,N]
    rows = torch.arange(jac.shape[0])
    g = torch.zeros((N, 3), dtype=x.dtype).index_add(0, idx_dep, M[rows, :, idx_ctr])
    lapd = torch.einsum("ij,ji->", lap, Ainv)
    trip = torch.zeros((), dtype=x.dtype)
    dep = idx_dep.numpy()
    order = np.argsort(dep, kind="stable")
    bounds = np.searchsorted(dep[order], np.arange(N + 1))
    for n in range(N):
        rws = torch.as_tensor(order[bounds[n]:bounds[n + 1]])
        c = idx_ctr[rws]
        B = M[rws][:, :, c]  # [r,3,r]
        trip = trip + (B * B.permute(2, 1, 0)).sum()
    return sign, FL(logdet, g.reshape(3 * N), lapd - trip)


def _cusp_fwd_lap(jp, cfg, r):
    """Pairwise cusp with forward-Laplacian (jastrow.py:397-403): per pair 6 tangents,
    scattered into a dense [3N] jacobian."""
    N, n_up = cfg.n_el, cfg.n_up
    (si, sj), (di, dj) = _all_pairs(N, n_up)
    c = jp["pairwise_cusps"]
    g = torch.zeros((N, 3), dtype=r.dtype)
    val = torch.zeros((), dtype=r.dtype)
    lap = torch.zeros((), dtype=r.dtype)
    for (ii, jj), a, f in (((si, sj), c["alpha_same"], c["factor_same"]), ((di, dj), c["alpha_diff"], c["factor_diff"])):
        if len(ii) == 0:
            continue
        ii_t, jj_t = torch.as_tensor(ii), torch.as_tensor(jj)
        d = fl.norm_last(_pair_fl(r[ii_t] - r[jj_t]))
        t = fl.reciprocal(d + a) * (a * a * f)
        val = val + t.x.sum()
        lap = lap + t.l.sum()
        g = g.index_add(0, ii_t, t.j[:3].T).index_add(0, jj_t, t.j[3:].T)
    return FL(val, g.reshape(3 * N), lap)


def jastrow_fwd_lap(p, cfg, r32, h: Node) -> FL:
    """Jastrow.apply_with_fwd_lap with a NodeWithFwdLap embedding (jastrow.py:387-459)."""
    jp = p["jastrow"]["params"]
    dtype = h.x.dtype
    N, n_up = cfg.n_el, cfg.n_up
    r = torch.as_tensor(np.asarray(r32, np.float64)).to(dtype)
    z = torch.zeros((), dtype=dtype)
    logpsi = FL(z, torch.zeros(3 * N, dtype=dtype), z)
    if cfg.e_e_cusps == "psiformer":
        logpsi = logpsi + _cusp_fwd_lap(jp, cfg, r)
    if cfg.use_mlp_jastrow or cfg.use_log_jastrow or cfg.use_attention:
        J = FL(torch.zeros(2, dtype=dtype), torch.zeros((3 * N, 2), dtype=dtype), torch.zeros(2, dtype=dtype))
        if cfg.use_attention:
            a = jp["att"]
            H, Dk = cfg.attention_heads, cfg.attention_dim
            q = h.linear(a["Q_W"]["kernel"], a["Q_W"]["bias"]).reshape_feat(H, Dk)
            v = h.linear(a["V_W"]["kernel"], a["V_W"]["bias"]).reshape_feat(H, Dk)
            qa = Node(q.x * a["register_keys"], q.jac * a["register_keys"], q.lap * a["register_keys"],
                      q.idx_ctr, q.idx_dep).sum_feat_last()
            sc = 1.0 / math.sqrt(Dk)
            att = qa.elementwise(lambda x: torch.exp(-x * sc), lambda x: -sc * torch.exp(-x * sc),
                                 lambda x: sc * sc * torch.exp(-x * sc))
            att_e = Node(att.x.unsqueeze(-1), att.jac.unsqueeze(-1), att.lap.unsqueeze(-1), att.idx_ctr, att.idx_dep)
            vals = att_e.mul(Node(v.x, v.jac, v.lap, v.idx_ctr, v.idx_dep))
            nu, nd = att.sum_from_to(0, n_up, N), att.sum_from_to(n_up, N, N)
            vu, vd = vals.sum_from_to(0, n_up, N), vals.sum_from_to(n_up, N, N)
            norm = FL(torch.stack([nu.x, nd.x]), torch.stack([nu.j, nd.j], 1), torch.stack([nu.l, nd.l]))
            values = FL(torch.stack([vu.x, vd.x]), torch.stack([vu.j, vd.j], 1), torch.stack([vu.l, vd.l]))
            J = J + att_readout(jp, cfg, norm, values)
        if cfg.use_mlp_jastrow or cfg.use_log_jastrow:
            x = h
            for W, b in jastrow_mlp_layers(jp, cfg):
                x = x.linear(W, b).silu()
            zb = torch.stack([torch.zeros((), dtype=dtype),
                              jp["log_bias"] if cfg.use_log_jastrow else torch.zeros((), dtype=dtype)])
            x = x.scale(jp["mlp_scale"]).add_const(zb)
            J = J + x.sum_over_nodes(N)
        _, lj = mlp_to_logpsi(cfg, J)
        if isinstance(lj, FL):
            logpsi = logpsi + lj
    return logpsi


def log_psi_fwd_lap(p, cfg, r32):
    """_logpsi_with_fwd_lap_sparse (wave_function.py:165-181): returns (sign, FL logpsi)
    with a dense [3N] gradient and the Laplacian."""
    h, nl = embedding_fwd_lap(p, cfg, r32)
    phi, jac, lap = orbitals_fwd_lap(p, cfg, r32, h)
    N, D = cfg.n_el, cfg.n_determinants
    signs, lds = [], []
    for d in range(D):
        s, ld = sparse_slogdet_fwd_lap(phi[d], jac[d], lap[d], h.idx_ctr, h.idx_dep, N)
        signs.append(s)
        lds.append(ld)
    sg = torch.stack(signs)
    L = FL(torch.stack([a.x for a in lds]), torch.stack([a.j for a in lds], 1), torch.stack([a.l for a in lds]))
    m = L.x.max().detach()
    w = (fl.exp(L - m) * sg).sum(0)
    logpsi = fl.log_abs(w) + m
    sign = torch.sign(w.x)
    lj = jastrow_fwd_lap(p, cfg, r32, h)
    return sign, logpsi + lj


def kinetic_energy(p, cfg, r32):
    """MoonLikeWaveFunction.kinetic_energy (wave_function.py:215-221):
    -1/2 (sum Laplacian + |grad|^2)."""
    _, lp = log_psi_fwd_lap(p, cfg, r32)
    return -0.5 * (lp.l.sum() + (lp.j * lp.j).sum())


def local_energy(p, cfg, r32):
    """make_local_energy (hamiltonian.py:69-81) for an all-electron molecule (ECP terms are
    zero).  Returns (E_loc, E_kin, E_pot)."""
    ek = float(kinetic_energy(p, cfg, r32))
    ep = potential_energy(r32, cfg.mol.R, cfg.mol.Z)
    return ek + ep, ek, ep


def kinetic_energy_dense(p, cfg, r32):
    """Dense Hessian-trace kinetic energy of the value path (the role of
    kinetic_energy_dense, wave_function.py:223-224 / hamiltonian.py:29-46), float64,
    with the neighbour graph frozen at ``r32``."""
    r = torch.as_tensor(np.asarray(r32, np.float64))

    def f(x):
        return log_psi_value_torch(p, cfg, x.reshape(-1, 3), r32)

    x = r.reshape(-1)
    g = torch.func.grad(f)(x)
    Hm = torch.func.jacfwd(torch.func.grad(f))(x)
    return -0.5 * (torch.trace(Hm) + (g * g).sum()), g, torch.trace(Hm)


# -----------------------------------------------------------------------------------------
# a11-a12: value path with state, low-rank update
# -----------------------------------------------------------------------------------------
def log_psi_with_state(p, cfg, r32):
    """MoonLikeWaveFunction.log_psi_with_state (wave_function.py:226-249)."""
    h, est = embedding_apply(p, cfg, r32, return_state=True)
    phi, ost = orbitals_apply(p, cfg, r32, h, return_state=True)
    (sl, ld), inv = slog_and_inverse(phi)
    sgn, lp = logsumexp_signed(ld, sl)
    r = torch.as_tensor(np.asarray(r32, np.float64)).to(h.dtype)
    (sJ, lJ), jst = jastrow_apply(p, cfg, r, h, return_state=True)
    state = dict(embedding=est, orbitals=ost,
                 determinant=dict(matrices=phi, inverses=inv, signs=sl, logdets=ld), jastrow=jst,
                 r32=np.asarray(r32, np.float32).copy())
    return (sgn * sJ, lp + lJ), state


def affected_electrons(r_old32, r_new32, idx, cutoff):
    """get_low_rank_pair_indices (new_sparse_model.py:439-453): electrons whose embedding
    changes = neighbours of the moved electron before or after the move, plus itself."""
    inc_old = pair_distances_f32(r_old32) < np.float32(cutoff)
    inc_new = pair_distances_f32(r_new32) < np.float32(cutoff)
    aff = inc_old[:, idx] | inc_new[:, idx]
    aff[idx] = True
    return np.nonzero(aff)[0]


def woodbury_update(sign, logdet, A, Ainv, newA, changed):
    """woodburry_update (utils.py:216-244): det(I + V A^-1 U) and the Woodbury inverse."""
    k = len(changed)
    V = newA[:, changed] - A[:, changed]  # [D,k,N]
    AU = Ainv[..., changed]  # [D,N,k]
    S = torch.eye(k, dtype=A.dtype) + V @ AU
    sd, ld = torch.linalg.slogdet(S)
    Sinv = torch.linalg.inv(S)
    new_inv = Ainv - AU @ Sinv @ V @ Ainv
    return sign * sd, logdet + ld, new_inv


def sherman_morrison_update(sign, logdet, A, Ainv, newA, changed):
    """sherman_morrison_update (utils.py:247-273): k sequential rank-1 updates."""
    for i in changed:
        v = newA[:, i, :] - A[:, i, :]  # [D,N]
        u = Ainv[:, :, i]  # [D,N]
        vA = torch.einsum("da,dab->db", v, Ainv)
        den = 1 + (v * u).sum(-1)
        Ainv = Ainv - u[:, :, None] * vA[:, None, :] / den[:, None, None]
        sign = sign * torch.sign(den)
        logdet = logdet + torch.log(torch.abs(den))
        A = A.clone()
        A[:, i, :] = newA[:, i, :]
    return sign, logdet, Ainv


def log_psi_low_rank_update(p, cfg, r_new32, idx, state, method="woodbury"):
    """MoonLikeWaveFunction.log_psi_low_rank_update (wave_function.py:251-287) with
    NewSparseEmbedding.low_rank_update (new_sparse_model.py:553-629),
    Orbitals.low_rank_update (orbitals.py:121-146) and Jastrow.low_rank_update
    (jastrow.py:309-361)."""
    est = state["embedding"]
    dtype = est["h0"].dtype
    N, n_up, F = cfg.n_el, cfg.n_up, cfg.feature_dim
    r_old32 = state["r32"]
    aff = affected_electrons(r_old32, r_new32, idx, cfg.cutoff)
    r = torch.as_tensor(np.asarray(r_new32, np.float64)).to(dtype)
    R = torch.as_tensor(np.asarray(cfg.mol.R, np.float64)).to(dtype)
    # h0 and messages of the moved electron
    mask = _en_mask(r_new32, cfg.mol.R, cfg.cutoff_1el)
    outs = elec_init(p, cfg, r[idx:idx + 1], R, mask[idx:idx + 1])
    names = ["h0", "msg_ct_same", "msg_ct_diff", "msg_nb_same", "msg_nb_diff"]
    new = {k: est[k].clone() for k in names}
    for k, o in zip(names, outs):
        new[k][idx] = o[0]
    # recontract all pairs of affected centres with the new neighbour lists
    nl = neighbour_lists(r_new32, n_up, cfg.cutoff)
    h = est["h_out"].clone()
    aff_t = torch.as_tensor(aff)
    h[aff_t] = new["h0"][aff_t]
    is_aff = np.zeros(N, bool)
    is_aff[aff] = True
    cut = learned_cutoff(p, cfg)
    ep = p["embedding"]
    for same, edge_p, mct, mnb in ((True, ep["edge_same"], new["msg_ct_same"], new["msg_nb_same"]),
                                   (False, ep["edge_diff"], new["msg_ct_diff"], new["msg_nb_diff"])):
        sel = (nl.same == same) & is_aff[nl.ctr]
        if not sel.any():
            continue
        ci, ni = torch.as_tensor(nl.ctr[sel]), torch.as_tensor(nl.nbr[sel])
        gamma, e = ee_edge(edge_p, cut, r[ci] - r[ni])
        h = h.index_add(0, ci, gamma * fl.silu(e + mct[ci] + mnb[ni]))
    s0, s1 = ep["scales"]
    up = ep["updates"][0]["params"]
    x = h[aff_t] * s0
    x = fl.silu(x @ up["Linear_0"]["kernel"] + up["Linear_0"]["bias"])
    x = fl.silu(x @ up["Linear_1"]["kernel"] + up["Linear_1"]["bias"])
    h[aff_t] = x * s1
    # orbitals: envelope row of the moved electron, orbital rows of affected electrons
    ost = state["orbitals"]
    op = p["to_orbitals"]["params"]
    env = ost["envelopes"].clone()
    orb = ost["orbitals"].clone()
    phi_full = orbitals_apply(p, cfg, r_new32, h, return_state=True)[1]
    env[idx] = phi_full["envelopes"][idx]
    orb[aff_t] = phi_full["orbitals"][aff_t]
    D = cfg.n_determinants
    phi = (orb * env).reshape(N, D, N).permute(1, 0, 2).contiguous()
    dst = state["determinant"]
    if method == "woodbury":
        sl, ld, inv = woodbury_update(dst["signs"], dst["logdets"], dst["matrices"], dst["inverses"], phi, aff_t)
    else:
        sl, ld, inv = sherman_morrison_update(dst["signs"], dst["logdets"], dst["matrices"], dst["inverses"], phi, aff)
    sgn, lp = logsumexp_signed(ld, sl)
    # Jastrow: cusp recomputed; attention/MLP rows of affected electrons
    (sJ, lJ), jst = jastrow_apply(p, cfg, r, h, return_state=True)
    new_state = dict(
        embedding=dict(electrons=r, h_out=h, **new), orbitals=dict(envelopes=env, orbitals=orb),
        determinant=dict(matrices=phi, inverses=inv, signs=sl, logdets=ld), jastrow=jst,
        r32=np.asarray(r_new32, np.float32).copy())
    return (sgn * sJ, lp + lJ), new_state, aff


# -----------------------------------------------------------------------------------------
# parameter gradient of log|psi| (the per-walker jacobian of preconditioner.py:191-194)
# -----------------------------------------------------------------------------------------
def grad_logpsi(params, cfg, r32):
    """d log|psi| / d theta for every leaf of the parameter tree (numpy FP64 params): the
    jax.grad(log_p) of make_dense_spring_preconditioner (preconditioner.py:191-194) for one
    walker, by torch autograd of :func:`signed` in FP64.  Returns {leaf path: array}, paths as
    '/embedding/elec_init/params/...' in the tree's order; leaves that do not reach log|psi|
    get zeros."""
    leaves, paths = [], []

    def mark(t, path):
        if isinstance(t, dict):
            return {k: mark(v, f"{path}/{k}") for k, v in t.items()}
        if isinstance(t, (list, tuple)):
            return [mark(v, f"{path}/{i}") for i, v in enumerate(t)]
        x = torch.as_tensor(np.asarray(t, np.float64)).clone().requires_grad_(True)
        leaves.append(x)
        paths.append(path)
        return x

    pt = mark(params, "")
    _, lp = signed(pt, cfg, r32)
    gs = torch.autograd.grad(lp, leaves, allow_unused=True)
    return {p: (np.zeros(tuple(x.shape)) if g is None else g.detach().numpy()) for p, x, g in zip(paths, leaves, gs)}
