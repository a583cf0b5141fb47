"""Parameter tree of the FiRE hot path: random initialisation, packing for the C ABI.

The tree mirrors the reference's ``MoonLikeParams(embedding, to_orbitals, jastrow)``
(src/sparse_wf/model/wave_function.py:54-57) with ``EmbeddingParams``
(new_sparse_model.py:218-225) and flax-style sub-trees for the modules.  Module names
follow flax's naming rules applied to the reference source (``beta_ee``
new_sparse_model.py:176, ``to_orbitals_up``/``env_up`` orbitals.py:53-81, ``Linear_i`` of
``SparseMLP`` sparse_fwd_lap.py:222, ...).

Initialisers restate the reference families with NumPy's PCG64:

* ``lecun_normal`` (model/utils.py:339-342): truncated normal on [-1, 1] times 1/sqrt(shape[0]);
* flax ``Dense``/``Linear`` kernels: flax ``lecun_normal`` (variance-scaling, truncated at 2 sigma);
* ``scale_initializer`` (utils.py:378-383), ``envelope_sigma_initializer`` (envelopes.py:28-32),
  isotropic-envelope ``pi`` ~ N(0, 1/n_envelopes) (envelopes.py:69-74);
* nucleus-dependent e-n filter params (new_sparse_model.py:382-397).

The normalisation scales ``embedding.scales = (s0, s1)`` are set from a forward pass as
``1/std`` (new_sparse_model.py:378, utils.py:461-468); see :func:`set_scales`.
"""
from __future__ import annotations

from typing import Callable

import numpy as np

from .config import FireConfig

_FLAX_TRUNC = 0.87962566103423978  # std of a unit normal truncated to [-2, 2]


def _truncnorm(rng, lo, hi, shape):
    out = rng.standard_normal(size=shape)
    bad = (out < lo) | (out > hi)
    while bad.any():
        out[bad] = rng.standard_normal(size=int(bad.sum()))
        bad = (out < lo) | (out > hi)
    return out


def lecun_normal_ref(rng, shape):
    """model/utils.py:339-342."""
    return (_truncnorm(rng, -1.0, 1.0, shape) / np.sqrt(shape[0])).astype(np.float32)


def flax_lecun_normal(rng, shape):
    fan_in = int(np.prod(shape[:-1]))
    std = np.sqrt(1.0 / fan_in) / _FLAX_TRUNC
    return (_truncnorm(rng, -2.0, 2.0, shape) * std).astype(np.float32)


def scale_initializer(rng, cutoff, shape):
    """model/utils.py:378-383."""
    n = shape[-1]
    max_len = min(20.0, 0.75 * cutoff)
    scale = np.linspace(0.5 * max_len, max_len, n)
    scale = scale * (1 + 0.1 * rng.standard_normal(size=shape))
    return scale.astype(np.float32)


def envelope_sigma_initializer(rng, shape, smin=1.0, smax=5.0):
    """envelopes.py:28-32 (the stored value is the softplus pre-image)."""
    scale = np.geomspace(smin, smax, shape[-1])
    # jax.random.truncated_normal(key, 0.8, 1.2): a unit normal truncated to [0.8, 1.2]
    scale = scale * _truncnorm(rng, 0.8, 1.2, shape)
    return np.log(np.expm1(scale)).astype(np.float32)


def _dense(rng, fan_in, fan_out, bias=True, bias_init=None):
    p = {"kernel": flax_lecun_normal(rng, (fan_in, fan_out))}
    if bias:
        p["bias"] = np.zeros(fan_out, np.float32) if bias_init is None else bias_init(rng, (fan_out,))
    return p


def _pairwise_filter_dense(rng, n_env, w0, w1):
    return {"Dense_0": _dense(rng, w0, w1, True), "Dense_1": _dense(rng, n_env, w1, False)}


def init_params(cfg: FireConfig, seed: int = 0) -> dict:
    """Random parameters in the reference layout; ``embedding.scales`` = (1, 1) until
    :func:`set_scales` is applied."""
    rng = np.random.default_rng(seed)
    F = cfg.feature_dim
    w0, w1 = cfg.pair_mlp_widths
    M = cfg.n_nuc
    E_en = cfg.pair_n_envelopes
    N, D, E = cfg.n_el, cfg.n_determinants, cfg.n_envelopes

    dyn = {
        "filter": {
            "scales": scale_initializer(rng, cfg.cutoff, (M, E_en)),
            "kernel": lecun_normal_ref(rng, (M, 4, w0)),
            "bias": (rng.standard_normal(size=(M, w0)) * 2.0).astype(np.float32),
        },
        "nuc_embedding": rng.standard_normal(size=(M, F)).astype(np.float32),
    }
    elec_init = {"params": {
        "ElecNucEdge_0": {
            "PairwiseFilter_0": _pairwise_filter_dense(rng, E_en, w0, w1),
            "Dense_0": _dense(rng, w1, F, False),
            "Dense_1": _dense(rng, 4, F, False),
        },
        "LayerNorm_0": {"scale": np.ones(F, np.float32), "bias": np.zeros(F, np.float32)},
        "GatedLinearUnit_0": {"Dense_0": _dense(rng, F, 2 * F, False)},
        "Dense_0": _dense(rng, F, F, True),
        "Dense_1": _dense(rng, F, 5 * F, True),
    }}

    def edge():
        return {"params": {
            "beta_ee": {
                "scale": scale_initializer(rng, cfg.cutoff, (w0,)),
                "kernel": lecun_normal_ref(rng, (4, w0)),
                "bias": rng.standard_normal(size=(w0,)).astype(np.float32),
                **_pairwise_filter_dense(rng, w0, w0, w1),
            },
            "Dense_0": _dense(rng, w1, F, False),
            "Dense_1": _dense(rng, 4, F, True),
        }}

    embedding = {
        "dynamic_params_en": dyn,
        "elec_init": elec_init,
        "edge_same": edge(),
        "edge_diff": edge(),
        "updates": [{"params": {"Linear_0": _dense(rng, F, F, True), "Linear_1": _dense(rng, F, F, True)}}],
        "scales": [np.float32(1.0), np.float32(1.0)],
        "cutoff_param": np.float32(10.0),
    }

    def trunc_bias(r, shape):
        return (_truncnorm(r, -2.0, 2.0, shape) * 0.01).astype(np.float32)

    orb = {}
    for s in ("up", "dn"):
        if s == "dn" and cfg.spin_restricted:
            continue
        orb[f"to_orbitals_{s}"] = _dense(rng, F, N * D, True, bias_init=trunc_bias)
        orb[f"env_{s}"] = {
            "sigma": envelope_sigma_initializer(rng, (M, D, E)),
            "pi": (rng.standard_normal(size=(M, D, E, N)) / np.sqrt(E)).astype(np.float32),
        }
    to_orbitals = {"params": orb}

    jp = {}
    if cfg.e_e_cusps == "psiformer":
        jp["pairwise_cusps"] = {
            "alpha_same": np.float32(1.0), "alpha_diff": np.float32(1.0),
            "factor_same": np.float32(0.0), "factor_diff": np.float32(0.0),
        }
    W = cfg.mlp_width
    if cfg.use_mlp_jastrow or cfg.use_log_jastrow:
        widths = [W] * cfg.mlp_depth + [2]
        mlp, fan = {}, F
        for i, w in enumerate(widths):
            mlp[f"Linear_{i}"] = _dense(rng, fan, w, True)
            fan = w
        jp["mlp"] = mlp
        jp["mlp_scale"] = np.zeros(2, np.float32)
        if cfg.use_log_jastrow:
            jp["log_bias"] = np.float32(1.0)
    if cfg.use_attention:
        H, Dk = cfg.attention_heads, cfg.attention_dim
        out, fan = {}, 2 * H * Dk
        for i, w in enumerate([W] * cfg.mlp_depth + [2]):
            out[f"Dense_{i}"] = _dense(rng, fan, w, True)
            fan = w
        jp["att"] = {
            "register_keys": rng.standard_normal(size=(H, Dk)).astype(np.float32),
            "Q_W": _dense(rng, F, H * Dk, True),
            "V_W": _dense(rng, F, H * Dk, True),
            "out": out,
            "scale": np.zeros(2, np.float32),
            "bias": np.ones(1, np.float32),
        }
    jastrow = {"params": jp}
    return {"embedding": embedding, "to_orbitals": to_orbitals, "jastrow": jastrow}


def exercise_jastrow(params: dict) -> dict:
    """Make the Jastrow non-trivial, as the reference test fixture does by setting every
    ``jastrow ... scale`` leaf to one (tests/utils.py:73-78); the e-e cusp factors are set
    to the cusp-condition values noted in jastrow.py:59."""
    jp = params["jastrow"]["params"]
    if "mlp_scale" in jp:
        jp["mlp_scale"] = np.ones_like(jp["mlp_scale"])
    if "att" in jp:
        jp["att"]["scale"] = np.ones_like(jp["att"]["scale"])
    if "pairwise_cusps" in jp:
        jp["pairwise_cusps"]["factor_same"] = np.float32(-0.25)
        jp["pairwise_cusps"]["factor_diff"] = np.float32(-0.5)
    return params


def set_scales(params: dict, embed_stages: Callable) -> dict:
    """``embed_stages(params) -> (h_contracted, h_mlp_unscaled_input_fn)``.

    Restates ``normalize(h, None)`` (utils.py:461-468) as used at init
    (new_sparse_model.py:378): s0 = 1/std(h after message passing); s1 = 1/std(MLP(s0 h)).
    ``embed_stages`` must return the two arrays in that order for the given params."""
    params["embedding"]["scales"] = [np.float32(1.0), np.float32(1.0)]
    h_contr, _ = embed_stages(params)
    s0 = 1.0 / np.std(np.asarray(h_contr, np.float64))
    s0 = np.float32(s0 if np.isfinite(s0) else 1.0)
    params["embedding"]["scales"] = [s0, np.float32(1.0)]
    _, h_mlp = embed_stages(params)
    s1 = 1.0 / np.std(np.asarray(h_mlp, np.float64))
    s1 = np.float32(s1 if np.isfinite(s1) else 1.0)
    params["embedding"]["scales"] = [s0, s1]
    return params


def tree_map(fn, tree):
    if isinstance(tree, dict):
        return {k: tree_map(fn, v) for k, v in tree.items()}
    if isinstance(tree, (list, tuple)):
        return type(tree)(tree_map(fn, v) for v in tree)
    return fn(tree)


def leaves_with_path(tree, prefix=""):
    if isinstance(tree, dict):
        for k in tree:
            yield from leaves_with_path(tree[k], f"{prefix}/{k}")
    elif isinstance(tree, (list, tuple)):
        for i, v in enumerate(tree):
            yield from leaves_with_path(v, f"{prefix}/{i}")
    else:
        yield prefix, tree


def _softplus(x):
    x = np.asarray(x, np.float64)
    return np.where(x > 20, x, np.log1p(np.exp(np.minimum(x, 20))))


def _sigmoid(x):
    return 1.0 / (1.0 + np.exp(-np.asarray(x, np.float64)))


def _swap_perm(N: int, D: int, n_up: int) -> np.ndarray:
    """Column permutation of _swap_spin_updn (orbitals.py:21-26): new[d,o] = old[d,(o+n_up)%N]."""
    o = np.arange(N)
    return np.concatenate([d * N + (o + n_up) % N for d in range(D)])


def padded_w1(w1: int) -> int:
    """Filter width the kernels are compiled for (4, 8 or 16); narrower filters are zero-padded,
    which adds exact zeros to every contraction."""
    for w in (4, 8, 16):
        if w1 <= w:
            return w
    raise ValueError(f"pair_mlp_widths[1] = {w1} > 16 is not supported")


def _pad_cols(a, n):
    a = np.asarray(a, np.float64)
    out = np.zeros(a.shape[:-1] + (n,))
    out[..., : a.shape[-1]] = a
    return out


def _pad_rows(a, n):
    a = np.asarray(a, np.float64)
    out = np.zeros((n,) + a.shape[1:])
    out[: a.shape[0]] = a
    return out


def pack_params(cfg: FireConfig, params: dict, dtype=np.float32) -> list:
    """Flatten ``params`` into the canonical float32 tensor list of ``fire_upload_params``.

    Transformations applied here (all exact re-arrangements or precomputations):
    softplus of the filter / envelope scales; the normalisation scale s0 folded into the first
    update-MLP weight and s1 into every first layer that consumes the embedding (orbitals,
    MLP Jastrow, attention Q/V) -- ``h_out = s1 * MLP(s0 * h)`` (new_sparse_model.py:691-693);
    Q and V weights concatenated; for ``spin_restricted`` the down-spin orbital / envelope
    parameters are the up-spin ones with the columns of _swap_spin_updn permuted
    (orbitals.py:88-107); learned cutoff = cutoff * sigmoid(cutoff_param) (:644)."""
    f32 = lambda a: np.ascontiguousarray(np.asarray(a, np.float64).astype(dtype))
    W1p = padded_w1(cfg.pair_mlp_widths[1])
    ep = params["embedding"]
    s0, s1 = (float(v) for v in ep["scales"])
    dyn = ep["dynamic_params_en"]
    ei = ep["elec_init"]["params"]
    edge = ei["ElecNucEdge_0"]
    out = [
        dyn["filter"]["kernel"], dyn["filter"]["bias"], _softplus(dyn["filter"]["scales"]), dyn["nuc_embedding"],
        _pad_cols(edge["PairwiseFilter_0"]["Dense_0"]["kernel"], W1p),
        _pad_cols(edge["PairwiseFilter_0"]["Dense_0"]["bias"], W1p),
        _pad_cols(edge["PairwiseFilter_0"]["Dense_1"]["kernel"], W1p),
        _pad_rows(edge["Dense_0"]["kernel"], W1p), edge["Dense_1"]["kernel"],
        ei["LayerNorm_0"]["scale"], ei["LayerNorm_0"]["bias"],
        ei["GatedLinearUnit_0"]["Dense_0"]["kernel"],
        ei["Dense_0"]["kernel"], ei["Dense_0"]["bias"], ei["Dense_1"]["kernel"], ei["Dense_1"]["bias"],
    ]
    for name in ("edge_same", "edge_diff"):
        q = ep[name]["params"]
        b = q["beta_ee"]
        out += [b["kernel"], b["bias"], _softplus(b["scale"]), _pad_cols(b["Dense_0"]["kernel"], W1p),
                _pad_cols(b["Dense_0"]["bias"], W1p), _pad_cols(b["Dense_1"]["kernel"], W1p),
                _pad_rows(q["Dense_0"]["kernel"], W1p), q["Dense_1"]["kernel"], q["Dense_1"]["bias"]]
    up = ep["updates"][0]["params"]
    out += [np.asarray(up["Linear_0"]["kernel"], np.float64) * s0, up["Linear_0"]["bias"],
            up["Linear_1"]["kernel"], up["Linear_1"]["bias"]]
    op = params["to_orbitals"]["params"]
    N, D = cfg.n_el, cfg.n_determinants
    W_up = np.asarray(op["to_orbitals_up"]["kernel"], np.float64)
    b_up = np.asarray(op["to_orbitals_up"]["bias"], np.float64)
    sp_up = _softplus(op["env_up"]["sigma"])
    pi_up = np.asarray(op["env_up"]["pi"], np.float64)
    if cfg.spin_restricted:
        perm = _swap_perm(N, D, cfg.n_up)
        W_dn, b_dn = W_up[:, perm], b_up[perm]
        sp_dn = sp_up
        pi_dn = pi_up[..., (np.arange(N) + cfg.n_up) % N]
    else:
        W_dn = np.asarray(op["to_orbitals_dn"]["kernel"], np.float64)
        b_dn = np.asarray(op["to_orbitals_dn"]["bias"], np.float64)
        sp_dn = _softplus(op["env_dn"]["sigma"])
        pi_dn = np.asarray(op["env_dn"]["pi"], np.float64)
    # pi as [det][nucleus * envelope][orbital] (one GEMM operand per determinant)
    pi_up = np.transpose(pi_up, (1, 0, 2, 3))
    pi_dn = np.transpose(pi_dn, (1, 0, 2, 3))
    out += [W_up * s1, b_up, W_dn * s1, b_dn, sp_up, pi_up, sp_dn, pi_dn]
    jp = params["jastrow"]["params"]
    if cfg.use_mlp_jastrow or cfg.use_log_jastrow:
        mlp = jp["mlp"]
        for l in range(cfg.mlp_depth):
            W = np.asarray(mlp[f"Linear_{l}"]["kernel"], np.float64)
            out += [W * s1 if l == 0 else W, mlp[f"Linear_{l}"]["bias"]]
        last = mlp[f"Linear_{cfg.mlp_depth}"]
        zb = [0.0, float(jp["log_bias"]) if cfg.use_log_jastrow else 0.0]
        out += [last["kernel"], last["bias"], jp["mlp_scale"], np.asarray(zb)]
    if cfg.use_attention:
        a = jp["att"]
        qv = np.concatenate([a["Q_W"]["kernel"], a["V_W"]["kernel"]], axis=1) * s1
        out += [qv, np.concatenate([a["Q_W"]["bias"], a["V_W"]["bias"]]), a["register_keys"]]
        for l in range(cfg.mlp_depth):
            out += [a["out"][f"Dense_{l}"]["kernel"], a["out"][f"Dense_{l}"]["bias"]]
        last = a["out"][f"Dense_{cfg.mlp_depth}"]
        out += [last["kernel"], last["bias"], a["scale"], np.asarray([0.0, float(np.asarray(a["bias"]).reshape(-1)[0])])]
    if cfg.e_e_cusps == "psiformer":
        c = jp["pairwise_cusps"]
        out.append(np.asarray([c["alpha_same"], c["alpha_diff"], c["factor_same"], c["factor_diff"]], np.float64))
    out.append(np.asarray([cfg.cutoff * _sigmoid(ep["cutoff_param"]), s1]))
    return [f32(t).reshape(-1) for t in out]


def pack_params_vjp(cfg: FireConfig, params: dict, g):
    """Adjoint of :func:`pack_params` for a batch of cotangents: ``g`` (torch [B, n], the packed
    tensors back to back, as ``fire_grad_logpsi`` writes them) -> a tree shaped like ``params``
    whose leaves are torch [B, *leaf shape] on ``g``'s device, the gradients with respect to the
    tree's own leaves.  Chain rules of the packing: softplus'd scales take sigmoid(x); folded
    normalisation scales give s0 * g, s1 * g to the weights and sum(W * g) to s0 / s1;
    zero-padded filter columns / rows are dropped; pi is transposed back; for
    ``spin_restricted`` the down-spin copies are scattered back through the column
    permutation; the learned cutoff takes cutoff * sigmoid'(cutoff_param); the Jastrow
    log-bias and attention bias are the second entries of their [0, bias] pairs."""
    import torch

    B = g.shape[0]
    sizes = [t.size for t in pack_params(cfg, params)]
    if g.shape[1] != sum(sizes):
        raise ValueError(f"gradient rows hold {g.shape[1]} floats, the packed parameters {sum(sizes)}")
    parts = iter(torch.split(g, sizes, dim=1))
    dev, dt = g.device, g.dtype
    out = tree_map(lambda x: torch.zeros((B,) + np.shape(x), dtype=dt, device=dev), params)
    T = lambda x: torch.as_tensor(np.asarray(x, np.float64), dtype=dt, device=dev)
    wsum = lambda W, gw: (T(W).reshape(1, -1) * gw.reshape(B, -1)).sum(1)
    W1p, w1 = padded_w1(cfg.pair_mlp_widths[1]), cfg.pair_mlp_widths[1]

    def take(shape=None):
        p = next(parts)
        return p if shape is None else p.reshape((B,) + tuple(shape))

    def add(node, key, v):
        node[key] += v.reshape(node[key].shape)

    def cols(x, leaf):  # _pad_cols
        return take(np.shape(leaf)[:-1] + (W1p,))[..., :w1]

    def rows(x, leaf):  # _pad_rows
        return take((W1p,) + np.shape(leaf)[1:])[:, :w1]

    ep, gp = params["embedding"], out["embedding"]
    s0, s1 = (float(v) for v in ep["scales"])
    gs0 = torch.zeros(B, dtype=dt, device=dev)
    gs1 = torch.zeros(B, dtype=dt, device=dev)
    dyn, gd = ep["dynamic_params_en"], gp["dynamic_params_en"]
    add(gd["filter"], "kernel", take())
    add(gd["filter"], "bias", take())
    add(gd["filter"], "scales", take() * T(_sigmoid(dyn["filter"]["scales"])).reshape(1, -1))
    add(gd, "nuc_embedding", take())
    ei, ge = ep["elec_init"]["params"], gp["elec_init"]["params"]
    e, ge_ = ei["ElecNucEdge_0"], ge["ElecNucEdge_0"]
    pf, gpf = e["PairwiseFilter_0"], ge_["PairwiseFilter_0"]
    add(gpf["Dense_0"], "kernel", cols(None, pf["Dense_0"]["kernel"]))
    add(gpf["Dense_0"], "bias", cols(None, pf["Dense_0"]["bias"]))
    add(gpf["Dense_1"], "kernel", cols(None, pf["Dense_1"]["kernel"]))
    add(ge_["Dense_0"], "kernel", rows(None, e["Dense_0"]["kernel"]))
    add(ge_["Dense_1"], "kernel", take())
    add(ge["LayerNorm_0"], "scale", take())
    add(ge["LayerNorm_0"], "bias", take())
    add(ge["GatedLinearUnit_0"]["Dense_0"], "kernel", take())
    for k in ("Dense_0", "Dense_1"):
        add(ge[k], "kernel", take())
        add(ge[k], "bias", take())
    for name in ("edge_same", "edge_diff"):
        q, gq = ep[name]["params"], gp[name]["params"]
        b, gb = q["beta_ee"], gq["beta_ee"]
        add(gb, "kernel", take())
        add(gb, "bias", take())
        add(gb, "scale", take() * T(_sigmoid(b["scale"])).reshape(1, -1))
        add(gb["Dense_0"], "kernel", cols(None, b["Dense_0"]["kernel"]))
        add(gb["Dense_0"], "bias", cols(None, b["Dense_0"]["bias"]))
        add(gb["Dense_1"], "kernel", cols(None, b["Dense_1"]["kernel"]))
        add(gq["Dense_0"], "kernel", rows(None, q["Dense_0"]["kernel"]))
        add(gq["Dense_1"], "kernel", take())
        add(gq["Dense_1"], "bias", take())
    up, gu = ep["updates"][0]["params"], gp["updates"][0]["params"]
    x = take()
    gs0 += wsum(up["Linear_0"]["kernel"], x)
    add(gu["Linear_0"], "kernel", s0 * x)
    add(gu["Linear_0"], "bias", take())
    add(gu["Linear_1"], "kernel", take())
    add(gu["Linear_1"], "bias", take())

    op, go = params["to_orbitals"]["params"], out["to_orbitals"]["params"]
    N, D, nu = cfg.n_el, cfg.n_determinants, cfg.n_up
    W_up = np.asarray(op["to_orbitals_up"]["kernel"], np.float64)
    gWu, gbu = take(W_up.shape), take()
    gWd, gbd = take(W_up.shape), take()
    gspu, gpiu, gspd, gpid = take(), take(), take(), take()
    pi_shape = np.shape(op["env_up"]["pi"])  # [nucleus, det, envelope, orbital]
    back = lambda p: p.reshape((B, pi_shape[1], pi_shape[0]) + tuple(pi_shape[2:])).transpose(1, 2)
    gs1 += wsum(W_up, gWu)
    add(go["to_orbitals_up"], "kernel", s1 * gWu)
    add(go["to_orbitals_up"], "bias", gbu)
    add(go["env_up"], "pi", back(gpiu))
    sig_up = T(_sigmoid(op["env_up"]["sigma"])).reshape(1, -1)
    add(go["env_up"], "sigma", gspu * sig_up)
    if cfg.spin_restricted:
        perm = torch.as_tensor(_swap_perm(N, D, nu), device=dev)
        gs1 += wsum(W_up[:, _swap_perm(N, D, nu)], gWd)
        t = torch.zeros_like(gWd)
        t[:, :, perm] = s1 * gWd
        add(go["to_orbitals_up"], "kernel", t)
        t = torch.zeros_like(gbd)
        t[:, perm] = gbd
        add(go["to_orbitals_up"], "bias", t)
        add(go["env_up"], "sigma", gspd * sig_up)
        t = torch.zeros((B,) + tuple(pi_shape), dtype=dt, device=dev)
        t[..., torch.as_tensor((np.arange(N) + nu) % N, device=dev)] = back(gpid)
        add(go["env_up"], "pi", t)
    else:
        gs1 += wsum(op["to_orbitals_dn"]["kernel"], gWd)
        add(go["to_orbitals_dn"], "kernel", s1 * gWd)
        add(go["to_orbitals_dn"], "bias", gbd)
        add(go["env_dn"], "sigma", gspd * T(_sigmoid(op["env_dn"]["sigma"])).reshape(1, -1))
        add(go["env_dn"], "pi", back(gpid))

    jp, gj = params["jastrow"]["params"], out["jastrow"]["params"]
    if cfg.use_mlp_jastrow or cfg.use_log_jastrow:
        mlp, gm = jp["mlp"], gj["mlp"]
        for l in range(cfg.mlp_depth):
            x = take()
            if l == 0:
                gs1 += wsum(mlp["Linear_0"]["kernel"], x)
                x = s1 * x
            add(gm[f"Linear_{l}"], "kernel", x)
            add(gm[f"Linear_{l}"], "bias", take())
        add(gm[f"Linear_{cfg.mlp_depth}"], "kernel", take())
        add(gm[f"Linear_{cfg.mlp_depth}"], "bias", take())
        add(gj, "mlp_scale", take())
        zb = take()
        if cfg.use_log_jastrow:
            add(gj, "log_bias", zb[:, 1])
    if cfg.use_attention:
        a, ga = jp["att"], gj["att"]
        hq = np.shape(a["Q_W"]["kernel"])[1]
        x = take((np.shape(a["Q_W"]["kernel"])[0], -1))
        qv = np.concatenate([a["Q_W"]["kernel"], a["V_W"]["kernel"]], axis=1)
        gs1 += wsum(qv, x)
        add(ga["Q_W"], "kernel", s1 * x[:, :, :hq])
        add(ga["V_W"], "kernel", s1 * x[:, :, hq:])
        x = take()
        add(ga["Q_W"], "bias", x[:, :hq])
        add(ga["V_W"], "bias", x[:, hq:])
        add(ga, "register_keys", take())
        for l in range(cfg.mlp_depth + 1):
            add(ga["out"][f"Dense_{l}"], "kernel", take())
            add(ga["out"][f"Dense_{l}"], "bias", take())
        add(ga, "scale", take())
        add(ga, "bias", take()[:, 1])
    if cfg.e_e_cusps == "psiformer":
        x, gc = take(), gj["pairwise_cusps"]
        for i, k in enumerate(("alpha_same", "alpha_diff", "factor_same", "factor_diff")):
            add(gc, k, x[:, i])
    x = take()
    sg = float(_sigmoid(ep["cutoff_param"]))
    add(gp, "cutoff_param", x[:, 0] * (cfg.cutoff * sg * (1.0 - sg)))
    gs1 += x[:, 1]
    gp["scales"][0] += gs0
    gp["scales"][1] += gs1
    return out
