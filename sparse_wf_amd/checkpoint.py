"""Reader (and writer) of the reference's training checkpoints.

The reference stores ``TrainingState`` with flax's msgpack serialisation
(``TrainingState.serialize`` src/sparse_wf/api.py:294-303 -> ``flax.serialization.to_bytes``;
written as ``chkpt{step:06d}.msgpk`` by ``loggers.py:188-197``).  flax is not needed to read
it: the format is plain msgpack of the state dict of the pytree, where

* a dataclass / ``struct.PyTreeNode`` (``TrainingState``) and a ``NamedTuple``
  (``MoonLikeParams``, ``EmbeddingParams``, ``NucleusDependentParams``, ``DynamicFilterParams``,
  ``WidthSchedulerState``) become maps keyed by field name (older flax wrote a NamedTuple as
  ``{"name": ..., "fields": {"0": ..., ...}}``; both are read);
* a tuple or list becomes a map keyed ``"0", "1", ...`` (``EmbeddingParams.updates`` and
  ``.scales``);
* an array is msgpack extension type 1 holding ``msgpack((shape, dtype name, C-order bytes))``,
  a NumPy scalar extension type 3 with the same payload, a Python complex extension type 2;
* an array above flax's chunk size is a map ``{"__msgpack_chunked_array__": True,
  "shape": {"0": ...}, "chunks": {"0": flat chunk, ...}}``.

The parameter tree ``MoonLikeParams(embedding, to_orbitals, jastrow)``
(model/wave_function.py:54-57) maps one to one onto the tree of :mod:`sparse_wf_amd.params`:
``embedding`` = ``EmbeddingParams(dynamic_params_en, elec_init, edge_same, edge_diff, updates,
scales, cutoff_param)`` (model/new_sparse_model.py:218-225), ``to_orbitals`` and ``jastrow`` the
flax variable dicts of ``Orbitals`` and ``Jastrow``.  Checkpoints written before
``cutoff_param`` existed (the case ``src/convert_checkpoint.py:91-101`` converts) get the
reference's initial value 10.0 (new_sparse_model.py:376).
"""
from __future__ import annotations

import os
from typing import Any

import numpy as np

from .config import FireConfig

EXT_NDARRAY, EXT_COMPLEX, EXT_NPSCALAR = 1, 2, 3
CHUNK_KEY = "__msgpack_chunked_array__"
TRAINING_STATE_FIELDS = ("key", "params", "electrons", "opt_state", "width_state", "spin_state", "step")
PARAM_FIELDS = ("embedding", "to_orbitals", "jastrow")
EMBEDDING_FIELDS = ("dynamic_params_en", "elec_init", "edge_same", "edge_diff", "updates", "scales", "cutoff_param")
CUTOFF_PARAM_INIT = 10.0


def _msgpack():
    try:
        import msgpack
    except ImportError as e:  # pragma: no cover - msgpack ships with the image
        raise ImportError("reading flax checkpoints needs the msgpack package") from e
    return msgpack


# ---------------------------------------------------------------------------------------------
# msgpack layer
# ---------------------------------------------------------------------------------------------

def _array_from_payload(data: bytes) -> np.ndarray:
    msgpack = _msgpack()
    shape, dtype_name, buf = msgpack.unpackb(data, raw=True)
    name = dtype_name.decode() if isinstance(dtype_name, bytes) else str(dtype_name)
    if name == "bfloat16":
        # stored as 16-bit words; widened exactly to float32
        u = np.frombuffer(buf, np.uint16).astype(np.uint32) << 16
        return u.view(np.float32).reshape(tuple(shape))
    return np.frombuffer(buf, dtype=np.dtype(name)).reshape(tuple(shape)).copy()


def _ext_hook(code: int, data: bytes):
    msgpack = _msgpack()
    if code == EXT_NDARRAY:
        return _array_from_payload(data)
    if code == EXT_NPSCALAR:
        return _array_from_payload(data)[()]
    if code == EXT_COMPLEX:
        re, im = msgpack.unpackb(data)
        return complex(re, im)
    raise ValueError(f"unknown msgpack extension type {code} in checkpoint")


def _index_map_to_list(d: dict) -> list:
    return [d[str(i)] for i in range(len(d))]


def _unchunk(tree):
    if isinstance(tree, dict):
        if tree.get(CHUNK_KEY) is True:
            shape = tuple(int(s) for s in _index_map_to_list(tree["shape"]))
            flat = np.concatenate([np.asarray(c).reshape(-1) for c in _index_map_to_list(tree["chunks"])])
            return flat.reshape(shape)
        return {k: _unchunk(v) for k, v in tree.items()}
    return tree


def msgpack_restore(data: bytes) -> Any:
    """State dict of a flax msgpack blob (maps with string keys, NumPy arrays, scalars)."""
    msgpack = _msgpack()
    tree = msgpack.unpackb(data, ext_hook=_ext_hook, raw=False, strict_map_key=False)
    return _unchunk(tree)


def _ext_default(x):
    msgpack = _msgpack()
    if isinstance(x, np.ndarray):
        payload = msgpack.packb((x.shape, x.dtype.name, np.ascontiguousarray(x).tobytes("C")), use_bin_type=True)
        return msgpack.ExtType(EXT_NDARRAY, payload)
    if isinstance(x, np.generic):
        a = np.asarray(x)
        payload = msgpack.packb((a.shape, a.dtype.name, a.tobytes("C")), use_bin_type=True)
        return msgpack.ExtType(EXT_NPSCALAR, payload)
    if isinstance(x, complex):
        return msgpack.ExtType(EXT_COMPLEX, msgpack.packb((x.real, x.imag)))
    raise TypeError(f"cannot serialise {type(x).__name__}")


def _chunk(tree, chunk_bytes: int):
    if isinstance(tree, dict):
        return {k: _chunk(v, chunk_bytes) for k, v in tree.items()}
    if isinstance(tree, np.ndarray) and tree.nbytes > chunk_bytes:
        per = max(1, chunk_bytes // tree.dtype.itemsize)
        flat = tree.reshape(-1)
        chunks = [flat[i:i + per] for i in range(0, flat.size, per)]
        return {CHUNK_KEY: True, "shape": {str(i): int(s) for i, s in enumerate(tree.shape)},
                "chunks": {str(i): c for i, c in enumerate(chunks)}}
    return tree


def msgpack_serialize(state_dict: Any, chunk_bytes: int = (1 << 30) - (1 << 16)) -> bytes:
    """flax msgpack blob of a state dict (maps, arrays, scalars); arrays larger than
    ``chunk_bytes`` are split into flat chunks the way flax does."""
    msgpack = _msgpack()
    return msgpack.packb(_chunk(state_dict, chunk_bytes), default=_ext_default, strict_types=True)


# ---------------------------------------------------------------------------------------------
# parameter tree
# ---------------------------------------------------------------------------------------------

def _fields(node, names, path):
    """Field map of a serialised NamedTuple (field names, or the older name/fields form)."""
    if not isinstance(node, dict):
        raise ValueError(f"checkpoint: {path} is {type(node).__name__}, expected a NamedTuple map")
    if set(node) == {"name", "fields"}:
        f = node["fields"]
        return {n: f[str(i)] for i, n in enumerate(names) if str(i) in f}
    return node


def _match(template, node, path):
    """``node`` reshaped onto the structure of ``template`` (dict / list / leaf); raises on a
    missing entry or a shape mismatch and names the path."""
    if isinstance(template, dict):
        if not isinstance(node, dict):
            raise ValueError(f"checkpoint: {path} is {type(node).__name__}, expected a map")
        missing = [k for k in template if k not in node]
        if missing:
            raise ValueError(f"checkpoint: {path} lacks {missing} (has {sorted(node)})")
        return {k: _match(template[k], node[k], f"{path}/{k}") for k in template}
    if isinstance(template, (list, tuple)):
        if isinstance(node, dict):
            node = [node[str(i)] for i in range(len(node)) if str(i) in node]
        if not isinstance(node, (list, tuple)) or len(node) != len(template):
            raise ValueError(f"checkpoint: {path} holds {len(node) if hasattr(node, '__len__') else node} entries, "
                             f"expected {len(template)}")
        return [_match(t, n, f"{path}/{i}") for i, (t, n) in enumerate(zip(template, node))]
    a = np.asarray(node)
    t = np.asarray(template)
    if a.shape != t.shape:
        raise ValueError(f"checkpoint: {path} has shape {a.shape}, the configuration expects {t.shape}")
    if not np.issubdtype(a.dtype, np.floating):
        raise ValueError(f"checkpoint: {path} has dtype {a.dtype}, expected floating point")
    out = a.astype(np.float32)
    return np.float32(out) if out.ndim == 0 and isinstance(template, np.generic) else out


def params_from_state_dict(node: dict, cfg: FireConfig) -> dict:
    """The parameter tree of :mod:`sparse_wf_amd.params` from the state dict of
    ``MoonLikeParams`` as stored in a checkpoint, checked entry by entry against the shapes
    ``cfg`` implies."""
    from .params import init_params

    template = init_params(cfg, seed=0)
    top = _fields(node, PARAM_FIELDS, "params")
    emb = dict(_fields(top["embedding"], EMBEDDING_FIELDS, "params/embedding"))
    if "cutoff_param" not in emb:  # checkpoints from before the learned cutoff
        emb["cutoff_param"] = np.float32(CUTOFF_PARAM_INIT)
    dyn = dict(_fields(emb["dynamic_params_en"], ("filter", "nuc_embedding"), "params/embedding/dynamic_params_en"))
    dyn["filter"] = _fields(dyn["filter"], ("scales", "kernel", "bias"), "params/embedding/dynamic_params_en/filter")
    emb["dynamic_params_en"] = dyn
    tree = {"embedding": emb, "to_orbitals": top["to_orbitals"], "jastrow": top["jastrow"]}
    return _match(template, tree, "params")


def params_to_state_dict(params: dict) -> dict:
    """State dict of ``MoonLikeParams`` (current flax layout) for a parameter tree."""
    def conv(x):
        if isinstance(x, dict):
            return {k: conv(v) for k, v in x.items()}
        if isinstance(x, (list, tuple)):
            return {str(i): conv(v) for i, v in enumerate(x)}
        a = np.asarray(x, np.float32)
        return a
    return conv(params)


def load_training_state(src, cfg: FireConfig) -> dict:
    """Read a reference checkpoint (path or bytes).

    Returns ``{"params", "electrons" [B, N, 3] float32, "step", "key", "width"}``; ``params``
    feeds :class:`~sparse_wf_amd.wave_function.FireWaveFunction` directly.  The optimiser
    and spin states are not on the hot path and are not returned."""
    if isinstance(src, (str, os.PathLike)):
        with open(src, "rb") as f:
            src = f.read()
    sd = msgpack_restore(bytes(src))
    if not isinstance(sd, dict) or "params" not in sd:
        raise ValueError("checkpoint: not a TrainingState (no 'params' entry)")
    out = {"params": params_from_state_dict(sd["params"], cfg)}
    el = sd.get("electrons")
    if el is not None:
        el = np.asarray(el, np.float32)
        el = el.reshape(-1, *el.shape[-2:])
        if el.shape[1:] != (cfg.n_el, 3):
            raise ValueError(f"checkpoint: electrons have shape {el.shape}, expected [B, {cfg.n_el}, 3]")
    out["electrons"] = el
    out["step"] = None if sd.get("step") is None else int(np.asarray(sd["step"]).reshape(-1)[0])
    out["key"] = None if sd.get("key") is None else np.asarray(sd["key"])
    ws = sd.get("width_state")
    if isinstance(ws, dict):
        ws = _fields(ws, ("width", "pmoves", "i"), "width_state")
        out["width"] = None if ws.get("width") is None else float(np.asarray(ws["width"]).reshape(-1)[0])
    else:
        out["width"] = None
    return out


def save_training_state(params: dict, electrons=None, step: int = 0, key=None, width: float | None = None) -> bytes:
    """A checkpoint blob in the reference layout holding the parameters, electrons, step, key
    and MCMC width (empty optimiser and spin states: enough for evaluation and sampling, not
    for resuming the reference's optimiser)."""
    sd = {
        "key": np.zeros(2, np.uint32) if key is None else np.asarray(key),
        "params": params_to_state_dict(params),
        "electrons": None if electrons is None else np.asarray(electrons, np.float32),
        "opt_state": {},
        "width_state": {"width": np.asarray(0.02 if width is None else width, np.float32),
                        "pmoves": np.zeros(0, np.float32), "i": np.asarray(0, np.int32)},
        "spin_state": {},
        "step": np.asarray(step, np.int32),
    }
    return msgpack_serialize(sd)
