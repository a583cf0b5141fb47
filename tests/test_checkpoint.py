"""Checkpoint reader against blobs written in the reference's flax-msgpack layout.

The blobs are assembled here with msgpack directly (not with the module's writer): a
``TrainingState`` map keyed by its field names (src/sparse_wf/api.py:285-292), the
``MoonLikeParams`` / ``EmbeddingParams`` / ``NucleusDependentParams`` / ``DynamicFilterParams``
NamedTuples as field-name maps (model/wave_function.py:54-57, model/new_sparse_model.py:134-136,
218-225, model/utils.py:372-375), tuples as "0", "1", ... maps, arrays as msgpack extension type
1 with (shape, dtype name, bytes), and the flax module names of the reference source
(``to_orbitals_up`` / ``env_up`` orbitals.py:53-81, ``beta_ee`` new_sparse_model.py:176,
``Linear_i`` of SparseMLP sparse_fwd_lap.py:209-225, ``Dense_i`` of MLP model/utils.py:25-52,
``pairwise_cusps`` / ``mlp`` / ``att`` jastrow.py:202-230).
"""
from __future__ import annotations

import msgpack
import numpy as np
import pytest

from sparse_wf_amd import checkpoint as C
from sparse_wf_amd import params as P
from tests.cases import small_cfg


def _arr(a):
    a = np.asarray(a)
    return msgpack.ExtType(1, msgpack.packb((list(a.shape), a.dtype.name, a.tobytes("C")), use_bin_type=True))


def _flax_tree(x):
    """State dict with every array as an extension object; lists become index maps."""
    if isinstance(x, dict):
        return {k: _flax_tree(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return {str(i): _flax_tree(v) for i, v in enumerate(x)}
    return _arr(np.asarray(x, np.float32))


def _blob(params, electrons, step=1234, legacy_namedtuple=False, drop_cutoff_param=False):
    emb = params["embedding"]
    emb_sd = {k: _flax_tree(emb[k]) for k in C.EMBEDDING_FIELDS if not (drop_cutoff_param and k == "cutoff_param")}
    moon = {"embedding": emb_sd, "to_orbitals": _flax_tree(params["to_orbitals"]),
            "jastrow": _flax_tree(params["jastrow"])}
    if legacy_namedtuple:
        moon = {"name": "MoonLikeParams", "fields": {str(i): moon[k] for i, k in enumerate(C.PARAM_FIELDS)}}
    state = {
        "key": _arr(np.array([0, 42], np.uint32)),
        "params": moon,
        "electrons": _arr(np.asarray(electrons, np.float32)),
        "opt_state": {"opt": {"0": {"count": _arr(np.int32(7))}}, "natgrad": {}},
        "width_state": {"width": _arr(np.float32(0.31)), "pmoves": _arr(np.zeros(10, np.float32)),
                        "i": _arr(np.int32(3))},
        "spin_state": None,
        "step": _arr(np.int32(step)),
    }
    return msgpack.packb(state, use_bin_type=True)


@pytest.fixture(scope="module")
def setup():
    cfg = small_cfg(False)
    prm = P.init_params(cfg, seed=5)
    prm["embedding"]["scales"] = [np.float32(0.7), np.float32(1.3)]
    prm["embedding"]["cutoff_param"] = np.float32(2.5)
    P.exercise_jastrow(prm)
    el = np.random.default_rng(0).normal(size=(6, cfg.n_el, 3)).astype(np.float32)
    return cfg, prm, el


def _same_packed(cfg, a, b):
    pa, pb = P.pack_params(cfg, a), P.pack_params(cfg, b)
    assert len(pa) == len(pb)
    for x, y in zip(pa, pb):
        np.testing.assert_array_equal(x, y)


def test_reads_reference_layout(setup):
    cfg, prm, el = setup
    st = C.load_training_state(_blob(prm, el), cfg)
    assert st["step"] == 1234 and abs(st["width"] - 0.31) < 1e-7
    np.testing.assert_array_equal(st["electrons"], el)
    np.testing.assert_array_equal(st["key"], [0, 42])
    for (pa, a), (pb, b) in zip(P.leaves_with_path(prm), P.leaves_with_path(st["params"])):
        assert pa == pb
        np.testing.assert_array_equal(np.asarray(a, np.float32), np.asarray(b))
    _same_packed(cfg, prm, st["params"])


def test_reads_legacy_namedtuple_layout(setup):
    cfg, prm, el = setup
    st = C.load_training_state(_blob(prm, el, legacy_namedtuple=True), cfg)
    _same_packed(cfg, prm, st["params"])


def test_checkpoint_without_cutoff_param(setup):
    """As src/convert_checkpoint.py:91-101: the learned cutoff gets its initial value."""
    cfg, prm, el = setup
    st = C.load_training_state(_blob(prm, el, drop_cutoff_param=True), cfg)
    assert float(st["params"]["embedding"]["cutoff_param"]) == 10.0
    packed = P.pack_params(cfg, st["params"])
    lc = packed[-1][0]  # learned cutoff = cutoff * sigmoid(10)
    assert abs(lc - cfg.cutoff / (1 + np.exp(-10.0))) < 1e-6


def test_chunked_electrons(setup):
    cfg, prm, el = setup
    blob = C.save_training_state(prm, el, step=9)
    small = C.msgpack_serialize(C.msgpack_restore(blob), chunk_bytes=64)  # split arrays into 16-float chunks
    assert small != blob
    st = C.load_training_state(small, cfg)
    np.testing.assert_array_equal(st["electrons"], el)
    _same_packed(cfg, prm, st["params"])


def test_writer_round_trip(setup, tmp_path):
    cfg, prm, el = setup
    path = tmp_path / "chkpt000009.msgpk"
    path.write_bytes(C.save_training_state(prm, el, step=9, width=0.2))
    st = C.load_training_state(str(path), cfg)
    assert st["step"] == 9 and abs(st["width"] - 0.2) < 1e-7
    _same_packed(cfg, prm, st["params"])


def test_mismatch_names_the_entry(setup):
    cfg, prm, el = setup
    bad = P.tree_map(lambda x: x, prm)
    bad["to_orbitals"]["params"]["to_orbitals_up"]["kernel"] = np.zeros((3, 3), np.float32)
    with pytest.raises(ValueError, match="to_orbitals_up/kernel"):
        C.load_training_state(_blob(bad, el), cfg)
    del bad["jastrow"]["params"]["att"]
    bad["to_orbitals"] = prm["to_orbitals"]
    with pytest.raises(ValueError, match="lacks"):
        C.load_training_state(_blob(bad, el), cfg)
    with pytest.raises(ValueError, match="electrons"):
        C.load_training_state(_blob(prm, el[:, :3]), cfg)


def test_bfloat16_arrays_widen_exactly():
    x = np.array([1.0, -2.5, 3.140625], np.float32)
    bits = (x.view(np.uint32) >> 16).astype(np.uint16)
    payload = msgpack.packb(([3], "bfloat16", bits.tobytes()), use_bin_type=True)
    out = C.msgpack_restore(msgpack.packb({"a": msgpack.ExtType(1, payload)}))
    np.testing.assert_array_equal(out["a"], x)
