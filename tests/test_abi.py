"""The C ABI library loads without a GPU and exports every entry point include/fire.h
declares, with the ctypes signatures' argument counts matching the header."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

from sparse_wf_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "fire.h")


def _declarations():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decl = {}
    for m in re.finditer(r"^\s*(?:int|const char\*)\s+(fire_\w+)\s*\(([^)]*)\)\s*;", text, flags=re.M):
        args = [a.strip() for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        decl[m.group(1)] = len(args)
    return decl


def test_header_parses():
    d = _declarations()
    assert "fire_create" in d and "fire_local_energy" in d and "fire_mcmc_steps" in d
    assert len(d) >= 15


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in _declarations():
        assert hasattr(lib, name), f"{name} declared in fire.h but not exported"


def test_ctypes_signatures_match_header():
    d = _declarations()
    assert set(d) == set(_lib.SIGNATURES), set(d) ^ set(_lib.SIGNATURES)
    for name, n in d.items():
        assert len(_lib.SIGNATURES[name][1]) == n, name


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="library not built")
def test_errors_are_reported_without_gpu():
    """Argument validation runs before any device call: a bad config fails with a message."""
    lib = _lib.load()
    cfg = _lib.FireConfigC(n_el=4, n_up=2, n_nuc=1, feature_dim=100, filter_w0=4, filter_w1=4, n_env_en=4,
                           n_det=1, n_env_iso=2, e_e_cusp=0, use_mlp_jastrow=0, use_log_jastrow=0,
                           use_attention=0, mlp_width=8, mlp_depth=1, att_heads=1, att_dim=1,
                           cutoff=3.0, cutoff_1el=20.0)
    R = (ctypes.c_float * 3)(0.0, 0.0, 0.0)
    Z = (ctypes.c_float * 1)(2.0)
    h = ctypes.c_void_p()
    rc = lib.fire_create(ctypes.byref(cfg), R, Z, 0, ctypes.byref(h))
    assert rc != 0
    assert b"feature_dim" in lib.fire_last_error()
    assert lib.fire_state_stats(None, None, None, None, None) != 0
    assert b"null handle" in lib.fire_last_error()
