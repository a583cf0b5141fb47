"""ctypes binding of ``libfire_amd.so`` (the C ABI declared in include/fire.h).

The library is built in-tree by ``sparse_wf_amd/build.py``; loading fails loudly when it is
missing -- there is no CPU fallback for the hot path.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfire_amd.so"

FIRE_FP32 = 0
FIRE_MIXED = 1
FIRE_BF16 = 2
FIRE_MIXED_FAST = 3


class FireConfigC(ctypes.Structure):
    _fields_ = [
        ("n_el", ctypes.c_int32), ("n_up", ctypes.c_int32), ("n_nuc", ctypes.c_int32),
        ("feature_dim", ctypes.c_int32),
        ("filter_w0", ctypes.c_int32), ("filter_w1", ctypes.c_int32),
        ("n_env_en", ctypes.c_int32), ("n_det", ctypes.c_int32), ("n_env_iso", ctypes.c_int32),
        ("e_e_cusp", ctypes.c_int32),
        ("use_mlp_jastrow", ctypes.c_int32), ("use_log_jastrow", ctypes.c_int32), ("use_attention", ctypes.c_int32),
        ("mlp_width", ctypes.c_int32), ("mlp_depth", ctypes.c_int32),
        ("att_heads", ctypes.c_int32), ("att_dim", ctypes.c_int32),
        ("cutoff", ctypes.c_float), ("cutoff_1el", ctypes.c_float),
    ]


# name -> (restype, argtypes); the exported symbol set of include/fire.h
_P = ctypes.c_void_p
SIGNATURES = {
    "fire_create": (ctypes.c_int, [ctypes.POINTER(FireConfigC), _P, _P, ctypes.c_int, ctypes.POINTER(_P)]),
    "fire_destroy": (ctypes.c_int, [_P]),
    "fire_reconfigure": (ctypes.c_int, [_P, ctypes.POINTER(FireConfigC), _P, _P]),
    "fire_workspace_bytes": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "fire_upload_params": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32]),
    "fire_neighbours": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32, _P, _P, ctypes.c_int64,
                                       ctypes.POINTER(ctypes.c_int64)]),
    "fire_neighbours_noalloc": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32, _P, _P, ctypes.c_int64,
                                               ctypes.POINTER(ctypes.c_int64)]),
    "fire_local_energy": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32, ctypes.c_int32, _P, _P, _P, _P, _P, _P, _P]),
    "fire_signed": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32, ctypes.c_int32, _P, _P]),
    "fire_embedding": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32, ctypes.c_int32, _P, _P]),
    "fire_profile_gemm": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_int64)]),
    "fire_state_init": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32, ctypes.c_int32, _P, _P]),
    "fire_state_propose": (ctypes.c_int, [_P, _P, ctypes.c_int32, _P, _P, _P]),
    "fire_state_commit": (ctypes.c_int, [_P, _P, _P]),
    "fire_mcmc_steps": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_float, ctypes.c_uint64, ctypes.c_int64, _P]),
    "fire_state_get": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "fire_state_stats": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "fire_autotune": (ctypes.c_int, [_P, ctypes.c_int32, _P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "fire_tile_forms": (ctypes.c_int, []),
    "fire_precision_plan": (ctypes.c_int, [_P, _P, ctypes.c_int32, _P]),
    "fire_mcmc_options": (ctypes.c_int, [_P, ctypes.c_int32]),
    "fire_det_options": (ctypes.c_int, [_P, ctypes.c_int32]),
    "fire_ecp_upload": (ctypes.c_int, [_P, ctypes.c_int32, _P, ctypes.c_int32, ctypes.c_int32, _P, _P, _P, _P, _P,
                                       ctypes.c_int32, _P, _P]),
    "fire_ecp_energy": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P, _P, ctypes.POINTER(ctypes.c_int64)]),
    "fire_energy_stats": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, _P, _P, _P]),
    "fire_grad_logpsi": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32, _P, _P]),
    "fire_param_count": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "fire_last_error": (ctypes.c_char_p, []),
}

_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library first (python -m sparse_wf_amd.build "
            "or __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class FireError(RuntimeError):
    pass


def check(status: int):
    if status != 0:
        msg = load().fire_last_error().decode(errors="replace")
        raise FireError(msg or f"fire call failed with status {status}")
