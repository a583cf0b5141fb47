"""fire_neighbours_noalloc (static scratch, no device allocation) against fire_neighbours,
called through the C ABI: bit-identical counts, neighbour lists and pair totals on two calls
in a row, the second with a smaller batch, and a batch beyond the static capacity rejected
before any launch."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from sparse_wf_amd import _lib
from sparse_wf_amd.wave_function import FireWaveFunction, _dptr, _stream_ptr

from cases import alkane_cfg, walkers

pytestmark = pytest.mark.gpu


def _call(wf, fn, r, cap):
    B, N = r.shape[0], wf.cfg.n_el
    counts = torch.full((B, N), -1, dtype=torch.int32, device=wf.device)
    nbr = torch.full((cap,), -1, dtype=torch.int32, device=wf.device)
    n = ctypes.c_int64(-1)
    _lib.check(fn(wf._h, _stream_ptr(), _dptr(r), B, _dptr(counts), _dptr(nbr), cap, ctypes.byref(n)))
    torch.cuda.synchronize()
    return counts.cpu(), nbr.cpu(), n.value


def test_noalloc_matches_fire_neighbours(gpu_device):
    cfg = alkane_cfg(cutoff=5.0)
    wf = FireWaveFunction(cfg, device=gpu_device)
    lib = _lib.load()
    for B, seed in ((64, 1), (24, 2)):  # the second call is smaller: stale scratch must not leak in
        r = torch.as_tensor(walkers(cfg, B, seed=seed), device=wf.device)
        cap = 64 * B * cfg.n_el
        c_ref, n_ref, p_ref = _call(wf, lib.fire_neighbours, r, cap)
        c_na, n_na, p_na = _call(wf, lib.fire_neighbours_noalloc, r, cap)
        print(f"  B={B}: {p_na} pairs", flush=True)
        assert p_na == p_ref > 0
        assert torch.equal(c_na, c_ref)
        assert torch.equal(n_na[:p_na], n_ref[:p_ref])
        assert torch.all(n_na[p_na:] == -1)


def test_noalloc_rejects_beyond_capacity(gpu_device):
    cfg = alkane_cfg(cutoff=5.0)
    wf = FireWaveFunction(cfg, device=gpu_device)
    B = (1 << 21) // cfg.n_el + 1
    counts = torch.empty(1, dtype=torch.int32, device=wf.device)  # never written: rejected first
    r = torch.empty((1, 3), device=wf.device)
    n = ctypes.c_int64(0)
    st = _lib.load().fire_neighbours_noalloc(wf._h, _stream_ptr(), _dptr(r), B, _dptr(counts), None, 0,
                                             ctypes.byref(n))
    assert st != 0 and b"static scratch capacity" in _lib.load().fire_last_error()
