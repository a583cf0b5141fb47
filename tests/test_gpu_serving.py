"""Continuous-batching server: every request's result is the same alone and in the mix
(bit-identical: kernels work per walker, random numbers are keyed by (seed, walker, step),
tile shapes are chosen once per contraction shape and process), concurrent requests overlap
on the device, the handle pool is reused across requests, and admission respects the HBM
budget."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from sparse_wf_amd.serving import Server, footprint_bytes, latency_summary, poisson_requests

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def requests(gpu_device):
    reqs = poisson_requests(5, qps=200.0, seed=11, n_el=(10, 40), walkers=(32, 96), sweeps=(1, 2),
                            precision="mixed", feature_dim=64)
    print("\n[serving] N:", [q.cfg.n_el for q in reqs], "B:", [len(q.walkers) for q in reqs], flush=True)
    return reqs


def test_same_alone_and_in_mix(requests):
    alone = {}
    for q in requests:
        res = Server(max_active=1).run([dataclasses.replace(q, arrival=0.0)])
        assert len(res) == 1 and res[0].error is None, res[0].error
        alone[q.rid] = res[0]
    srv = Server(max_active=4)
    mix = {r.rid: r for r in srv.run(requests)}
    assert set(mix) == set(alone)
    for rid, a in alone.items():
        m = mix[rid]
        assert m.error is None, m.error
        np.testing.assert_array_equal(m.walkers, a.walkers)
        np.testing.assert_array_equal(m.e_loc, a.e_loc)
        assert np.isfinite(m.e_mean)
    assert srv.max_concurrent >= 2
    # the pool's handles are retargeted between requests, not created per request
    assert srv.handles_created <= 4 < len(requests) + 1
    # execution intervals of at least two requests overlapped
    iv = sorted((r.started, r.finished) for r in mix.values())
    assert any(iv[k + 1][0] < iv[k][1] for k in range(len(iv) - 1))
    s = latency_summary(list(mix.values()))
    print(f"  p50 {s['p50_s']:.3f} s p99 {s['p99_s']:.3f} s walkers/s {s['walkers_per_s']:.0f}", flush=True)


def test_admission_by_footprint(requests):
    """With a budget below two requests' footprint, requests run one at a time and all
    finish (a request larger than the whole budget still runs when nothing else does)."""
    need = min(footprint_bytes(q.cfg, len(q.walkers)) for q in requests)
    srv = Server(max_active=4, hbm_budget=need + 1)
    res = srv.run([dataclasses.replace(q, arrival=0.0) for q in requests[:3]])
    assert len(res) == 3 and all(r.error is None for r in res)
    assert srv.max_concurrent == 1
