"""Every tile form of the split contractions gives the same local energy and log|psi|.

The autotuner picks a tile form per contraction shape by timing, so any form may end up in use.
Each form is forced in turn for every contraction of the handle (``autotune(v)``; where a form
does not apply to a contraction, the base tile runs) and compared with the base tile (form 0).
Every form accumulates each output over K in the same chunk order with the same products, so
the results agree bit for bit (the bounds below leave room for rounding); the handle's tile
table shows that each form did run.
F = W = 256 gives K = 256 contractions with N = 256, 272 (joint Jastrow layer with attention
values), 512 and the fused activation epilogues; the value path (``signed``) runs the
contractions of the value rows (their FP16 planes with row maxima taken from the rows).
"""
from __future__ import annotations

import numpy as np
import pytest

from sparse_wf_amd.wave_function import FireWaveFunction

from cases import make_params, small_cfg, walkers

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(gpu_device):
    cfg = small_cfg(False, feature_dim=256, mlp_width=256)
    prm = make_params(cfg, seed=0)
    wf = FireWaveFunction(cfg, device=gpu_device)
    wf.upload(prm)
    r = walkers(cfg, 8, seed=5)
    return cfg, prm, wf, r


def _run(wf, prm, r, form):
    wf.autotune(form)
    out = {}
    for prec in ("mixed", "mixed_fast"):
        t = wf.local_energy_terms(prm, r, precision=prec)
        out[prec] = (t["e_loc"].cpu().numpy(), t["logpsi"].cpu().numpy())
    _, lp = wf.signed(prm, r, precision="mixed")
    out["value"] = (None, lp.cpu().numpy())
    return out


def test_every_tile_form_matches_the_base_tile(setup):
    cfg, prm, wf, r = setup
    n = FireWaveFunction.tile_forms()
    assert n >= 27
    ref = _run(wf, prm, r, 0)
    worst = {}
    used = set()
    try:
        for form in range(1, n):
            print(f"[tile form {form}]", flush=True)
            got = _run(wf, prm, r, form)
            used |= {c[4] for c in wf.tile_choices()}
            for key, (e_ref, lp_ref) in ref.items():
                e, lp = got[key]
                dlp = float(np.max(np.abs(lp - lp_ref)))
                assert dlp <= 2e-4, (form, key, dlp)
                de = 0.0
                if e is not None:
                    de = float(np.max(np.abs(e - e_ref) / (1.0 + np.abs(e_ref))))
                    assert de <= 2e-4, (form, key, de)
                w = worst.setdefault(key, [0.0, 0.0])
                w[0], w[1] = max(w[0], dlp), max(w[1], de)
    finally:
        wf.autotune("on")
    print("largest differences from form 0 (log|psi| abs, E_loc rel):", worst)
    # every form applies to some contraction of this configuration, except the 2 x 2
    # A-in-registers form (15), which is offered to single BF16 planes only
    assert used >= set(range(n)) - {15}, sorted(set(range(n)) - used)
