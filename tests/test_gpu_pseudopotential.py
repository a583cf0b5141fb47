"""Pseudopotential energy on the GPU (fire_ecp_energy) against the oracle: C2H6 with carbon
ECPs (14 valence electrons), the same Philox quadrature rotations, psi ratios of the GPU's
low-rank proposals against full FP64 oracle evaluations.  Bounds: local part 1e-10 relative
(the same FP64 tables, another summation order); non-local part 3e-5 (FP32) or 1e-4 (mixed)
of the sum of |coefficient x ratio| it cancels from: the GPU's log|psi| of each point is off
the FP64 one by up to about 3e-5 in FP32, and these terms cancel to a sum up to 60 times
smaller (measured on MI355X: the GPU sum equals the oracle's points and coefficients with the
GPU's own fresh FP32 ratios to 1e-8)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import fire
from oracle import pseudopotentials as O
from sparse_wf_amd import pseudopotentials as P
from sparse_wf_amd.wave_function import FireWaveFunction, make_local_energy
from tests.cases import make_params
from tests.ecp_cases import CARBON_ECP, ethane_valence, valence_walkers

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ecp_setup(gpu_device):
    cfg = ethane_valence()
    prm = make_params(cfg, seed=0, r_init=valence_walkers(cfg, 1, seed=9)[0])
    pp = P.make_pseudopotential(cfg.mol.Z, {"C": CARBON_ECP})
    wf = FireWaveFunction(cfg, device=gpu_device)
    wf.set_pseudopotential(pp)
    tab = {"atoms": pp.atoms, "grid": pp.grid, "v_loc": pp.v_loc, "v_nl": pp.v_nl, "cutoffs": pp.cutoffs,
           "n_quad": pp.n_quad, "quad_dirs": pp.quad_dirs}
    return cfg, prm, pp, wf, tab


@pytest.mark.parametrize("precision,tol", [("fp32", 3e-5), ("mixed", 1e-4)])
def test_ecp_energy_matches_oracle(ecp_setup, precision, tol):
    cfg, prm, pp, wf, tab = ecp_setup
    B, seed = 6, 11
    r = valence_walkers(cfg, B, seed=1)
    out = wf.pseudopotential_energy(prm, r, seed=seed, precision=precision)
    pt = fire.to_torch(prm)
    R = np.asarray(cfg.mol.R, np.float32)
    n_tot = 0
    for w in range(B):
        vl = O.local_energy_pp(r[w], R, tab)
        vn, n, scale = O.nonlocal_energy_pp(lambda x: fire.signed(pt, cfg, x), r[w], R, tab, seed, w, True)
        n_tot += n
        gl, gn = float(out["v_loc"][w]), float(out["v_nl"][w])
        print(f"  {precision} walker {w}: V_loc {gl:.8f} vs {vl:.8f}, V_nl {gn:.8f} vs {vn:.8f} ({n} points)",
              flush=True)
        assert abs(gl - vl) <= 1e-6 * max(1.0, abs(vl))
        assert abs(gn - vn) <= tol * scale
    assert out["n_points"] == n_tot and n_tot > 0


def test_local_energy_with_pseudopotential(ecp_setup):
    cfg, prm, pp, wf, tab = ecp_setup
    r = valence_walkers(cfg, 4, seed=2)
    terms = wf.local_energy_terms(prm, r, precision="fp32")
    R = np.asarray(cfg.mol.R, np.float32)
    for w in range(4):  # Coulomb terms with the effective charges Z - n_core
        ref = fire.potential_energy(r[w], R, pp.effective_charges)
        assert abs(float(terms["e_pot"][w]) - ref) < 1e-9 * max(1.0, abs(ref))
    (_, lp0), _ = wf.log_psi_with_state(prm, r, precision="fp32")
    le = make_local_energy(wf, precision="fp32", pseudopotential=pp)
    e, _ = le(5, prm, r)
    ppe = wf.pseudopotential_energy(prm, r, seed=5, precision="fp32")
    np.testing.assert_allclose(e.cpu().numpy(), (terms["e_loc"] + ppe["v_loc"] + ppe["v_nl"]).cpu().numpy(),
                               rtol=1e-12, atol=1e-12)
    # the resident state is left as it was built (the proposals are dropped)
    lp_state = wf.state_values()[2]
    np.testing.assert_allclose(lp_state.cpu().numpy(), lp0.cpu().numpy(), rtol=0, atol=1e-12)
    wf.set_pseudopotential(None)
    t2 = wf.local_energy_terms(prm, r[:1], precision="fp32")
    ref = fire.potential_energy(r[0], R, cfg.mol.Z)
    assert abs(float(t2["e_pot"][0]) - ref) < 1e-9 * max(1.0, abs(ref))
    wf.set_pseudopotential(pp)
