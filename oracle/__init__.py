"""CPU oracle for the FiRE hot path -- TEST INFRASTRUCTURE ONLY.

This package restates the reference's algorithm (MScherbela/sparse_wf, FiRE
``new_sparse`` model) on the CPU with PyTorch in float64 or float32.  It exists to check
the HIP implementation in ``sparse_wf_amd``; only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and never as the thing that
is measured or shipped.  Nothing under ``sparse_wf_amd/`` imports it.

Modules
-------
``fwdlap``   dense forward-Laplacian arrays (the role folx plays in the reference).
``fire``     value path, sparse forward-Laplacian path, energies, low-rank update.

Pinning (tests/test_oracle.py):

* forward-Laplacian results vs a dense ``torch.func`` Hessian trace of the value path
  (float64, small N) -- the reference's own test strategy (tests/test_fwd_lap.py:57-117);
* low-rank update vs full recompute (tests/test_low_rank.py:29-61);
* known answers: hydrogen-like atoms (E_loc = -Z^2/2 exactly), Coulomb energies of
  hand-placed charges, the cutoff polynomial's closed-form boundary values, silu
  derivatives, slogdet vs numpy, neighbour lists vs brute force;
* Woodbury vs Sherman-Morrison vs direct slogdet.

No golden vectors exist in the reference (its tests are all relative), and the reference
cannot be imported in this environment (no jax/flax/folx), so nothing is checked against
reference outputs directly.
"""
