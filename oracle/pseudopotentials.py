"""CPU restatement of the pseudopotential energy terms (oracle; test infrastructure only).

Follows make_local_pseudopotential, build_all_integration_points and
make_nonlocal_pseudopotential (src/sparse_wf/pseudopotentials.py:204-218, 269-398) for one
walker, in FP64 numpy, on tables in the layout of ``sparse_wf_amd.pseudopotentials`` (radial
grid, local and non-local channels per ECP nucleus, cutoffs, quadrature directions).  The
psi ratios come from a caller-supplied ``signed(r) -> (sign, log|psi|)`` (the FiRE oracle, or an
analytic wavefunction in the known-answer tests).

The reference draws the quadrature rotations from JAX's threefry stream, which cannot run
here; the HIP kernel draws them from Philox4x32-10 (key (seed, walker), counter (electron,
0x45435000 + block), Box-Muller), restated in :func:`rotation` with ``oracle.mcmc``'s Philox.
"""
from __future__ import annotations

import math

import numpy as np

from .mcmc import _u01, philox4x32_10


def rotation(seed: int, walker: int, electron: int) -> np.ndarray:
    """Rows v1, v2, v3 of random_rot_mat (pseudopotentials.py:220-226) from the kernel's
    Philox normals."""
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    key = np.array([(seed & 0xFFFFFFFF) ^ (seed >> 32), walker], np.uint64)
    n = []
    for b in range(2):
        a = philox4x32_10(np.array([electron, 0x45435000 + b, 0, 0], np.uint64), key)
        r1 = math.sqrt(-2.0 * math.log(_u01(a[0])))
        r2 = math.sqrt(-2.0 * math.log(_u01(a[2])))
        t1, t2 = 2.0 * math.pi * _u01(a[1]), 2.0 * math.pi * _u01(a[3])
        n += [r1 * math.cos(t1), r1 * math.sin(t1), r2 * math.cos(t2), r2 * math.sin(t2)]
    v1 = np.array(n[0:3]); v2 = np.array(n[3:6])
    v1 = v1 / np.linalg.norm(v1)
    v2 = v2 - v1 * np.dot(v1, v2)
    v2 = v2 / np.linalg.norm(v2)
    return np.stack([v1, v2, np.cross(v1, v2)])


def legendre(l: int, x):
    """P_0 .. P_3 (leg_l0 .. leg_l3, pseudopotentials.py:229-242)."""
    return [np.ones_like(x), x, 0.5 * (3 * x ** 2 - 1), 0.5 * (5 * x ** 3 - 3 * x)][l]


def dist_f32(a, b):
    """float32 |a - b|, every operation rounded to float32 ((dx^2 + dy^2) + dz^2)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    d = (a - b).astype(np.float32)
    s = (d[..., 0] * d[..., 0]).astype(np.float32) + (d[..., 1] * d[..., 1]).astype(np.float32)
    s = s.astype(np.float32) + (d[..., 2] * d[..., 2]).astype(np.float32)
    return np.sqrt(s.astype(np.float32)).astype(np.float32)


def local_energy_pp(r, R, tab) -> float:
    """Sum over electrons and ECP nuclei of the interpolated local channel."""
    r = np.asarray(r, np.float32)
    R = np.asarray(R, np.float32)
    v = 0.0
    for e, J in enumerate(tab["atoms"]):
        d = dist_f32(r, R[J]).astype(np.float64)
        v += float(np.interp(d, tab["grid"], tab["v_loc"][e]).sum())
    return v


def integration_points(r, R, tab, seed: int, walker: int):
    """Active (electron, point float32 [3], coefficient) of one walker: the point on the
    closest ECP nucleus's rotated quadrature at the electron's distance, the coefficient
    sum_l (2l+1) V_l(|r' - R|) P_l(cos theta) / n_quad."""
    r = np.asarray(r, np.float32)
    R = np.asarray(R, np.float32)
    out = []
    for i in range(r.shape[0]):
        dists = [dist_f32(r[i], R[J]) for J in tab["atoms"]]
        e = int(np.argmin(dists))
        J = tab["atoms"][e]
        d = dists[e]
        if not float(d) < tab["cutoffs"][e]:
            continue
        rot = rotation(seed, walker, i)
        direction = ((r[i] - R[J]).astype(np.float32) / d).astype(np.float32).astype(np.float64)
        n = int(tab["n_quad"][e])
        for q in range(n):
            g = rot @ tab["quad_dirs"][e, q]
            p = (R[J].astype(np.float64) + float(d) * g).astype(np.float32)
            cs = float(g @ direction)
            dae = float(np.linalg.norm(p.astype(np.float64) - R[J]))
            c = sum((2 * l + 1) * float(np.interp(dae, tab["grid"], tab["v_nl"][e, l])) * float(legendre(l, cs))
                    for l in range(tab["v_nl"].shape[1])) / n
            out.append((i, p, c))
    return out


def nonlocal_energy_pp(signed, r, R, tab, seed: int, walker: int, return_scale: bool = False):
    """(V_nl, n_points) of one walker: the coefficients times psi(r') / psi(r).  With
    ``return_scale`` also sum |coefficient * ratio|, the magnitude the terms cancel from."""
    r = np.asarray(r, np.float32)
    s0, l0 = signed(r)
    v = scale = 0.0
    pts = integration_points(r, R, tab, seed, walker)
    for i, p, c in pts:
        rr = r.copy()
        rr[i] = p
        s, lp = signed(rr)
        t = c * float(s) * float(s0) * math.exp(float(lp) - float(l0))
        v += t
        scale += abs(t)
    return (v, len(pts), scale) if return_scale else (v, len(pts))
