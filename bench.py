"""Benchmark: FiRE local-energy walkers/s on one MI355X.

Workload (default): zigzag alkane C16H34, 130 electrons (65/65), 1024 walkers, cutoff 5 bohr,
F = 256, 4 determinants, default Jastrow (MLP 256x2 + 16x16 attention, psiformer cusp).
A step = one sparse forward-Laplacian local-energy evaluation (kinetic + Coulomb) of the
whole batch, walkers resident in HBM.  The mixed-precision path is the headline value; the
plain-FP32 kernels are timed the same way beside it.

Prints one JSON line (see the field list at the bottom).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "local-energy walkers/s at 130 e-, mixed vs plain FP32 at <=1 mHa; p99 latency"
PEAK = {"bf16": 2500.0, "fp32": 157.3, "fp64": 78.6, "hbm_gbs": 8000.0}  # MI355X dense peaks (TFLOP/s, GB/s)


def algorithmic_flops(cfg, counts_per_walker):
    """Algorithmic FLOPs per walker of the local energy (formula terms: h0, edges+messages,
    update MLP, orbitals, envelopes, FP64 determinant part, Jastrow, Coulomb)."""
    N, M, F = cfg.n_el, cfg.n_nuc, cfg.feature_dim
    w0, w1 = cfg.pair_mlp_widths
    Een, D, Ei = cfg.pair_n_envelopes, cfg.n_determinants, cfg.n_envelopes
    W, HD = cfg.mlp_width, cfg.attention_heads * cfg.attention_dim
    c = np.asarray(counts_per_walker, np.float64)  # [N] neighbour counts
    P = c.sum()
    T = ((c + 1) ** 2).sum()
    R = 2 * N + 3 * (N + P)
    K = M
    terms = {
        "h0": 5 * N * (16 * F * F + K * 2 * (4 * w0 + w0 * w1 + Een * w1 + 12 * F)),
        "edges_messages": 16 * P * (4 * w0 + 2 * w0 * w1 + 12 * F) + 48 * P * F,
        "update_mlp": 4 * R * F * F,
        "orbitals": 2 * R * F * N * D,
        "envelopes": 10 * N * M * D * Ei * N,
        "det_fp64": D * (2.7 * N ** 3 + 6 * T * N + 2 * N * N),
        "jastrow": 2 * R * (F * W + W * W + 2 * W) + 4 * R * F * HD,
        "coulomb": 10 * (N * N / 2 + N * M),
    }
    return float(sum(terms.values())), {k: float(v) for k, v in terms.items()}, dict(P=P, T=T, R=R)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c16h34")
    ap.add_argument("--walkers", type=int, default=None)
    ap.add_argument("--cpu-sample", type=int, default=1, help="walkers timed on the CPU oracle (FP32)")
    ap.add_argument("--bias-sample", type=int, default=8, help="walkers checked against the FP64 oracle")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (walkers would shard "
                                   "across GPUs as independent data-parallel batches)"}))
        sys.exit(2)

    import torch

    from sparse_wf_amd import systems
    from sparse_wf_amd import params as P
    from sparse_wf_amd.config import FireConfig
    from sparse_wf_amd.wave_function import FireWaveFunction
    from sparse_wf_amd import _lib

    mol, over, n_walk, sigma = systems.workload(args.workload)
    B = args.walkers or n_walk
    cfg = FireConfig(mol=mol, **over)
    r_np = systems.init_electrons(mol, B, stddev=sigma, seed=0)
    wf = FireWaveFunction(cfg, device=0)
    params = wf.init(0, r_np[0])
    P.exercise_jastrow(params)
    wf.upload(params)
    r = torch.as_tensor(r_np, device="cuda")
    counts, _, _ = wf.neighbours(r)
    mean_counts = counts.double().mean(0).cpu().numpy()
    flops_w, flop_terms, sizes = algorithmic_flops(cfg, mean_counts)

    def run(prec, steps, warmup, profile=False):
        for _ in range(warmup):
            wf.local_energy_terms(params, r, precision=prec)
        torch.cuda.synchronize()
        if profile:
            _prof(True)
        lat = []
        t0 = time.perf_counter()
        for _ in range(steps):
            ts = time.perf_counter()
            out = wf.local_energy_terms(params, r, precision=prec)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - ts)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        prof = _prof(False) if profile else None
        return out, dt, lat, prof

    import ctypes

    def _prof(enable):
        ms, fl, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        _lib.check(wf._lib.fire_profile_gemm(wf._h, int(enable), ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(n)))
        return ms.value, fl.value, n.value

    out_m, dt_m, lat_m, prof_m = run("mixed", args.steps, args.warmup, profile=True)
    out_f, dt_f, lat_f, _ = run("fp32", max(1, args.steps // 2), 1)
    wps_m = B * args.steps / dt_m
    wps_f = B * max(1, args.steps // 2) / dt_f

    # dominant kernel family: the MFMA contractions (HIP events on the launch stream)
    g_ms, g_flops, g_n = prof_m
    g_achieved = (g_flops / (g_ms * 1e-3)) / 1e12 if g_ms > 0 else None
    roofline = {
        "bound": "mfma",
        "kernel": "k_gemm_bf16 (all dense contractions of one step, BF16 MFMA)",
        "achieved": g_achieved,
        "peak": PEAK["bf16"],
        "unit": "TFLOP/s",
        "frac": (g_achieved / PEAK["bf16"]) if g_achieved else None,
        "traffic": None,
        "launches_timed": g_n,
        "whole_step_tflops": flops_w * wps_m / 1e12,
    }

    e_m = out_m["e_loc"].cpu().numpy()
    e_f = out_f["e_loc"].cpu().numpy()
    result = {
        "metric": METRIC,
        "value": wps_m,
        "unit": "walkers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt_m / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded walkers, random-init params in the reference layout)",
        "config": {"workload": f"{mol.name} N={cfg.n_el} walkers={B} cutoff={cfg.cutoff} F={cfg.feature_dim} "
                               f"D={cfg.n_determinants}", "global_batch": B, "parallelism": "none (1 GPU)"},
        "fp32_walkers_per_s": wps_f,
        "mixed_over_fp32": wps_m / wps_f,
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat_m, 50)), "p99": 1e3 * float(np.percentile(lat_m, 99)),
                       "what": "one local-energy call of the whole batch"},
        "mixed_minus_fp32_mean_mHa": 1e3 * float(np.mean(e_m - e_f)),
        "algorithmic_gflop_per_walker": flops_w / 1e9,
        "graph": {"pairs_per_walker": sizes["P"], "triplets_per_walker": sizes["T"], "rows_per_walker": sizes["R"]},
        "roofline": roofline,
    }

    if not args.no_cpu:
        # CPU baseline leg: the oracle (FP32) timed on a bounded sample of the same walkers,
        # and the FP64 oracle as the reference for the energy bias of both GPU paths.
        from oracle import fire

        torch.set_num_threads(len(os.sched_getaffinity(0)))
        p32 = fire.to_torch(params, torch.float32)
        t0 = time.perf_counter()
        for w in range(args.cpu_sample):
            fire.local_energy(p32, cfg, r_np[w])
        dt_cpu = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": args.cpu_sample / dt_cpu, "unit": "walkers/s",
                                  "cores": torch.get_num_threads(), "kind": "port",
                                  "sample": f"{args.cpu_sample} walker(s) of the same batch, oracle in FP32 "
                                            f"(torch CPU), {dt_cpu:.1f} s"}
        if args.bias_sample > 0:
            p64 = fire.to_torch(params, torch.float64)
            ref = np.array([fire.local_energy(p64, cfg, r_np[w])[0] for w in range(args.bias_sample)])
            dm = e_m[: args.bias_sample] - ref
            dfp = e_f[: args.bias_sample] - ref
            result["energy_error_vs_fp64_mHa"] = {
                "walkers": args.bias_sample,
                "mixed": {"bias": 1e3 * float(np.mean(dm)), "rms": 1e3 * float(np.sqrt(np.mean(dm ** 2))),
                          "max": 1e3 * float(np.max(np.abs(dm)))},
                "fp32": {"bias": 1e3 * float(np.mean(dfp)), "rms": 1e3 * float(np.sqrt(np.mean(dfp ** 2))),
                         "max": 1e3 * float(np.max(np.abs(dfp)))},
            }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
