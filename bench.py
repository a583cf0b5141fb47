"""Benchmark: FiRE local-energy walkers/s on one MI355X.

Workload (default): zigzag alkane C16H34, 130 electrons (65/65), 1024 walkers, cutoff 5 bohr,
F = 256, 4 determinants, default Jastrow (MLP 256x2 + 16x16 attention, psiformer cusp),
random-init parameters in the reference layout.

1. Sampling: the walkers are equilibrated by single-electron Metropolis sweeps (Gaussian
   steps of 0.3 bohr, FP64 Woodbury low-rank determinant updates, mixed precision) on the
   GPU; acceptance, throughput and the drift of the resident log psi are reported.
2. Timed step = one sparse forward-Laplacian local-energy evaluation (kinetic + Coulomb) of
   the whole batch, walkers resident in HBM.  The mixed-precision path is the headline value;
   the plain-FP32 kernels (and the faster two-plane split) are timed the same way beside it.
3. Energy error of each path against the FP64 oracle on a sample of the equilibrated
   walkers; the oracle in FP32 timed on the host cores as the CPU baseline.

Prints one JSON line (fields at the bottom of main()).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "local-energy walkers/s at 130 e-, mixed vs plain FP32 at <=1 mHa; p99 latency"
PEAK = {"bf16": 2500.0, "fp32": 157.3, "fp64": 78.6, "hbm_gbs": 8000.0}  # MI355X dense peaks (TFLOP/s, GB/s)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


TRAFFIC_PROFILE = "profiles/local_energy_c16h34_hbm_traffic_mfma.md"


def measured_traffic():
    """HBM bytes per split-contraction launch of the C16H34 mixed evaluation, from the
    rocprofv3 counter summary committed with the repository (None if it is absent)."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_PROFILE)) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def err_stats(d):
    d = np.asarray(d, np.float64) * 1e3
    return {"bias": float(np.mean(d)), "rms": float(np.sqrt(np.mean(d ** 2))), "median_abs": float(np.median(np.abs(d))),
            "max_abs": float(np.max(np.abs(d)))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c16h34")
    ap.add_argument("--walkers", type=int, default=None)
    ap.add_argument("--sweeps", type=int, default=50, help="Metropolis sweeps (n_el single-electron steps each)")
    ap.add_argument("--width", type=float, default=0.3, help="Gaussian step (bohr)")
    ap.add_argument("--cpu-sample", type=int, default=192, help="walkers timed on the CPU oracle (FP32)")
    ap.add_argument("--bias-sample", type=int, default=512, help="walkers checked against the FP64 oracle")
    ap.add_argument("--config-legs", default="h2o_like,h64",
                    help="other benchmark configurations evaluated (throughput, error against FP64); '' to skip")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--grad-walkers", type=int, default=1024,
                    help="walkers of the parameter-gradient leg (0: skip)")
    ap.add_argument("--ragged", type=int, default=32, help="geometries of the ragged-batch leg (0: skip)")
    ap.add_argument("--ragged-sweeps", type=int, default=2, help="timed Metropolis sweeps of the ragged leg")
    ap.add_argument("--serve-requests", type=int, default=8, help="requests of the serving leg (0: skip)")
    ap.add_argument("--serve-qps", type=float, default=4.0, help="Poisson arrival rate of the serving leg")
    ap.add_argument("--serve-nel", default="20,400", help="electron-count range of the serving requests")
    ap.add_argument("--serve-walkers", default="128,512", help="walker range of the serving requests")
    ap.add_argument("--serve-sweeps", default="10,12", help="sweep range of the serving requests")
    ap.add_argument("--plan", default=None,
                    help="operand planes per contraction class of the mixed precision, comma-separated "
                         "(ElecInit MLP, update MLP, orbitals, MLP Jastrow, attention Q/V, readout)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device copy of this many MiB (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (walkers would shard "
                                   "across GPUs as independent data-parallel batches)"}))
        sys.exit(2)

    import torch

    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 4
        src = torch.randn(n, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n, "launches": 3}))
        return

    from sparse_wf_amd import systems
    from sparse_wf_amd import params as P
    from sparse_wf_amd.config import FireConfig
    from sparse_wf_amd.wave_function import FireWaveFunction
    from sparse_wf_amd import _lib

    mol, over, n_walk, sigma = systems.workload(args.workload)
    B = args.walkers or n_walk
    cfg = FireConfig(mol=mol, **over)
    N = cfg.n_el
    r0 = systems.init_electrons(mol, B, stddev=sigma, seed=0)
    wf = FireWaveFunction(cfg, device=0)
    params = wf.init(0, r0[0])
    P.exercise_jastrow(params)
    wf.upload(params)
    plan = wf.precision_plan([int(x) for x in args.plan.split(",")] if args.plan else None)

    # 1. equilibrate with the GPU sampler (mixed precision state)
    mcmc = None
    r = torch.as_tensor(r0, device="cuda")
    if args.sweeps > 0:
        wf.log_psi_with_state(params, r, precision="mixed")
        n_acc = torch.zeros(B, dtype=torch.int32, device="cuda")
        t_mc = 0.0
        t_first = None
        for s in range(args.sweeps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            # the resident state is rebuilt every 10 sweeps (inside the timed region): long chains
            # of FP64 low-rank updates drift on the few ill-conditioned walkers otherwise
            r, st = wf.mcmc_steps_low_rank(params, r, N, width=args.width, seed=1, step0=s * N,
                                           keep_state=(s % 10 != 0))
            torch.cuda.synchronize()
            if s == 0:
                t_first = time.perf_counter() - t0  # includes capturing the sweep graph
            else:
                t_mc += time.perf_counter() - t0
            n_acc += st["n_accept"]
            if s == 0 or (s + 1) % 10 == 0:
                log(f"sweep {s + 1}/{args.sweeps}: pmove {st['pmove']:.3f}, {t_first + t_mc:.1f} s")
        timed = max(args.sweeps - 1, 1)
        if args.sweeps == 1:
            t_mc = t_first
        st = wf.state_stats()
        lp_state = wf.state_values()[2]
        lp_fresh = wf.signed(params, r, precision="mixed")[1]
        mcmc = {
            "sweeps": args.sweeps, "steps": args.sweeps * N, "width_bohr": args.width,
            "walker_moves_per_s": B * N * timed / t_mc, "ms_per_step": 1e3 * t_mc / (N * timed),
            "first_sweep_s": t_first, "timed_sweeps": timed,
            "pmove": float(n_acc.double().mean()) / (N * args.sweeps),
            "changed_electrons_per_move": {"mean": st["mean_changed"], "max": st["max_changed"]},
            "logpsi_drift_max": float((lp_state - lp_fresh).abs().max()),
            "what": "single-electron Metropolis steps, FP64 Woodbury updates of the resident state, mixed precision; "
                    "each sweep replayed from one captured hipGraph (the first sweep, which captures it, is timed "
                    "separately); the state is rebuilt from the positions every 10 sweeps (timed)",
        }

    counts, _, _ = wf.neighbours(r)
    mean_counts = counts.double().mean(0).cpu().numpy()
    flops_w, flop_terms, sizes = algorithmic_flops(cfg, mean_counts)

    def _prof(enable):
        ms, fl, by, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        _lib.check(wf._lib.fire_profile_gemm(wf._h, int(enable), ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(by),
                                             ctypes.byref(n)))
        return ms.value, fl.value, by.value, n.value

    # 2. timed local-energy steps
    def run(prec, steps, warmup, profile=False):
        for _ in range(warmup):
            wf.local_energy_terms(params, r, precision=prec)
        torch.cuda.synchronize()
        lat = []
        t0 = time.perf_counter()
        for _ in range(steps):
            ts = time.perf_counter()
            out = wf.local_energy_terms(params, r, precision=prec)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - ts)
        dt = time.perf_counter() - t0
        prof = None
        if profile:  # one more step with HIP events around every dense-contraction launch
            _prof(True)
            wf.local_energy_terms(params, r, precision=prec)
            torch.cuda.synchronize()
            prof = _prof(False)
        return out, dt, lat, prof

    log("timing mixed")
    out_m, dt_m, lat_m, prof_m = run("mixed", args.steps, args.warmup, profile=True)
    log("timing fp32")
    out_f, dt_f, lat_f, _ = run("fp32", args.steps, 1)
    log("timing mixed_fast")
    out_x, dt_x, _, _ = run("mixed_fast", max(1, args.steps // 2), 1)
    # the determinant Laplacian's triplet products in FP64 (FP64 transposed inverse on the FP64
    # matrix cores) instead of on an FP32 copy of the inverse
    log("timing mixed with FP64 determinant-Laplacian products")
    wf.det_options(True)
    out_d, dt_d, _, _ = run("mixed", max(1, args.steps // 2), 1)
    wf.det_options(False)
    wps_m = B * args.steps / dt_m
    wps_f = B * args.steps / dt_f
    wps_x = B * max(1, args.steps // 2) / dt_x
    wps_d = B * max(1, args.steps // 2) / dt_d

    # 3. per-walker parameter gradient of log|psi| (FP32 value path and its reverse mode; the
    # rows over the packed parameter tensors, as the preconditioner's Jacobian consumes them)
    grad_leg = None
    if args.grad_walkers > 0:
        Bg = min(B, args.grad_walkers)
        rg = r[:Bg].contiguous()
        log(f"timing parameter gradient ({Bg} walkers)")
        wf.grad_logpsi(params, rg, packed=True)  # sizes the workspace
        torch.cuda.synchronize()
        n_rep = 2
        t0 = time.perf_counter()
        for _ in range(n_rep):
            g_rows, lp_g = wf.grad_logpsi(params, rg, packed=True)
        torch.cuda.synchronize()
        dt_g = (time.perf_counter() - t0) / n_rep
        lp_v = wf.signed(params, rg, precision="fp32")[1]
        grad_leg = {
            "walkers": Bg, "parameters": int(g_rows.shape[1]), "seconds_per_batch": dt_g,
            "walkers_per_s": Bg / dt_g, "finite": bool(torch.isfinite(g_rows).all()),
            "logpsi_minus_fp32_value_max": float((lp_g - lp_v).abs().max()),
            "what": "d log|psi| / d theta per walker over every packed parameter float (fire_grad_logpsi), "
                    "FP32 forward with FP64 determinants and its reverse mode, walkers in chunks of up to 128",
        }
        del g_rows, lp_g
        torch.cuda.empty_cache()

    g_ms, g_flops, g_bytes, g_n = prof_m
    g_achieved = (g_flops / (g_ms * 1e-3)) / 1e12 if g_ms > 0 else None
    g_gbs = (g_bytes / (g_ms * 1e-3)) / 1e9 if g_ms > 0 else None
    traffic = measured_traffic() if args.workload == "c16h34" and B == 1024 else None
    # the contractions' algorithmic intensity (2MNK flops over A, C and the weights moved once)
    # is far below MI355X's ridge point (2.5 PFLOP/s over 8 TB/s = 312 flop/byte), so their
    # roofline ceiling is HBM; the matrix-core view is kept beside it
    roofline = {
        "bound": "hbm",
        "kernel": "k_gemm_bf16* (every dense contraction of one mixed step: ElecInit-MLP, update-MLP and "
                  "orbital contractions on two FP16 planes of power-of-two-scaled operands, Jastrow and "
                  "attention on a BF16 hi + lo pair, three products each; autotuned tile shapes; the "
                  "update-MLP and first Jastrow layers with the forward-Laplacian silu in their epilogue)",
        "achieved": g_gbs,
        "peak": PEAK["hbm_gbs"],
        "unit": "GB/s",
        "frac": (g_gbs / PEAK["hbm_gbs"]) if g_gbs else None,
        "traffic": traffic,
        "algorithmic_bytes_per_launch": g_bytes / g_n if g_n else None,
        "intensity_flop_per_byte": g_flops / g_bytes if g_bytes else None,
        "ridge_flop_per_byte": PEAK["bf16"] * 1e12 / (PEAK["hbm_gbs"] * 1e9),
        "mfma_view": {"achieved": g_achieved, "peak": PEAK["bf16"], "unit": "TFLOP/s",
                      "frac": (g_achieved / PEAK["bf16"]) if g_achieved else None,
                      "what": "algorithmic 2*M*N*K per launch; the matrix cores execute three 16-bit products "
                              "per multiply-add on top of these"},
        "traffic_source": TRAFFIC_PROFILE if traffic else None,
        "avg_launch_ms": g_ms / g_n if g_n else None,
        "launches_per_step": g_n,
        "algorithmic_flop_per_step": g_flops,
        "note": "achieved = algorithmic bytes per launch (A read once, C written once, the weight planes once: "
                "4*M*(K+N) + planes*2*N*K) over the HIP-event duration of the launches on their stream; "
                "traffic = HBM bytes per launch from two rocprofv3 --pmc runs "
                "(FETCH_SIZE x2, WRITE_SIZE x1 by a 1 GiB copy calibration) of this workload, read from the "
                "committed profile summary; whole-step rate in whole_step_tflops",
        "whole_step_tflops": flops_w * wps_m / 1e12,
    }

    # tile shapes the autotuner chose for the BF16 contractions of this run
    cap = 256
    table = (ctypes.c_int32 * (7 * cap))()
    n_tab = ctypes.c_int32()
    _lib.check(wf._lib.fire_autotune(wf._h, -1, table, cap, ctypes.byref(n_tab)))
    # log2 M carries +64 for the fused-activation epilogue and +32 for device-counted rows
    roofline["autotuned_tiles"] = [
        {"split": table[7 * i], "N": table[7 * i + 1], "K": table[7 * i + 2], "log2_M": table[7 * i + 3] % 32,
         "fused_silu": table[7 * i + 3] >= 64, "device_rows": (table[7 * i + 3] // 32) % 2 == 1,
         "tile": f"{table[7 * i + 5]}x{table[7 * i + 6]}"} for i in range(min(n_tab.value, cap))]

    e_m = out_m["e_loc"].cpu().numpy()
    e_f = out_f["e_loc"].cpu().numpy()
    e_x = out_x["e_loc"].cpu().numpy()
    e_d = out_d["e_loc"].cpu().numpy()
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
        "data": "synthetic (seeded walkers equilibrated by the GPU sampler, random-init params in the reference layout)",
        "config": {"workload": f"{mol.name} N={cfg.n_el} walkers={B} cutoff={cfg.cutoff} F={cfg.feature_dim} "
                               f"D={cfg.n_determinants}", "global_batch": B, "parallelism": "none (1 GPU)"},
        "fp32_walkers_per_s": wps_f,
        "mixed_over_fp32": wps_m / wps_f,
        "mixed_plan": plan,
        "mixed_fast_walkers_per_s": wps_x,
        "mixed_fp64_laplacian_walkers_per_s": wps_d,
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat_m, 50)), "p99": 1e3 * float(np.percentile(lat_m, 99)),
                       "what": "one local-energy call of the whole batch (mixed)"},
        "mixed_minus_fp32_mHa": err_stats(e_m - e_f),
        "mixed_fast_minus_fp32_mHa": err_stats(e_x - e_f),
        "algorithmic_gflop_per_walker": flops_w / 1e9,
        "gflop_terms_per_walker": {k: v / 1e9 for k, v in flop_terms.items()},
        "graph": {"pairs_per_walker": sizes["P"], "triplets_per_walker": sizes["T"], "rows_per_walker": sizes["R"]},
        "roofline": roofline,
        "mcmc": mcmc,
        "param_grad": grad_leg,
    }

    if args.ragged > 0:
        # ragged-batch leg: seeded random H/C/N/O chains of 10-200 electrons, 256 walkers each,
        # every geometry's sweep replayed from its own captured graph on its own stream (all in
        # flight at once), against the same steps issued eagerly one geometry after the other
        from sparse_wf_amd.ragged import random_geometries, run_ragged

        log(f"ragged leg: {args.ragged} geometries")
        geoms = random_geometries(args.ragged, seed=3, n_el=(10, 200), walkers=256, feature_dim=256)
        g_run = run_ragged(geoms, args.ragged_sweeps, precision="mixed", use_graphs=True, local_energy=True)
        e_run = run_ragged(geoms, args.ragged_sweeps, precision="mixed", use_graphs=False)
        same = all(np.array_equal(a.cpu().numpy(), b.cpu().numpy())
                   for a, b in zip(g_run["walkers"], e_run["walkers"]))
        result["ragged"] = {
            "workload": f"{args.ragged} random H/C/N/O chains, electrons uniform in [10, 200], 256 walkers each, "
                        f"F=256, D=4, cutoff 5, {args.ragged_sweeps} Metropolis sweeps (sigma 0.3 bohr), mixed",
            "electrons": g_run["electrons"],
            "walker_moves_per_s": g_run["walker_moves_per_s"], "seconds": g_run["seconds"],
            "eager_walker_moves_per_s": e_run["walker_moves_per_s"], "eager_seconds": e_run["seconds"],
            "graph_over_eager": g_run["walker_moves_per_s"] / e_run["walker_moves_per_s"],
            "pmove_mean": float(np.mean(g_run["pmove"])),
            "walkers_identical_graph_vs_eager": bool(same),
            "local_energy_walkers_per_s": g_run["local_energy_walkers_per_s"],
            "local_energy_seconds": g_run["local_energy_seconds"],
            "local_energy_what": "mixed local energy of every geometry's final walkers: two host threads, each "
                                 "with one evaluation handle and stream retargeted per geometry",
        }
        log(f"ragged leg: {g_run['walker_moves_per_s']:.3g} moves/s (graphs), "
            f"{e_run['walker_moves_per_s']:.3g} (eager)")
        del g_run, e_run, geoms
        torch.cuda.empty_cache()

    if args.serve_requests > 0:
        # serving leg: seeded Poisson stream of random H/C chains, continuous batching,
        # the same stream once in mixed precision and once in FP32
        from sparse_wf_amd.serving import Server, latency_summary, poisson_requests

        rng2 = lambda s: tuple(int(x) for x in s.split(","))
        serving = {"workload": f"{args.serve_requests} requests, Poisson {args.serve_qps}/s, random H/C chains "
                               f"N={args.serve_nel}, walkers={args.serve_walkers}, sweeps={args.serve_sweeps}, "
                               "F=256, D=4, cutoff 5; result = local energy after the sweeps; one host thread "
                               "dispatching sweep-graph replays over a pool of reusable handles; walker and sweep "
                               "ranges scaled down from 128-4096 walkers and 10-100 sweeps so the leg stays within a "
                               "minute (a 4096-walker, 400-electron, 100-sweep request alone is 164 M walker-moves, "
                               "about 7 minutes at the sampler's rate)"}
        for prec in ("mixed", "fp32"):
            log(f"serving leg ({prec}): preparing {args.serve_requests} requests")
            reqs = poisson_requests(args.serve_requests, args.serve_qps, seed=5, n_el=rng2(args.serve_nel),
                                    walkers=rng2(args.serve_walkers), sweeps=rng2(args.serve_sweeps),
                                    precision=prec)
            srv = Server(device=0)
            t0 = time.perf_counter()
            res = srv.run(reqs)
            log(f"serving leg ({prec}): {time.perf_counter() - t0:.1f} s")
            s = latency_summary(res)
            s["max_concurrent"] = srv.max_concurrent
            serving[prec] = s
        result["serving"] = serving
        result["latency_ms"]["serving_p50"] = 1e3 * serving["mixed"]["p50_s"]
        result["latency_ms"]["serving_p99"] = 1e3 * serving["mixed"]["p99_s"]

    if not args.no_cpu:
        # CPU baseline leg: the oracle (FP32) timed on a bounded sample of the same walkers over a
        # pool of single-threaded processes (one per host core of the share, at most 16), and the
        # FP64 oracle the same way as the reference for the energy error of the GPU paths.
        from oracle import pool

        procs = pool.default_procs()
        r_np = r.cpu().numpy()
        log(f"CPU baseline on {args.cpu_sample} walkers ({procs} processes)")
        with pool.OraclePool(params, cfg, "float32", procs) as p32:
            p32.warm()
            t0 = time.perf_counter()
            p32(r_np[:args.cpu_sample])
            dt_cpu = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": args.cpu_sample / dt_cpu, "unit": "walkers/s",
                                  "cores": procs, "kind": "port",
                                  "sample": f"{args.cpu_sample} walkers of the same batch, oracle in FP32 (torch CPU, "
                                            f"sparse forward Laplacian), {procs} single-threaded processes, "
                                            f"{dt_cpu:.1f} s (process start-up excluded)"}
        if args.bias_sample > 0:
            n = min(args.bias_sample, B)
            log(f"FP64 oracle on {n} walkers")
            ref, _, _ = pool.local_energies(params, cfg, r_np[:n], procs=procs)
            result["energy_error_vs_fp64_mHa"] = {
                "walkers": n, "mixed": err_stats(e_m[:n] - ref), "fp32": err_stats(e_f[:n] - ref),
                "mixed_fast": err_stats(e_x[:n] - ref),
                "mixed_fp64_laplacian": err_stats(e_d[:n] - ref),
                "gate": "|bias| <= 1 mHa; bias_se = standard error of the bias",
            }
            for k in ("mixed", "fp32", "mixed_fast", "mixed_fp64_laplacian"):
                e = {"mixed": e_m, "fp32": e_f, "mixed_fast": e_x, "mixed_fp64_laplacian": e_d}[k][:n]
                result["energy_error_vs_fp64_mHa"][k]["bias_se"] = float(np.std((e - ref) * 1e3, ddof=1) / np.sqrt(n))
        legs = [x for x in args.config_legs.split(",") if x]
        if legs:
            result["configs"] = {mol.name or args.workload: {
                "walkers": B, "mixed_walkers_per_s": wps_m, "fp32_walkers_per_s": wps_f,
                "energy_error_vs_fp64_mHa": result.get("energy_error_vs_fp64_mHa")}}
            for name in legs:
                result["configs"][name] = config_leg(name, procs, args)
    print(json.dumps(result))


def config_leg(name, procs, args):
    """Another benchmark configuration at its own size: walkers equilibrated by 5 sweeps of the
    GPU sampler, mixed and FP32 local energy of the whole batch (timed), and the error of both
    against the FP64 oracle on a sample of 256 walkers."""
    import torch

    from oracle import pool
    from sparse_wf_amd import params as P
    from sparse_wf_amd import systems
    from sparse_wf_amd.config import FireConfig
    from sparse_wf_amd.wave_function import FireWaveFunction

    mol, over, n_walk, sigma = systems.workload(name)
    cfg = FireConfig(mol=mol, **over)
    log(f"config leg {name}: N={cfg.n_el}, {n_walk} walkers")
    r0 = systems.init_electrons(mol, n_walk, stddev=sigma, seed=0)
    wf = FireWaveFunction(cfg, device=0)
    params = wf.init(0, r0[0])
    P.exercise_jastrow(params)
    r, _ = wf.mcmc_steps_low_rank(params, r0, 5 * cfg.n_el, width=args.width, seed=1, precision="mixed")
    out = {"workload": f"{mol.name} N={cfg.n_el} walkers={n_walk} cutoff={cfg.cutoff} F={cfg.feature_dim} "
                       f"D={cfg.n_determinants}", "walkers": n_walk}
    e = {}
    for prec in ("mixed", "fp32"):
        wf.local_energy_terms(params, r, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            e[prec] = wf.local_energy_terms(params, r, precision=prec)["e_loc"]
        torch.cuda.synchronize()
        out[f"{prec}_walkers_per_s"] = 3 * n_walk / (time.perf_counter() - t0)
        e[prec] = e[prec].cpu().numpy()
    n = min(256, n_walk)
    ref, _, _ = pool.local_energies(params, cfg, r[:n].cpu().numpy(), procs=procs)
    out["energy_error_vs_fp64_mHa"] = {"walkers": n}
    for prec in ("mixed", "fp32"):
        d = e[prec][:n] - ref
        st = err_stats(d)
        st["bias_se"] = float(np.std(d * 1e3, ddof=1) / np.sqrt(n))
        out["energy_error_vs_fp64_mHa"][prec] = st
    return out


if __name__ == "__main__":
    main()
