"""Benchmark: transmission points per second of the nested-dissection engine on one MI355X.

Workload (the equal-size batch): BHZ ribbon 200 x 64 sites (51,200 orbitals, two spin blocks),
polarized magnetic impurities (concentration 0.2, exchange J = 0.4, opposite g), configurations
seeded 1..8 as `caroli run` does; 33 energies in [-0.16, 0.16]; 264 points per step. The lead
self-energies are computed once per energy on the host (FP64) and shared by the configurations
(crates/caroli-cli/src/run.rs:113-127); they are inputs, resident in HBM with the Hamiltonians.

A step is one batched launch over all 264 points. The JSON line carries the per-kernel roofline
(HIP events on the engine's stream during the timed steps) and the CPU baseline: the oracle (FP64
C++ restatement of the reference path) on the host cores for a bounded sample of the same points.
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6e12  # MI355X dense FP64 matrix
PEAK_FP32_MFMA = 157.3e12  # MI355X dense FP32 matrix
PEAK_HBM = 8.0e12


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="double", choices=["double", "single", "mixed"])
    ap.add_argument("--configs", type=int, default=8)
    ap.add_argument("--energies", type=int, default=33)
    ap.add_argument("--width", type=int, default=64)
    ap.add_argument("--length", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU time of the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the engine has no multi-GPU path; "
              "independent points would shard by energy across GPUs.", file=sys.stderr)
        sys.exit(2)

    import caroli_amd as ca

    if not ca.device_available():
        raise SystemExit("no HIP device: the engine has no CPU path")
    prec = {"double": ca.Precision.Double, "single": ca.Precision.Single, "mixed": ca.Precision.Mixed}[args.precision]
    threads = max(1, min(os.cpu_count() or 1, 16))

    t0 = time.time()
    systems = [ca.NestedTwoTerminal.preset(ca.Model.BHZ, [args.length, args.width, 0], disorder=2, strength=0.4,
                                           concentration=0.2, seed=1 + c) for c in range(args.configs)]
    build_s = time.time() - t0
    energies = np.linspace(-0.16, 0.16, args.energies)
    t0 = time.time()
    with ThreadPoolExecutor(threads) as ex:
        sigmas = list(ex.map(systems[0].self_energies, energies))
    sigma_s = time.time() - t0
    pts_sys = [s for s in systems for _ in energies]
    pts_sig = [g for _ in systems for g in sigmas]
    npts = len(pts_sys)
    batch = ca.ResidentBatch(pts_sys, pts_sig, prec)
    for _ in range(args.warmup):
        batch.run()
    ca.kernel_timing(True)
    times, stats = [], ca.capi.BatchStats()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        t, st = batch.run(stats)
        times.append(time.perf_counter() - t0)
    ks = ca.kernel_stats()
    ca.kernel_timing(False)
    step = float(np.mean(times))
    value = npts / step
    flops_pt, bytes_pt = systems[0].work()

    # Dominant kernel family by time, and its roofline.
    # The bound is the roof the kernel's algorithmic intensity meets first: flops/byte against
    # the ridge point peak_flops / peak_bandwidth.
    dom = max(ks.items(), key=lambda kv: kv[1][1])
    name, (launches, secs, fl, by) = dom
    peak = PEAK_FP32_MFMA if prec != ca.Precision.Double else PEAK_FP64_MFMA
    bound = "mfma" if fl > 0 and fl / max(by, 1) >= peak / PEAK_HBM else "hbm"
    if bound == "mfma":
        achieved = fl / secs
        roof = {"bound": "mfma", "kernel": name, "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None, "launches": launches,
                "avg_launch_ms": 1e3 * secs / max(launches, 1)}
    else:
        achieved = by / secs
        roof = {"bound": "hbm", "kernel": name, "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": achieved / PEAK_HBM, "traffic": None, "launches": launches,
                "avg_launch_ms": 1e3 * secs / max(launches, 1)}
    kernels = {k: {"launches": v[0], "ms": 1e3 * v[1] / max(args.steps, 1),
                   "tflops": (v[2] / v[1] / 1e12) if v[1] > 0 else 0.0,
                   "gbs": (v[3] / v[1] / 1e9) if v[1] > 0 else 0.0} for k, v in ks.items()}

    cpu = None
    max_dt = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from _oracle import OracleSystem, transmissions_parallel

        osys = [OracleSystem(s.arrays()) for s in systems]
        flats = [OracleSystem.pack([g.block(b)[:2] for b in range(systems[0].block_count())]) for g in sigmas]
        # Single-thread latency of one point, then a bounded sample on all threads.
        t0 = time.perf_counter()
        one, _ = transmissions_parallel([osys[0]], [energies[0]], [flats[0]], 1)
        lat = time.perf_counter() - t0
        per_thread = max(1, int(args.cpu_seconds / max(lat, 1e-6)))
        sample = min(npts, threads * per_thread)
        idx = np.linspace(0, npts - 1, sample).astype(int)
        t0 = time.perf_counter()
        tc, failed = transmissions_parallel([osys[i // len(energies)] for i in idx], [energies[i % len(energies)] for i in idx],
                                            [flats[i % len(energies)] for i in idx], threads)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": sample / cpu_s, "unit": "points/s", "cores": threads, "kind": "port",
               "sample": f"{sample} of the {npts} points (evenly spaced), FP64 oracle, one point per thread",
               "single_thread_s_per_point": lat}
        max_dt = float(np.max(np.abs(tc - t[idx])))

    line = {
        "metric": "T(E) points/sec at |dT|<=1e-8 vs FP64 CPU",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"double": "complex128", "single": "complex64", "mixed": "complex64+f64 refinement"}[args.precision],
        "data": "synthetic (seeded disorder, host-built Hamiltonians)",
        "config": {"workload": f"BHZ {args.length}x{args.width} magnetic c=0.2 J=0.4 opposite-g, "
                               f"{args.energies} E x {args.configs} configs", "points_per_step": npts,
                   "precision": args.precision, "orbitals": systems[0].sizes()[0], "blocks": systems[0].block_count()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "max_abs_dT_vs_cpu": max_dt,
        "status_ok": int((st == 0).sum()),
        "delayed_points": int(stats.delayed_points),
        "refinement_iters": int(stats.refinement_iters),
        "max_refine_correction": float(stats.max_refine_resid),
        "device_seconds_per_step": stats.seconds,
        "algorithmic_flops_per_point": flops_pt,
        "algorithmic_bytes_per_point": bytes_pt,
        "whole_step_tflops": flops_pt * npts / step / 1e12,
        "sigma_host_seconds": sigma_s,
        "build_seconds": build_s,
        "kernels": kernels,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
