"""Benchmark: transmission points per second of the nested-dissection engine on one MI355X.

Default workload (the equal-size batch, `--config c2`): BHZ ribbon 200 x 64 sites (51,200 orbitals,
two spin blocks), polarized magnetic impurities (concentration 0.2, exchange J = 0.4, opposite g),
configurations seeded 1..8 as `caroli run` does; 33 energies in [-0.16, 0.16]; 264 points per
step, complex128. The lead self-energies are computed once per energy on the host (FP64) and shared
by the configurations (crates/caroli-cli/src/run.rs:113-127); they are inputs, resident in HBM with
the Hamiltonians before the timed region.

Other workloads: `--config c1` the clean BHZ ribbon 16 x 10 at 33 energies (complex128); `--config
c3` zigzag graphene 512 x 64 two-site cells, Anderson w = 1, 256
energies x 32 configurations (8,192 points; refined mixed precision by default); `--config c4`
the adjoint of BHZ 100 x 32 (Anderson w = 40/72.9, E = 0.01; a step is one backward sweep giving
dT/d(eps) for all 12,800 orbitals); `--config c5` the uneven cubic batch (W = 16.5, lengths 2L,
cross-sections L drawn from {16, 24, 32, 40, 48} with xoshiro256** seeded 5, 4 energies in
[-1, 1]) on the first `--systems` systems whose L is at most `--max-l`.

Before the warmups the engine's autotuner times the kernel choices per front class on the batch's
own points (`--autotune`, on by default except for mixed precision) and keeps the fastest; the JSON
line reports the choice under `tuning`.

A step is one batched launch over all points; the engine runs up to four launch units (spin blocks,
halves of a block's points, systems of other sizes) at once on four streams, and replays each unit's
sweep from a HIP graph after the warmups. The JSON line carries the roofline of the dominant kernel
family: HIP events on the engine's streams during `--roofline-steps` steps that follow the timed
ones and run the units one after the other (with several sweeps in flight, a launch's duration would
include the other streams' work); the HBM traffic per launch comes from the rocprofv3 counter
summary in profiles/ when one matches. Unless `--no-compare`, the same points also run in the other
precision (complex128 against refined mixed). The CPU baseline is the oracle (FP64 C++ restatement
of the reference path) on the host cores for a bounded sample of the same points.
"""
import argparse
import glob
import json
import os
import re
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6e12  # MI355X dense FP64 matrix
PEAK_FP32_MFMA = 157.3e12  # MI355X dense FP32 matrix
PEAK_HBM = 8.0e12


def workload_tag(args):
    if args.config == "c5":
        return f"cubic_uneven_l{args.max_l}_s{args.systems}"
    return {"c1": "bhz16x10", "c2": "bhz200x64", "c3": "graphene512x64", "c4": "bhz100x32_adjoint"}[args.config]


def traffic_from_profiles(family, tag, precision):
    """HBM bytes per launch of `family` on this workload from the committed counter summaries.

    The summaries (profiles/*.md) carry lines `traffic_bytes_per_launch: <family> <workload>
    <precision> <bytes>`, from two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) scaled by the
    same counters read on a plain device copy of known size."""
    pat = re.compile(r"^traffic_bytes_per_launch:\s+(\S+)\s+(\S+)\s+(\S+)\s+([0-9.eE+]+)\s*$")
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                m = pat.match(line.strip())
                if m and m.group(1) == family and m.group(2) == tag and m.group(3) == precision:
                    return float(m.group(4)), os.path.relpath(path, ROOT)
    return None, None


def roofline_of(ks, prec_name, tag):
    """Roofline object of the dominant kernel family (by time), from the engine's HIP-event timers.

    The bound is the roof the family's algorithmic intensity meets first: flops per byte against
    the ridge point peak_flops / peak_bandwidth."""
    if not ks or max(v[1] for v in ks.values()) <= 0:
        return None  # no timed launches (--roofline-steps 0)
    name, (launches, secs, fl, by) = max(ks.items(), key=lambda kv: kv[1][1])
    peak = PEAK_FP64_MFMA if prec_name == "double" else PEAK_FP32_MFMA
    bound = "mfma" if fl > 0 and fl / max(by, 1) >= peak / PEAK_HBM else "hbm"
    traffic, source = traffic_from_profiles(name, tag, prec_name)
    if bound == "mfma":
        achieved = fl / secs
        roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": achieved / peak}
    else:
        achieved = by / secs
        roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": achieved / PEAK_HBM}
    roof.update({"traffic": traffic, "traffic_source": source, "kernel": name, "launches": launches,
                 "avg_launch_ms": 1e3 * secs / max(launches, 1),
                 "algorithmic_bytes_per_launch": by / max(launches, 1),
                 "algorithmic_flops_per_launch": fl / max(launches, 1)})
    return roof


def host_threads():
    """Threads of the CPU baseline: the cores of this process's affinity mask, up to 16 (a one-GPU
    box's CPU share, whatever nproc reports there)."""
    try:
        mask = len(os.sched_getaffinity(0))
    except AttributeError:
        mask = os.cpu_count() or 1
    return max(1, min(mask, 16))


def host_cores():
    """(nproc, cores in the affinity mask) of this host."""
    try:
        mask = len(os.sched_getaffinity(0))
    except AttributeError:
        mask = None
    return os.cpu_count(), mask


def progress(*a):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print("[bench]", *a, file=sys.stderr, flush=True)


def kernels_of(ks, steps):
    return {k: {"launches": v[0], "ms": 1e3 * v[1] / max(steps, 1),
                "tflops": (v[2] / v[1] / 1e12) if v[1] > 0 else 0.0,
                "gbs": (v[3] / v[1] / 1e9) if v[1] > 0 else 0.0} for k, v in ks.items()}


def build_points(args, ca, threads):
    """Systems, per-point (system, self-energies, energy) lists and the description of a T workload."""
    if args.config == "c1":
        systems = [ca.NestedTwoTerminal.preset(ca.Model.BHZ, [16, 10, 0])]
        energies = np.linspace(-0.16, 0.16, args.energies)
        desc = f"clean BHZ 16x10, {args.energies} E"
        groups = [(systems, energies)]
    elif args.config == "c2":
        systems = [ca.NestedTwoTerminal.preset(ca.Model.BHZ, [args.length, args.width, 0], disorder=2, strength=0.4,
                                               concentration=0.2, seed=1 + c) for c in range(args.configs)]
        energies = np.linspace(-0.16, 0.16, args.energies)
        desc = (f"BHZ {args.length}x{args.width} magnetic c=0.2 J=0.4 opposite-g, "
                f"{args.energies} E x {args.configs} configs")
        groups = [(systems, energies)]
    elif args.config == "c3":
        systems = [ca.NestedTwoTerminal.preset(ca.Model.Graphene, [512, 64, 0], disorder=1, strength=1.0, seed=1 + c)
                   for c in range(args.configs)]
        energies = np.linspace(-0.5, 0.5, args.energies)
        desc = f"zigzag graphene 512x64 cells (128 atoms wide), Anderson w=1, {args.energies} E x {args.configs} configs"
        groups = [(systems, energies)]
    else:  # c5
        import ctypes as C
        u = np.zeros(512)
        ca.capi.cg_uniforms(0, 5, None, 512, u.ctypes.data_as(C.POINTER(C.c_double)))
        sizes = [[16, 24, 32, 40, 48][int(5 * x)] for x in u]
        chosen = [(s, L) for s, L in enumerate(sizes) if L <= args.max_l][:args.systems]
        energies = np.linspace(-1.0, 1.0, 4)
        by_l = {}
        for s, L in chosen:
            by_l.setdefault(L, []).append(ca.NestedTwoTerminal.preset(ca.Model.Cubic, [2 * L, L, L], disorder=1,
                                                                      strength=16.5, seed=1000 + s))
            progress(f"system {s} (L = {L}) built")
        groups = [(by_l[L], energies) for L in sorted(by_l)]
        desc = (f"cubic Anderson W=16.5, 2L x L x L, systems {[s for s, _ in chosen]} of the 512 (L <= {args.max_l}: "
                f"L = {[L for _, L in chosen]}), 4 E")
    pts_sys, pts_sig, pts_e = [], [], []
    t0 = time.time()
    for systems, energies in groups:
        with ThreadPoolExecutor(threads) as ex:
            sigmas = list(ex.map(systems[0].self_energies, energies))
        progress(f"self-energies of {len(energies)} energies done ({time.time() - t0:.1f} s)")
        for s in systems:
            for e, g in zip(energies, sigmas):
                pts_sys.append(s)
                pts_sig.append(g)
                pts_e.append(e)
    return pts_sys, pts_sig, pts_e, desc, time.time() - t0


def cpu_baseline(pts_sys, pts_sig, pts_e, t_gpu, args, threads):
    """The oracle on the host cores, one point per thread, on a bounded sample of the points."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import OracleSystem, transmissions_parallel

    npts = len(pts_sys)
    cache = {}

    def osys(s):
        if id(s) not in cache:
            cache[id(s)] = OracleSystem(s.arrays())
        return cache[id(s)]

    def flat(p):
        s, g = pts_sys[p], pts_sig[p]
        return OracleSystem.pack([g.block(b)[:2] for b in range(s.block_count())])

    # Points differ in cost when the systems differ in size (the cubic batch): the single-thread
    # probe takes the cheapest point, and the sample the points whose estimated time fits the
    # budget, evenly spaced; the rate of the whole workload is extrapolated from the sample's
    # measured flop rate.
    flops = np.array([s.work()[0] for s in pts_sys])
    p0 = int(np.argmin(flops))
    progress(f"cpu baseline: single-thread probe on point {p0} ({flops[p0]:.2e} flops)")
    t0 = time.perf_counter()
    transmissions_parallel([osys(pts_sys[p0])], [pts_e[p0]], [flat(p0)], 1)
    lat = time.perf_counter() - t0
    rate = flops[p0] / max(lat, 1e-9)  # flops per second of one thread
    fits = np.nonzero(flops / rate <= args.cpu_seconds)[0]
    if fits.size == 0:
        fits = np.array([p0])
    per_thread = max(1, int(args.cpu_seconds / max(float(np.mean(flops[fits])) / rate, 1e-6)))
    sample = min(fits.size, threads * per_thread)
    idx = fits[np.linspace(0, fits.size - 1, sample).astype(int)]
    progress(f"cpu baseline: {sample} points on {threads} threads")
    t0 = time.perf_counter()
    tc, _ = transmissions_parallel([osys(pts_sys[i]) for i in idx], [pts_e[i] for i in idx], [flat(i) for i in idx],
                                   threads)
    cpu_s = time.perf_counter() - t0
    progress(f"cpu baseline: {cpu_s:.1f} s")
    sample_rate = float(np.sum(flops[idx])) / cpu_s  # flops per second on all threads
    nproc, mask = host_cores()
    cpu = {"value": sample / cpu_s, "unit": "points/s", "cores": threads, "kind": "port",
           "host_nproc": nproc, "host_affinity_cores": mask,
           "sample": (f"{sample} of the {npts} points ({'evenly spaced' if fits.size == npts else 'evenly spaced among the ' + str(fits.size) + ' whose estimated single-thread time fits ' + str(args.cpu_seconds) + ' s'}), "
                      "FP64 oracle, one point per thread"),
           "single_thread_s_per_point": lat,
           "workload_points_per_s": sample_rate / float(np.mean(flops))}
    d = np.abs(tc - t_gpu[idx])
    return cpu, float(np.max(d)), float(np.max(d / (1.0 + np.abs(tc))))


def run_transmission(args, ca):
    prec = {"double": ca.Precision.Double, "single": ca.Precision.Single, "mixed": ca.Precision.Mixed}[args.precision]
    threads = host_threads()
    t0 = time.time()
    pts_sys, pts_sig, pts_e, desc, sigma_s = build_points(args, ca, threads)
    build_s = time.time() - t0 - sigma_s
    npts = len(pts_sys)
    progress(f"{npts} points: {desc}; systems built in {build_s:.1f} s, self-energies {sigma_s:.1f} s")
    batch = ca.ResidentBatch(pts_sys, pts_sig, prec)
    progress("batch resident")
    tuning = None
    if args.autotune == "on" or (args.autotune == "auto" and args.precision != "mixed"):
        # Kernel choices timed on these points (eager sweeps, before the warmups and outside the
        # timed steps); the fastest stays in force for everything after it in this process.
        t0 = time.perf_counter()
        choice, sec = batch.autotune()
        tuning = {"choice": choice, "seconds_per_sweep": sec, "tuning_seconds": time.perf_counter() - t0}
        progress(f"autotune: {choice}, {sec:.3f} s per eager sweep ({tuning['tuning_seconds']:.1f} s)")
    for w in range(args.warmup):
        t0 = time.perf_counter()
        batch.run()
        progress(f"warmup {w + 1}: {time.perf_counter() - t0:.3f} s")
    times, stats = [], ca.capi.BatchStats()
    for k in range(args.steps):
        t0 = time.perf_counter()
        t, st = batch.run(stats)
        times.append(time.perf_counter() - t0)
        progress(f"step {k + 1}: {times[-1]:.3f} s")
    step = float(np.mean(times))
    # Kernel roofline: the same launches with the HIP-event timers on, which run the launch units
    # (the two spin blocks) one after the other so that each launch's duration is its own.
    ca.kernel_timing(True)
    for _ in range(args.roofline_steps):
        batch.run()
    ks = ca.kernel_stats()
    ca.kernel_timing(False)
    progress("roofline steps done")
    other = None
    if args.compare and args.precision in ("double", "mixed"):
        other = compare_precision(ca, pts_sys, pts_sig, t, npts / step, args)
    per_point = plain_per_point(ca, pts_sys, pts_sig, args) if args.per_point > 0 else None
    if per_point is not None:
        # Batched rates over the plain complex128 solver called one point at a time (resident
        # one-point batches replayed from their graphs, the faster of its two forms).
        mixed_rate = npts / step if args.precision == "mixed" else (other or {}).get("value")
        plain_rate = npts / step if args.precision == "double" else (other or {}).get("value")
        per_point["batched_plain_over_per_point"] = plain_rate / per_point["resident_points_per_s"] \
            if plain_rate else None
        per_point["batched_mixed_over_per_point"] = mixed_rate / per_point["resident_points_per_s"] \
            if mixed_rate else None
    fixture = fixture_check(args, pts_e, t)
    cpu, max_dt, max_rel = (None, None, None) if args.no_cpu else cpu_baseline(pts_sys, pts_sig, pts_e, t, args, threads)
    if cpu is not None and other is not None:
        # the other precision against the same CPU sample, through its difference to this run
        other["max_abs_dT_vs_cpu_bound"] = max_dt + other["max_abs_dT_vs_this"]
    flops_pt = np.mean([s.work()[0] for s in pts_sys])
    bytes_pt = np.mean([s.work()[1] for s in pts_sys])
    return {
        "metric": "T(E) points/sec at |dT|<=1e-8 vs FP64 CPU",
        "value": npts / step,
        "unit": "points/s",
        "ms_per_step": 1e3 * step,
        "higher_is_better": True,
        "dtype": {"double": "complex128", "single": "complex64", "mixed": "complex64+f64 refinement"}[args.precision],
        "config": {"workload": desc, "points_per_step": npts, "precision": args.precision,
                   "orbitals": int(max(s.sizes()[0] for s in pts_sys))},
        "roofline": roofline_of(ks, args.precision, workload_tag(args)),
        "cpu_baseline": cpu,
        "max_abs_dT_vs_cpu": max_dt,
        "max_dT_over_1_plus_T_vs_cpu": max_rel,
        "status_ok": int((st == 0).sum()),
        "delayed_points": int(stats.delayed_points),
        "recomputed_points": int(stats.recomputed_points),
        "refinement_iters": int(stats.refinement_iters),
        "max_refine_correction": float(stats.max_refine_resid),
        "device_seconds_per_step": stats.seconds,
        "algorithmic_flops_per_point": flops_pt,
        "algorithmic_bytes_per_point": bytes_pt,
        "whole_step_tflops": flops_pt * npts / step / 1e12,
        "whole_step_algorithmic_gbs": bytes_pt * npts / step / 1e9,
        "sigma_host_seconds": sigma_s,
        "build_seconds": build_s,
        "tuning": tuning if tuning is not None else {"choice": ca.get_tuning(), "autotuned": False},
        "kernels": kernels_of(ks, args.roofline_steps),
        "roofline_steps": args.roofline_steps,
        ("mixed" if args.precision == "double" else "plain"): other,
        "plain_per_point": per_point,
        "long_double_fixture": fixture,
    }


def plain_per_point(ca, pts_sys, pts_sig, args):
    """The plain complex128 solver one point at a time, on `--per-point` points spread over the
    batch: (1) through the per-point API, transmission_with (nested.rs:682), as a caller looping
    over points would (inputs uploaded per call); (2) as a resident one-point batch replayed from
    its HIP graph (the per-point latency floor). Rates in points/s."""
    npts = len(pts_sys)
    idx = np.linspace(0, npts - 1, min(args.per_point, npts)).astype(int)
    for i in idx[:2]:
        pts_sys[i].transmission_with(pts_sig[i])  # warm
    t0 = time.perf_counter()
    for i in idx:
        pts_sys[i].transmission_with(pts_sig[i])
    api = len(idx) / (time.perf_counter() - t0)
    batches = [ca.ResidentBatch([pts_sys[i]], [pts_sig[i]], ca.Precision.Double) for i in idx]
    for b in batches:
        b.run()
        b.run()  # eager, then captured
    t0 = time.perf_counter()
    for b in batches:
        b.run()
    resident = len(idx) / (time.perf_counter() - t0)
    progress(f"plain per point: {api:.1f} points/s through transmission_with, {resident:.1f} resident")
    return {"points": int(len(idx)), "api_points_per_s": api, "resident_points_per_s": resident,
            "unit": "points/s", "precision": "double"}


def fixture_check(args, pts_e, t):
    """|T - T_long| against the long-double oracle values committed for the full equal-size batch
    (tests/golden/bhz200x64_magnetic.json), when this run is that workload."""
    path = os.path.join(ROOT, "tests", "golden", "bhz200x64_magnetic.json")
    if args.config != "c2" or (args.length, args.width, args.configs, args.energies) != (200, 64, 8, 33):
        return None
    if not os.path.exists(path) or args.precision != "double":
        return None
    pts = np.array(json.load(open(path))["points"])
    if not np.array_equal(pts[:, 1], np.asarray(pts_e)):
        return None
    err = np.abs(t - pts[:, 3])
    # Two double-precision evaluations of the same elimination (plain and with contracted
    # multiply-adds) sample the rounding error of double precision on each point; the per-point
    # bar is 1e-10, or four times that error where it is larger (tests/test_gpu_headline.py).
    eps = np.maximum(np.abs(pts[:, 2] - pts[:, 3]), np.abs(pts[:, 4] - pts[:, 3]))
    bar = np.maximum(1e-10, 4.0 * eps)
    return {"max_abs_dT_vs_long_double": float(err.max()), "median": float(np.median(err)),
            "points_above_1e-10": int((err > 1e-10).sum()),
            "double_evaluations_max_error": float(eps.max()),
            "points_with_bar_above_1e-10": int((bar > 1e-10).sum()),
            "max_err_over_bar": float(np.max(err / bar)),
            "within_bar": bool(np.all(err <= bar))}


def compare_precision(ca, pts_sys, pts_sig, t_this, rate_this, args):
    """The same resident points in the other precision (complex128 <-> refined mixed): its rate, the
    ratio mixed / plain, and its largest |dT| against this run."""
    name = "mixed" if args.precision == "double" else "double"
    prec = ca.Precision.Mixed if name == "mixed" else ca.Precision.Double
    batch = ca.ResidentBatch(pts_sys, pts_sig, prec)
    for _ in range(args.warmup):
        batch.run()
    times, stats = [], ca.capi.BatchStats()
    for k in range(args.steps):
        t0 = time.perf_counter()
        t, st = batch.run(stats)
        times.append(time.perf_counter() - t0)
        progress(f"{name} step {k + 1}: {times[-1]:.3f} s")
    rate = len(pts_sys) / float(np.mean(times))
    mixed_rate, plain_rate = (rate, rate_this) if name == "mixed" else (rate_this, rate)
    return {"precision": name, "value": rate, "unit": "points/s", "ms_per_step": 1e3 * float(np.mean(times)),
            "mixed_over_plain": mixed_rate / plain_rate,
            "max_abs_dT_vs_this": float(np.max(np.abs(t - t_this))), "status_ok": int((st == 0).sum()),
            "recomputed_points": int(stats.recomputed_points), "delayed_points": int(stats.delayed_points),
            "refinement_iters": int(stats.refinement_iters),
            "max_refine_correction": float(stats.max_refine_resid)}


def run_adjoint(args, ca):
    """One backward sweep (T and dT/d(eps) of every orbital) per step, BHZ 100 x 32."""
    s = ca.NestedTwoTerminal.preset(ca.Model.BHZ, [100, 32, 0], disorder=1, strength=40 / 72.9, seed=1)
    sig = s.self_energies(0.01)
    for _ in range(args.warmup):
        s.backward(sig)
    ca.kernel_timing(True)
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        t, grad = s.backward(sig)
        times.append(time.perf_counter() - t0)
    ks = ca.kernel_stats()
    ca.kernel_timing(False)
    step = float(np.mean(times))
    cpu, err = None, None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from helpers import oracle_gradient

        t0 = time.perf_counter()
        ref = oracle_gradient(s, 0.01)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": 1.0 / cpu_s, "unit": "gradients/s", "cores": 1, "kind": "port",
               "sample": "one FP64 gradient from the oracle's G columns (G E_R and G^T E_L, all orbitals)"}
        err = float(np.abs(grad - ref).max() / np.abs(ref).max())
    # The batched adjoint: 16 disorder configurations of the same structure at E = 0.01 in one
    # backward sweep (cg_backward_batch), gradients per second.
    batch = [ca.NestedTwoTerminal.preset(ca.Model.BHZ, [100, 32, 0], disorder=1, strength=40 / 72.9, seed=1 + c)
             for c in range(16)]
    sigs = [sig] * len(batch)  # clean leads: the same self-energies serve every configuration
    ca.backward_batch(batch, sigs)
    t0 = time.perf_counter()
    reps = max(1, args.steps)
    for _ in range(reps):
        tb, gb, sb = ca.backward_batch(batch, sigs)
    batched = {"points": len(batch), "value": reps * len(batch) / (time.perf_counter() - t0), "unit": "gradients/s",
               "status_ok": int((sb == 0).sum()),
               "max_abs_grad_vs_single": float(np.abs(gb[0] - grad).max())}
    progress(f"batched adjoint: {batched['value']:.1f} gradients/s over {len(batch)} configurations")
    return {
        "metric": "adjoint sweeps/sec (T and dT/d(eps) for all orbitals)",
        "batched": batched,
        "value": 1.0 / step,
        "unit": "gradients/s",
        "ms_per_step": 1e3 * step,
        "higher_is_better": True,
        "dtype": "complex128",
        "config": {"workload": "BHZ 100x32 Anderson w=40/72.9, E=0.01, adjoint over 12,800 orbitals", "T": t},
        "roofline": roofline_of(ks, "double", workload_tag(args)),
        "cpu_baseline": cpu,
        "max_rel_grad_vs_cpu": err,
        "kernels": kernels_of(ks, args.steps),
    }


def copy_calibration():
    """A plain device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    x = torch.empty(1 << 27, dtype=torch.complex128, device="cuda")  # 2 GiB
    x.fill_(1.0)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": x.numel() * 16}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2,
                    help="untimed steps (the first runs eagerly, the second captures the sweeps into HIP graphs)")
    ap.add_argument("--roofline-steps", type=int, default=1,
                    help="steps after the timed ones with the per-kernel timers on (units one at a time)")
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--precision", default=None, choices=["double", "single", "mixed"])
    ap.add_argument("--configs", type=int, default=None, help="disorder configurations (c2: 8, c3: 32)")
    ap.add_argument("--energies", type=int, default=None, help="energies (c2: 33, c3: 256)")
    ap.add_argument("--width", type=int, default=64)
    ap.add_argument("--length", type=int, default=200)
    ap.add_argument("--systems", type=int, default=6, help="c5: systems taken from the 512")
    ap.add_argument("--max-l", type=int, default=40, help="c5: largest cross-section taken")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU time of the baseline sample")
    ap.add_argument("--autotune", default="auto", choices=["auto", "on", "off"],
                    help="time the kernel choices on the batch first (auto: complex128 and complex64 runs; "
                         "a mixed run re-runs its refinement per candidate)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--per-point", type=int, default=16,
                    help="points timed one at a time by the plain complex128 solver (0: skip)")
    ap.add_argument("--no-compare", dest="compare", action="store_false",
                    help="skip the run of the same points in the other precision (mixed vs plain)")
    ap.add_argument("--copy-calibration", action="store_true", help="run a 2 GiB device copy and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the engine has no multi-GPU path; "
              "independent points would shard by energy across GPUs.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    defaults = {"c1": ("double", 1, 33), "c2": ("double", 8, 33), "c3": ("mixed", 32, 256), "c4": ("double", 1, 1),
                "c5": ("double", 1, 4)}
    p, c, e = defaults[args.config]
    args.precision = args.precision or p
    args.configs = args.configs or c
    args.energies = args.energies or e

    import caroli_amd as ca

    if not ca.device_available():
        raise SystemExit("no HIP device: the engine has no CPU path")
    line = run_adjoint(args, ca) if args.config == "c4" else run_transmission(args, ca)
    line.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "scaling": "weak", "vs_baseline": None,
                 "data": "synthetic (seeded disorder, host-built Hamiltonians)"})
    invalid = validity(line, args)
    if invalid:
        line["invalid"] = invalid
    keys = ["metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
            "vs_baseline", "dtype", "data", "config"]
    print(json.dumps({**{k: line[k] for k in keys}, **{k: v for k, v in line.items() if k not in keys}}), flush=True)
    if invalid:
        progress("INVALID: " + invalid)
        sys.exit(1)


def validity(line, args):
    """Why the value may not be published, or None: the metric counts points at |dT| <= 1e-8
    against the FP64 CPU oracle, so every point must have succeeded and the sampled points must
    meet that bound; on the equal-size batch in complex128 every point must also meet the parity
    bar against the long-double oracle."""
    if args.config == "c4":
        err = line.get("max_rel_grad_vs_cpu")
        return None if err is None or err <= 1e-6 else f"gradient off by {err:.1e} relative"
    npts = line["config"]["points_per_step"]
    if line.get("status_ok") != npts:
        return f"{npts - line.get('status_ok', 0)} of {npts} points failed"
    d = line.get("max_abs_dT_vs_cpu")
    if d is not None and not d <= 1e-8:
        return f"max |dT| against the CPU oracle {d:.1e} above 1e-8"
    fx = line.get("long_double_fixture")
    if fx is not None and not fx["max_abs_dT_vs_long_double"] <= 1e-8:
        return f"max |dT| against the long-double oracle {fx['max_abs_dT_vs_long_double']:.1e} above 1e-8"
    if fx is not None and not fx["within_bar"]:
        return (f"complex128 parity bar missed against the long-double oracle "
                f"(max |dT| / bar {fx['max_err_over_bar']:.2f})")
    plain = line.get("plain")  # a mixed headline: its difference to complex128 on every point
    if plain is not None and not plain["max_abs_dT_vs_this"] <= 1e-8:
        return f"max |dT| against complex128 {plain['max_abs_dT_vs_this']:.1e} above 1e-8"
    return None


if __name__ == "__main__":
    main()
