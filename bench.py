"""Benchmark of the stencil path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

A step is one fused sweep (out += S(in), in_next = in + 1) over the whole
grid, the per-iteration work of Cxx11/stencil.cc:183-190.  Default workload:
FP64 star R=2 on a 30000 x 30000 grid with the reference's analytic init,
100 timed iterations after 1 untimed warm-up sweep (the reference times
iterations 1..K after sweep 0, Cxx11/stencil.cc:185).  The run is validated
with the reference's closed form, norm = 2*(sweeps) within 1e-8.

Prints ONE JSON line.  `value` is the algorithmic HBM bandwidth of the
sweep (GB/s): B = 2*s*n^2 + 2*s*(n-2R)^2 bytes per sweep (read in, write
in_next on every point; read and write out on the interior), inputs resident
in HBM, timed with HIP events on the stream the sweeps run on.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "star-2 stencil GB/s & % HBM peak at n=30000; served solves/s + p99 latency"
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth

WORKLOADS = {
    # name: (pattern, radius, dtype, n)
    "star2_f64_n30000": (0, 2, 0, 30000),
    "grid2_f64_n30000": (1, 2, 0, 30000),
    "star2_f32_n30000": (0, 2, 1, 30000),
    "star2_bf16_n30000": (0, 2, 2, 30000),
    # FP64 grid stencils bound by their tap chains (radius/precision sweep sizes)
    "grid3_f64_n16384": (1, 3, 0, 16384),
    "grid5_f64_n16384": (1, 5, 0, 16384),
    # a star stencil swept from the register ring
    "star8_f64_n16384": (0, 8, 0, 16384),
    # the slowest stencil of the radius/precision sweep's table
    "grid3_bf16_n16384": (1, 3, 2, 16384),
}
DT_NAME = {0: "f64", 1: "f32", 2: "bf16"}
BATCH_WORKLOADS = ("batch4096_f32",)
SERVE_WORKLOADS = ("serve",)
SERVE_RATES = (500.0, 2000.0, 8000.0)


def config5_requests(rate: float, seconds: float, seed: int = 11):
    """Poisson arrivals at `rate` per second for `seconds`: n log-uniform in
    [128, 8192], radius uniform in 1..4, star/grid 50/50, iterations uniform
    in 1..200, storage 60% FP32 / 30% FP64 / 10% BF16, seeded U(-1,1) grids.
    Returns (arrival offsets in s, solve specs)."""
    import numpy as np
    rng = np.random.default_rng([seed, int(rate)])
    t, arrivals = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        arrivals.append(t)
    m = len(arrivals)
    ns = np.exp(rng.uniform(np.log(128), np.log(8192), m)).round().astype(int)
    radii = rng.integers(1, 5, m)
    pats = rng.integers(0, 2, m)
    iters = rng.integers(1, 201, m)
    dts = rng.choice([1, 0, 2], size=m, p=[0.6, 0.3, 0.1])
    specs = [(int(p), int(r), int(d), int(n), int(k), 1, 1000003 * int(rate) + i)
             for i, (p, r, d, n, k) in enumerate(zip(pats, radii, dts, ns, iters))]
    return arrivals, specs


def bench_serve(a) -> int:
    """Open-loop serving run per offered rate: a client submits at the Poisson
    arrival times; latency is completion minus scheduled arrival."""
    import numpy as np
    import torch
    from kernels_amd.server import Server

    torch.cuda.set_device(0)
    lines = []
    for rate in a.serve_rates:
        arrivals, specs = config5_requests(rate, a.serve_seconds)
        srv = Server(budget_bytes=int(a.serve_budget_gb * (1 << 30)), max_pending=a.serve_max_pending)
        t0 = Server.clock() + 0.05
        ids, rejected = {}, 0
        for off, spec in zip(arrivals, specs):
            target = t0 + off
            while True:
                d = target - Server.clock()
                if d <= 0:
                    break
                time.sleep(min(d, 0.001))
            rid, rc = srv.try_submit(spec)
            if rid is None:
                rejected += 1
            else:
                ids[rid] = (target, spec)
        t_offer_end = Server.clock()
        srv.drain(600)
        res = srv.poll(len(ids) + 16)
        st = srv.stats()
        srv.close()
        torch.cuda.empty_cache()
        lat = np.array([r["t_done"] - ids[r["id"]][0] for r in res])
        qwait = np.array([r["t_admit"] - ids[r["id"]][0] for r in res])
        t_end = max(r["t_done"] for r in res)
        served_bytes = sum(solve_bytes(ids[r["id"]][1]) for r in res)
        span = t_end - t0
        lines.append({
            "offered_per_s": rate, "offered": len(specs), "accepted": len(ids), "rejected": rejected,
            "served": len(res), "served_per_s": round(len(res) / span, 1),
            "served_per_s_during_offer": round(sum(1 for r in res if r["t_done"] <= t_offer_end)
                                              / (t_offer_end - t0), 1),
            "p50_s": round(float(np.percentile(lat, 50)), 3), "p99_s": round(float(np.percentile(lat, 99)), 3),
            "mean_s": round(float(lat.mean()), 3), "queue_wait_p50_s": round(float(np.percentile(qwait, 50)), 3),
            "effective_gbs": round(served_bytes / span / 1e9, 1), "windows": st["windows"],
            "launches": st["launches"], "host_build_s": round(st["build_seconds"], 2),
            "host_wait_s": round(st["wait_seconds"], 2), "span_s": round(span, 2),
        })
        print(f"# rate {rate}: {json.dumps(lines[-1])}", file=sys.stderr, flush=True)
    best = max(lines, key=lambda x: x["served_per_s"])
    res = {
        "metric": METRIC,
        "value": best["served_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": len(lines),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "mixed (60% f32, 30% f64, 10% bf16)",
        "data": "synthetic (Poisson arrivals, seeded U(-1,1) grids, generator seed 11)",
        "config": {"workload": "served request stream: n log-uniform [128,8192], R 1..4, star/grid, "
                               "iterations 1..200, continuous batching under an HBM budget",
                   "seconds_per_rate": a.serve_seconds, "budget_gb": a.serve_budget_gb,
                   "max_pending": a.serve_max_pending},
        "rates": lines,
    }
    print(json.dumps(res))
    return 0


def config4_solves(seed: int = 7, count: int = 4096):
    """4096 uneven grids: n log-uniform in [64, 2048], star/grid 50/50, radius
    uniform in 1..4, iterations uniform in 1..50, FP32 storage, seeded
    U(-1,1) initial values (solve i uses seed i)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(64), np.log(2048), count)).round().astype(int)
    pats = rng.integers(0, 2, count)
    radii = rng.integers(1, 5, count)
    iters = rng.integers(1, 51, count)
    return [(int(p), int(r), 1, int(n), int(k), 1, i) for i, (p, r, n, k) in enumerate(zip(pats, radii, ns, iters))]


def solve_bytes(spec) -> int:
    """Algorithmic bytes of one solve: init writes in and out, each of the
    K+1 fused sweeps reads in, writes in_next, reads and writes the interior
    of out, the norm reads the interior once."""
    pattern, radius, dtype, n, iters = spec[:5]
    s = {0: 8, 1: 4, 2: 2}[dtype]
    m = n - 2 * radius
    return 2 * s * n * n + (iters + 1) * (2 * s * n * n + 2 * s * m * m) + s * m * m


FP32_BATCH_BOUND = 3.5e-8  # normwise error against FP64 after K+1 sweeps: at most this * (K + 2)


def batch_errors_vs_f64(Batch, K, specs, b) -> dict:
    """Errors of every reduced-precision result of batch `b` (run and
    waited for) against the same solves run in FP64, a second batch outside
    any timed region.  The device computes, per solve, the max and mean
    absolute error and the normwise L2 error over the interior
    (prk_batch_errors).  Returns the worst over all solves, percentiles of
    the normwise error, how many FP32 solves are within 3.5e-8 * (K + 2)
    normwise, and the worst per (pattern, radius, storage) class."""
    import numpy as np
    b64 = Batch([(s[0], s[1], K.F64) + tuple(s[3:]) for s in specs])
    b64.run()
    b64.wait()
    e = b.errors(b64)
    b64.close()
    keys = ("max_abs", "mean_abs", "normwise_l2")
    lp = [i for i, sp in enumerate(specs) if sp[2] != K.F64]
    f32 = [i for i in lp if specs[i][2] == K.F32]
    fmt = lambda v: float(f"{v:.3e}")
    by_class = {}
    for i in lp:
        pattern, radius, dtype = specs[i][:3]
        name = f"{DT_NAME[dtype]} {'star' if pattern == 0 else 'grid'}{radius}"
        c = by_class.setdefault(name, {"solves": 0, **dict.fromkeys(keys, 0.0)})
        c["solves"] += 1
        for k, v in zip(keys, e[i]):
            c[k] = max(c[k], float(v))
    nw = e[lp, 2] if lp else np.zeros(1)
    within = sum(1 for i in f32 if e[i, 2] <= FP32_BATCH_BOUND * (specs[i][4] + 2))
    return {"solves": len(lp), **{k: fmt(float(e[lp, j].max())) if lp else 0.0 for j, k in enumerate(keys)},
            "normwise_l2_p50": fmt(float(np.percentile(nw, 50))),
            "normwise_l2_p99": fmt(float(np.percentile(nw, 99))),
            "fp32_within_bound": within, "fp32_solves": len(f32), "fp32_bound": "3.5e-8*(K+2) normwise",
            "by_class": {k: {kk: (fmt(vv) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                         for k, v in sorted(by_class.items())}}


def bench_batch(a) -> int:
    import numpy as np
    import torch
    import kernels_amd as K
    from kernels_amd.batch import Batch, one_by_one

    from kernels_amd.batch import autotune, tuning_get

    specs = config4_solves()
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    if a.batch_heights:
        # replay the tile heights another run recorded (its JSON line)
        from kernels_amd.batch import tuning_set
        rec = json.load(open(a.batch_heights))["tile_heights"]["by_pattern_radius_log2n"]
        for key, rows in rec.items():
            pr, k = key.split("/2^")
            tuning_set(0 if pr.startswith("star") else 1, int(pr[4:]), 1, 1 << int(k), int(rows))
    t0 = time.perf_counter()
    keys = 0 if (a.no_tune or a.batch_heights) else autotune(specs, sample=16, reps=2)
    t_tune = time.perf_counter() - t0
    heights = {}
    for p_, r_, d_, n_, *_ in specs:
        k = (p_, r_, n_.bit_length() - 1)
        if k not in heights:
            heights[k] = tuning_get(p_, r_, d_, n_)
    t0 = time.perf_counter()
    b = Batch(specs)
    t_create = time.perf_counter() - t0
    info = b.info()
    for _ in range(a.warmup):
        b.run(stream.cuda_stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(a.steps):
        b.run(stream.cuda_stream)
    e1.record(stream)
    norms = b.wait()
    ms = e0.elapsed_time(e1)
    # every FP32 result against the same solve in FP64
    errors = batch_errors_vs_f64(Batch, K, specs, b)
    b.close()
    torch.cuda.empty_cache()
    solves_s = len(specs) * a.steps / (ms * 1e-3)
    total_bytes = sum(solve_bytes(s) for s in specs)
    # unbatched baselines: the same solves one grid per launch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n32 = one_by_one(specs)
    t32 = time.perf_counter() - t0
    t0 = time.perf_counter()
    n64 = one_by_one(specs, dtype_override=K.F64)
    t64 = time.perf_counter() - t0
    rel32 = float(np.max(np.abs(norms - n32) / np.abs(n32)))
    rel64 = float(np.max(np.abs(norms - n64) / np.abs(n64)))
    res = {
        "metric": METRIC,
        "value": round(solves_s, 1),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (4096 seeded U(-1,1) grids, generator seed 7)",
        "config": {"workload": "batched uneven grids: 4096 solves, n log-uniform [64,2048], star/grid 50/50, "
                               "R in 1..4, iterations 1..50, FP32 storage; one hipGraph of round launches",
                   "solves": len(specs), "launches_per_run": info["launches"], "tiles_per_run": info["tiles"],
                   "device_bytes": info["device_bytes"]},
        "effective_gbs": round(total_bytes * a.steps / (ms * 1e-3) / 1e9, 1),
        "algorithmic_bytes_per_run": total_bytes,
        "one_by_one_f64": {"solves_per_s": round(len(specs) / t64, 1), "seconds": round(t64, 3),
                           "max_rel_norm_diff_vs_batch": rel64},
        "one_by_one_f32": {"solves_per_s": round(len(specs) / t32, 1), "seconds": round(t32, 3),
                           "max_rel_norm_diff_vs_batch": rel32},
        "speedup_vs_one_by_one_f64": round(solves_s * t64 / len(specs), 2),
        "speedup_vs_one_by_one_f32": round(solves_s * t32 / len(specs), 2),
        "create_s": round(t_create, 2),
        "fp32_vs_fp64_all": errors,
        "tile_heights": {"autotuned_keys": keys, "tune_s": round(t_tune, 2),
                         "by_pattern_radius_log2n": {f"{'star' if k[0] == 0 else 'grid'}{k[1]}/2^{k[2]}": v
                                                    for k, v in sorted(heights.items())}},
    }
    print(json.dumps(res))
    return 0 if rel32 < 1e-12 else 1


TABLE_WORKLOADS = {
    # the radius/precision sweep's stencils, and the radii beyond the reference's table
    "sweep_table_n16384": [(0, r) for r in range(1, 6)] + [(1, r) for r in range(1, 4)],
    "sweep_table_ext_n16384": [(0, r) for r in range(6, 9)] + [(1, r) for r in range(4, 9)],
}


def bench_table(a) -> int:
    """Stencils at n=16384 in FP64, FP32 and BF16 storage (the radius/
    precision sweep's star R=1..5 and grid R=1..3, or the radii beyond the
    reference's table): autotune, then time min(steps, 20) fused sweeps
    after one warm-up sweep."""
    import torch
    import kernels_amd as K

    n = 16384
    steps = min(a.steps, 20)
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    cases = []
    for dtype in (0, 1, 2):
        bufs = [K.alloc_grid(n, dtype), K.alloc_grid(n, dtype)]
        out = K.alloc_grid(n, dtype)
        for pattern, radius in TABLE_WORKLOADS[a.workload]:
            tuning = K.autotune(pattern, radius, dtype, bufs[0], bufs[1], out, 2, stream)
            K.init(K.INIT_ANALYTIC, 0, dtype, bufs[0], out, stream)
            cur = 0
            K.sweep(pattern, radius, dtype, bufs[cur], bufs[cur ^ 1], out, stream)
            cur ^= 1
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(steps):
                K.sweep(pattern, radius, dtype, bufs[cur], bufs[cur ^ 1], out, stream)
                cur ^= 1
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_time(e1) / steps
            norm = K.l1norm(dtype, radius, out, stream)
            gbs = K.stencil_bytes_per_iter(n, radius, dtype) / (ms * 1e-3) / 1e9
            case = {"stencil": f"{'star' if pattern == 0 else 'grid'}{radius}", "dtype": DT_NAME[dtype],
                    "ms_per_sweep": round(ms, 4), "gbs": round(gbs, 1), "frac_of_8tbs": round(gbs / HBM_PEAK_GBS, 4),
                    "points_per_s": round((n - 2 * radius) ** 2 / (ms * 1e-3), 0),
                    "mflops_reference_formula": round(K.nominal_flops(pattern, radius, n) / (ms * 1e-3) / 1e6, 1),
                    "tuning": {k: v for k, v in tuning.as_dict().items() if k != "gbs"},
                    "validated": None if dtype == 2 else abs(norm - 2.0 * (steps + 1)) <= (1e-8 if dtype == 0 else 1e-4)}
            cases.append(case)
            print(f"# {json.dumps(case)}", file=sys.stderr, flush=True)
        del bufs, out
        torch.cuda.empty_cache()
    head = cases[0]
    peak = {} if a.no_peak else measure_hbm_peak(torch, K)
    res = {
        "metric": METRIC, "value": head["gbs"], "unit": "GB/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": head["ms_per_sweep"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (reference analytic init in[i][j]=i+j, out=0)",
        "config": {"workload": f"{a.workload}: n=16384 fused sweeps of "
                               f"{', '.join(c['stencil'] for c in cases if c['dtype'] == 'f64')}, "
                               "FP64/FP32/BF16 storage, each autotuned"},
        "cases": cases,
    }
    if peak:
        res["hbm_measured"] = {k: (round(v, 1) if isinstance(v, float) else v) for k, v in peak.items()}
    print(json.dumps(res))
    return 0 if all(c["validated"] is not False for c in cases) else 1


TRANSPOSE_WORKLOADS = {"transpose_order32768": 32768}


def bench_transpose(a) -> int:
    """B += A^T, A += 1 on order x order FP64 matrices (Cxx11/transpose.cc):
    a step is one transpose; W untimed, then K timed; the reference's
    validation over all W + K transposes."""
    import torch
    import kernels_amd as K

    order = TRANSPOSE_WORKLOADS[a.workload]
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    A = torch.empty((order, order), dtype=torch.float64, device="cuda")
    B = torch.empty_like(A)
    K.transpose_init(A, B, stream)
    for _ in range(a.warmup):
        K.transpose(A, B, stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(a.steps):
        K.transpose(A, B, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    abserr = K.transpose_check(a.warmup + a.steps - 1, B, stream)
    del A, B
    torch.cuda.empty_cache()
    if a.copy_calibration:
        copy_calibration(torch, K, stream)
    bytes_iter = K.transpose_bytes_per_iter(order)
    gbs = bytes_iter / (ms * 1e-3) / 1e9
    peak = {} if a.no_peak else measure_hbm_peak(torch, K)
    res = {
        "metric": METRIC, "value": round(gbs, 1), "unit": "GB/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (reference init A = 0..order^2-1, B = 0)",
        "config": {"workload": f"matrix transpose B += A^T, A += 1, order={order}", "order": order},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(a.workload)[0],
                     "algorithmic_bytes_per_launch": bytes_iter},
        "validated": abserr < 1e-8, "abserr": abserr,
        "rate_mb_s_reference_formula": round(1e-6 * bytes_iter / (ms * 1e-3), 1),
    }
    if peak:
        res["hbm_measured"] = {k: (round(v, 1) if isinstance(v, float) else v) for k, v in peak.items()}
        res["frac_of_measured_peak"] = round(gbs / max(v for v in peak.values() if isinstance(v, float)), 4)
        res["ratio_to_sweep_mix"] = round(gbs / peak["sweep_mix"], 4)
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_baseline_transpose(order)
    print(json.dumps(res))
    return 0 if res["validated"] else 1


SLAB_WORKLOADS = {
    # name: (pattern, radius, dtype, n, slabs): the middle slab of n rows cut
    # into `slabs` equal slabs, the share of one GPU of a row-strip split
    "star2_f64_slab8_n30000": (0, 2, 0, 30000, 8),
}


def bench_slab(a) -> int:
    """One GPU's share of a row-strip split: the fused sweep of one interior
    slab (prk_sweep_slab) at its real shape, n/P rows plus R halo rows above
    and below, n columns.  A step is one slab sweep.  The halo exchange with
    the neighbouring GPUs is not part of the step (one GPU has no
    neighbours); the halo rows of the new `in` therefore stay as they are, so
    the run is timed but not validated here -- tests/test_gpu_slab.py checks
    slabs with a full exchange against the whole grid bit for bit.  The same
    sweeps over the whole grid are timed after it for comparison."""
    import torch
    import kernels_amd as K

    pattern, radius, dtype, n, parts = SLAB_WORKLOADS[a.workload]
    rows = n // parts
    row0 = rows * (parts // 2)
    top, bot = K.slab_halos(n, radius, row0, rows)
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    # the slab's grids before the tuner's n x n scratch grids come and go
    bufs = [K.alloc_slab(n, radius, row0, rows, dtype) for _ in range(2)]
    out = K.alloc_slab(n, radius, row0, rows, dtype)
    tuning = K.Tuning()
    if a.tuning_file and a.no_tune and os.path.exists(a.tuning_file):
        loaded = K.lib().prk_tuning_load(a.tuning_file.encode())  # entries read, or < 0
        if loaded < 0:
            K.check(loaded)
    if not a.no_tune:
        K.check(K.lib().prk_autotune(pattern, radius, dtype, n, 2, tuning))
        if a.tuning_file:
            K.check(K.lib().prk_tuning_save(a.tuning_file.encode()))
    else:
        K.check(K.lib().prk_tuning_get(pattern, radius, dtype, n, tuning))
    # the reference's analytic values at the slab's global rows
    gi = torch.arange(row0 - top, row0 + rows + bot, device="cuda", dtype=torch.float64)
    gj = torch.arange(n, device="cuda", dtype=torch.float64)
    bufs[0].copy_(gi[:, None] + gj[None, :])
    bufs[1].copy_(bufs[0])
    out.zero_()

    def timed(step):
        for _ in range(a.warmup):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(a.steps):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / a.steps

    cur = [0]

    def slab_step():
        K.sweep_slab(pattern, radius, dtype, n, row0, rows, bufs[cur[0]], bufs[cur[0] ^ 1], out, stream)
        cur[0] ^= 1

    ms = timed(slab_step)
    del bufs, out, gi, gj
    torch.cuda.empty_cache()
    if a.copy_calibration:
        copy_calibration(torch, K, stream)
    ms_whole = None
    if not a.no_whole:
        whole = [K.alloc_grid(n, dtype) for _ in range(3)]
        K.init(K.INIT_ANALYTIC, 0, dtype, whole[0], whole[2], stream)
        wc = [0]

        def whole_step():
            K.sweep(pattern, radius, dtype, whole[wc[0]], whole[wc[0] ^ 1], whole[2], stream)
            wc[0] ^= 1

        ms_whole = timed(whole_step)
        del whole
        torch.cuda.empty_cache()
    bytes_iter = K.slab_bytes_per_iter(n, radius, dtype, row0, rows)
    gbs = bytes_iter / (ms * 1e-3) / 1e9
    halo_bytes = (top + bot) * n * {0: 8, 1: 4, 2: 2}[dtype]
    res = {
        "metric": METRIC, "value": round(gbs, 1), "unit": "GB/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": DT_NAME[dtype],
        "data": "synthetic (reference analytic init at the slab's global rows, out=0)",
        "config": {"workload": f"{'star' if pattern == 0 else 'grid'} R={radius} {DT_NAME[dtype]} n={n}: "
                               f"fused sweep of one of {parts} row slabs (rows {row0}..{row0 + rows - 1} "
                               f"plus {top}+{bot} halo rows), exchange not included",
                   "n": n, "slabs": parts, "row0": row0, "rows": rows, "halo_rows": [top, bot]},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(a.workload)[0],
                     "algorithmic_bytes_per_launch": bytes_iter},
        "tuning": {**tuning.as_dict(), "autotuned": not a.no_tune},
        "validated": None,
        "whole_grid_ms_per_sweep": None if ms_whole is None else round(ms_whole, 4),
        "slab_ms_over_whole_ms_per_slab": None if ms_whole is None else round(ms / (ms_whole / parts), 4),
        "halo_bytes_received_per_sweep": halo_bytes,
    }
    print(json.dumps(res))
    return 0


def cpu_baseline_transpose(order: int):
    """The reference's OpenMP transpose (Cxx11/transpose-openmp.cc, built
    from the reference sources into oracle/_ref) on this host's cores, 3
    timed iterations (plus the untimed first) at the same order."""
    exe = ROOT / "oracle" / "_ref" / "transpose-openmp"
    if not exe.exists():
        return None
    iters = 3
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([str(exe), str(iters), str(order), "32"], capture_output=True, text=True, env=env,
                       timeout=900)
    m = re.search(r"Avg time \(s\):\s*([0-9.eE+-]+)", r.stdout)
    t = re.search(r"Number of threads\s*=\s*(\d+)", r.stdout)
    if r.returncode != 0 or not m or "Solution validates" not in r.stdout:
        return None
    avg = float(m.group(1))
    return {"value": 4 * 8 * order * order / avg / 1e9, "unit": "GB/s", "cores": int(t.group(1)) if t else threads,
            "kind": "reference", "sample": f"{iters} timed + 1 untimed transposes, order={order}",
            "avg_s_per_iter": avg}


def traffic_from_profiles(workload: str, tiling=None):
    """HBM bytes per sweep launch measured with rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE, scaled by the copy-kernel calibration (profiles/).  A
    profile lists one `traffic_bytes_per_launch[variant rows wgs order]:`
    line per tiling it measured; the figure returned is the one of `tiling`
    (variant, rows_per_wave, wgs_per_cu, order), or None when that tiling was
    not profiled.  Returns (bytes or None, the profiled tilings)."""
    p = ROOT / "profiles" / f"sweep_{workload}.md"
    if not p.exists():
        p = ROOT / "profiles" / f"{workload}.md"
    if not p.exists():
        return None, []
    text = p.read_text()
    rows = {tuple(int(x) for x in m.group(1).split()): float(m.group(2))
            for m in re.finditer(r"^traffic_bytes_per_launch\[([0-9 ]+)\]:\s*([0-9.eE+]+)", text, re.M)}
    profiled = [" ".join(map(str, k)) for k in sorted(rows)]
    if tiling is None:
        m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", text, re.M)
        return (float(m.group(1)) if m else None), profiled
    return rows.get(tuple(tiling)), profiled


def copy_calibration(torch, K, stream) -> None:
    """One prk_copy of exactly 2 GiB (2 GiB read, 2 GiB written): the known
    traffic that scales the FETCH_SIZE / WRITE_SIZE counters."""
    nb = 2 << 30
    src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    K.check(K.lib().prk_copy(nb, dst.data_ptr(), src.data_ptr(), stream.cuda_stream))
    torch.cuda.synchronize()
    del src, dst
    torch.cuda.empty_cache()


# Copy shapes (16-byte vectors per lane, non-temporal, workgroups; 0 = one
# per chunk) of the copy-shape sweep; the default line measures the
# COPY_SHAPES_PEAK subset, the fastest of the sweep.
COPY_SHAPES = [(v, nt, b) for v in (1, 2, 4, 8) for nt in (1, 0) for b in (0, 1024, 2048, 4096)]
COPY_SHAPES_PEAK = [(1, 1, 0), (2, 1, 0), (4, 1, 0), (4, 0, 0), (2, 1, 2048), (4, 1, 2048)]


def time_copy_shapes(torch, K, shapes, nbytes: int = 4 << 30, reps: int = 10) -> dict:
    """GB/s (read + written bytes) of prk_copy_shape per shape, on nbytes."""
    s = torch.cuda.current_stream()
    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    out = {}
    for vpl, nt, blocks in shapes:
        def fn():
            K.check(K.lib().prk_copy_shape(nbytes, dst.data_ptr(), src.data_ptr(), vpl, nt, blocks, s.cuda_stream))
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        e1.synchronize()
        out[f"vpl{vpl}_nt{nt}_wg{blocks}"] = 2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    ok = bool(torch.equal(src, dst))
    del src, dst
    torch.cuda.empty_cache()
    if not ok:
        raise RuntimeError("copy shape produced a wrong copy")
    return out


def bench_copy_shapes(a) -> int:
    """The copy probe in every shape of COPY_SHAPES on 4 GiB (10 copies
    each): the fastest is the measured peak the stencil is held against."""
    import torch
    import kernels_amd as K

    torch.cuda.set_device(0)
    rates = time_copy_shapes(torch, K, COPY_SHAPES)
    best = max(rates, key=rates.get)
    res = {
        "metric": METRIC, "value": round(rates[best], 1), "unit": "GB/s", "n_gpus": 1, "steps": 10,
        "warmup": 1, "ms_per_step": round(2 * (4 << 30) / (rates[best] * 1e9) * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic (4 GiB of ones)",
        "config": {"workload": "copy probe, 4 GiB, every shape (vectors per lane, non-temporal, workgroups)"},
        "best_shape": best, "peak_fraction_of_8tbs": round(rates[best] / HBM_PEAK_GBS, 4),
        "shapes_gbs": {k: round(v, 1) for k, v in sorted(rates.items(), key=lambda kv: -kv[1])},
    }
    print(json.dumps(res))
    return 0


def measure_hbm_peak(torch, K, nbytes: int = 4 << 30, reps: int = 10):
    """Own copy, triad and sweep-mix kernels: the measured HBM peak on this
    GPU (the fastest of the copy shapes in COPY_SHAPES_PEAK, the fastest of
    the full copy-shape sweep), and the rate of a plain elementwise kernel
    with the sweep's mix of two read and two write streams (a comparison,
    not a bound: the tiled sweep moves its bytes a little faster than it)."""
    shapes = time_copy_shapes(torch, K, COPY_SHAPES_PEAK, nbytes, reps)
    s = torch.cuda.current_stream()
    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    c = torch.empty_like(src).fill_(2.0)
    best = max(shapes, key=shapes.get)
    res = {"copy_best": shapes[best], "copy_best_shape": best}
    for name, fn, bytes_per in (
            ("copy", lambda: K.check(K.lib().prk_copy(nbytes, dst.data_ptr(), src.data_ptr(), s.cuda_stream)),
             2 * nbytes),
            ("triad", lambda: K.check(K.lib().prk_triad(src.numel(), 3.0, dst.data_ptr(), src.data_ptr(),
                                                         c.data_ptr(), s.cuda_stream)), 4 * nbytes),
            ("sweep_mix", lambda: K.check(K.lib().prk_stream_mix(src.numel(), src.data_ptr(), dst.data_ptr(),
                                                                  c.data_ptr(), s.cuda_stream)), 4 * nbytes)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        e1.synchronize()
        res[name] = bytes_per * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst, c
    torch.cuda.empty_cache()
    return res


def cpu_baseline(pattern: int, radius: int, n: int):
    """The reference's OpenMP driver (Cxx11/stencil-openmp.cc, built from the
    reference sources into oracle/_ref) on this host's cores; if it was not
    built, the oracle's OpenMP restatement.  Bounded sample: 8 timed sweeps
    (plus the warm-up sweep) of the same n x n grid, about 1.6 s on 16 host
    cores at n=30000 (some 26 core-seconds of work)."""
    iters = 8
    bytes_iter = 2 * 8 * n * n + 2 * 8 * (n - 2 * radius) ** 2
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    exe = ROOT / "oracle" / "_ref" / "stencil-openmp"
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    sample = f"{iters} timed + 1 warm-up sweeps, FP64 {'star' if pattern == 0 else 'grid'} R={radius}, n={n}"
    if exe.exists():
        r = subprocess.run([str(exe), str(iters), str(n), "32", "star" if pattern == 0 else "grid", str(radius)],
                           capture_output=True, text=True, env=env, timeout=600)
        m = re.search(r"Avg time \(s\):\s*([0-9.eE+-]+)", r.stdout)
        t = re.search(r"Number of threads\s*=\s*(\d+)", r.stdout)
        if r.returncode == 0 and m and "Solution validates" in r.stdout:
            avg = float(m.group(1))
            return {"value": bytes_iter / avg / 1e9, "unit": "GB/s", "cores": int(t.group(1)) if t else threads,
                    "kind": "reference", "sample": sample, "avg_s_per_iter": avg,
                    "mflops": float(re.search(r"Rate \(MFlops/s\):\s*([0-9.eE+-]+)", r.stdout).group(1))}
    from tests._oracle import Oracle
    avg, norm, team = Oracle().omp_run(pattern, radius, n, 32, iters, threads)
    return {"value": bytes_iter / avg / 1e9, "unit": "GB/s", "cores": team, "kind": "port", "sample": sample,
            "avg_s_per_iter": avg}


def fp32_storage_section(a, torch, K, ms_f64: float) -> dict:
    """The FP32-storage form of the headline stencil (FP32 star R=2,
    n=30000): its points/s against the FP64 sweep just timed (analytic init,
    tuned on its own grids, min(steps, 20) sweeps after one warm-up sweep,
    validated to 1e-4), and its error against FP64 after 50 sweeps from
    seeded U(-1,1) values (seed 42) over the whole interior: max and mean
    absolute and relative (relative to max(|y|, 1e-3 mean|y|)) and the
    normwise L2 error, beside the 1e-6 the FP32 form aims for."""
    n, radius, pattern = 30000, 2, 0
    stream = torch.cuda.current_stream()
    steps = min(a.steps, 20)
    bufs = [K.alloc_grid(n, K.F32), K.alloc_grid(n, K.F32)]
    out = K.alloc_grid(n, K.F32)
    if not a.no_tune:
        K.autotune(pattern, radius, K.F32, bufs[0], bufs[1], out, 2, stream)
    K.init(K.INIT_ANALYTIC, 0, K.F32, bufs[0], out, stream)
    K.sweep(pattern, radius, K.F32, bufs[0], bufs[1], out, stream)
    cur = 1
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        K.sweep(pattern, radius, K.F32, bufs[cur], bufs[cur ^ 1], out, stream)
        cur ^= 1
    e1.record(stream)
    e1.synchronize()
    ms32 = e0.elapsed_time(e1) / steps
    norm = K.l1norm(K.F32, radius, out, stream)
    validated = abs(norm - 2.0 * (steps + 1)) <= 1e-4
    # seeded init, 50 sweeps after the first, FP32 then FP64
    iters = 50
    K.init(K.INIT_SEEDED, 42, K.F32, bufs[0], out, stream)
    K.run(pattern, radius, K.F32, iters, bufs[0], bufs[1], out, stream)
    del bufs
    torch.cuda.empty_cache()
    b64 = [K.alloc_grid(n, K.F64), K.alloc_grid(n, K.F64)]
    y = K.alloc_grid(n, K.F64)
    K.init(K.INIT_SEEDED, 42, K.F64, b64[0], y, stream)
    K.run(pattern, radius, K.F64, iters, b64[0], b64[1], y, stream)
    del b64
    torch.cuda.empty_cache()
    tot = {"sum_abs": 0.0, "sum_rel": 0.0, "max_abs": 0.0, "max_rel": 0.0, "d2": 0.0, "y2": 0.0, "l1x": 0.0, "l1y": 0.0}
    inner = slice(radius, n - radius)
    tau = 1e-3 * float(y[inner, inner].abs().mean())
    for r0 in range(radius, n - radius, 2048):
        r1 = min(r0 + 2048, n - radius)
        yy = y[r0:r1, inner]
        d = (out[r0:r1, inner].double() - yy).abs()
        rel = d / torch.clamp(yy.abs(), min=tau)
        tot["sum_abs"] += float(d.sum()); tot["sum_rel"] += float(rel.sum())
        tot["max_abs"] = max(tot["max_abs"], float(d.max())); tot["max_rel"] = max(tot["max_rel"], float(rel.max()))
        tot["d2"] += float((d * d).sum()); tot["y2"] += float((yy * yy).sum())
        tot["l1x"] += float(out[r0:r1, inner].double().abs().sum()); tot["l1y"] += float(yy.abs().sum())
    pts = (n - 2 * radius) ** 2
    normwise = (tot["d2"] / tot["y2"]) ** 0.5
    del out, y
    torch.cuda.empty_cache()
    return {
        "workload": f"star R=2 f32 storage (f32 accumulation) n={n} fused sweep",
        "ms_per_step": round(ms32, 4), "steps": steps,
        "gbs": round(K.stencil_bytes_per_iter(n, radius, K.F32) / (ms32 * 1e-3) / 1e9, 1),
        "points_per_s": round((n - 2 * radius) ** 2 / (ms32 * 1e-3), 0),
        "points_per_s_vs_f64": round(ms_f64 / ms32, 3), "target_points_per_s_vs_f64": 1.8,
        "validated": bool(validated),
        "error_vs_f64_seeded": {"iterations": iters, "seed": 42, "normwise_l2": float(f"{normwise:.3e}"),
                                "max_abs": float(f"{tot['max_abs']:.3e}"),
                                "mean_abs": float(f"{tot['sum_abs'] / pts:.3e}"),
                                "max_rel": float(f"{tot['max_rel']:.3e}"),
                                "mean_rel": float(f"{tot['sum_rel'] / pts:.3e}"),
                                # the validated quantity: the interior L1 norm, FP32 against FP64
                                "l1_norm_rel": float(f"{abs(tot['l1x'] - tot['l1y']) / tot['l1y']:.3e}"),
                                "target_normwise": 1e-6, "target_met": bool(normwise <= 1e-6)},
    }


def batch_section(a, torch, K) -> dict:
    """The 4096 uneven grids of the batched workload (FP32 storage) in one
    hipGraph of round launches, against the same solves one grid per launch
    in FP64: solves/s of each (after tuning the tile heights unless
    --no-tune; one warm-up run, then three timed runs of the batch)."""
    import numpy as np
    from kernels_amd.batch import Batch, autotune, one_by_one

    specs = config4_solves()
    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    if not a.no_tune:
        autotune(specs, sample=16, reps=2)
    t_tune = time.perf_counter() - t0
    b = Batch(specs)
    b.run(stream.cuda_stream)
    b.wait()
    runs = 3
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(runs):
        b.run(stream.cuda_stream)
    e1.record(stream)
    norms = b.wait()
    ms = e0.elapsed_time(e1) / runs
    errors = batch_errors_vs_f64(Batch, K, specs, b)
    b.close()
    torch.cuda.empty_cache()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n64 = one_by_one(specs, dtype_override=K.F64)
    t64 = time.perf_counter() - t0
    solves_s = len(specs) / (ms * 1e-3)
    return {
        "workload": "4096 grids (seed 7): n log-uniform [64,2048], star/grid, R 1..4, iterations 1..50, "
                    "FP32 storage, one hipGraph of round launches",
        "solves_per_s": round(solves_s, 1), "ms_per_run": round(ms, 2), "runs": runs,
        "effective_gbs": round(sum(solve_bytes(s) for s in specs) / (ms * 1e-3) / 1e9, 1),
        "one_by_one_f64_solves_per_s": round(len(specs) / t64, 1),
        "speedup_vs_one_by_one_f64": round(solves_s * t64 / len(specs), 2), "target_speedup": 5.0,
        "max_rel_norm_diff_vs_f64": float(f"{float(np.max(np.abs(norms - n64) / np.abs(n64))):.3e}"),
        "fp32_vs_fp64_all": errors,
        "tile_tune_s": round(t_tune, 2),
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="star2_f64_n30000",
                    choices=sorted(WORKLOADS) + list(BATCH_WORKLOADS) + list(SERVE_WORKLOADS) + list(TABLE_WORKLOADS)
                            + list(TRANSPOSE_WORKLOADS) + list(SLAB_WORKLOADS) + ["copy_shapes"])
    ap.add_argument("--serve-rates", type=float, nargs="+", default=list(SERVE_RATES))
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-max-pending", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-peak", action="store_true", help="skip the copy/triad peak measurement")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (use the built-in default tiling)")
    ap.add_argument("--tuning-file", default=None,
                    help="save the autotuned table here (or, with --no-tune, load it from here)")
    ap.add_argument("--no-whole", action="store_true",
                    help="slab workloads: skip the whole-grid comparison sweeps")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the sweeps, run one 2 GiB copy (known bytes) for counter calibration")
    ap.add_argument("--batch-heights", default=None,
                    help="batch4096_f32: replay the tile heights recorded in this bench JSON line (no tuning)")
    ap.add_argument("--no-extra", action="store_true",
                    help="default workload: skip the FP32-storage and batched (4096 grids) sections")
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): the halo exchange between GPUs is "
              "not implemented; one GPU's share of a row-strip split is measured by "
              "--workload star2_f64_slab8_n30000", file=sys.stderr)
        return 2
    if a.steps < 1 or a.warmup < 0:
        print("need --steps >= 1 and --warmup >= 0", file=sys.stderr)
        return 2

    if a.workload in BATCH_WORKLOADS:
        return bench_batch(a)
    if a.workload in TABLE_WORKLOADS:
        return bench_table(a)
    if a.workload in TRANSPOSE_WORKLOADS:
        return bench_transpose(a)
    if a.workload in SLAB_WORKLOADS:
        return bench_slab(a)
    if a.workload == "copy_shapes":
        return bench_copy_shapes(a)
    if a.workload in SERVE_WORKLOADS:
        # The first run of this open-loop serving benchmark on MI355X (500
        # req/s for 10 s, 200 GiB pool) ended in a host segmentation fault
        # with no output.  The server's host logic runs clean under
        # AddressSanitizer and ThreadSanitizer on the CPU with a recording HIP
        # stand-in at the same arrival rates, so the cause is not found; the
        # workload stays switched off until it is.
        print("the serving benchmark is switched off: its first run on MI355X ended in a host "
              "segmentation fault whose cause has not been found", file=sys.stderr)
        return 2

    import torch
    import kernels_amd as K

    pattern, radius, dtype, n = WORKLOADS[a.workload]
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    # The grids come first and the tuner runs on them: tuning on scratch
    # grids freed before these are allocated left the sweep 3 % slower.
    bufs = [K.alloc_grid(n, dtype), K.alloc_grid(n, dtype)]
    out = K.alloc_grid(n, dtype)
    tuning = K.Tuning()
    t_tune = time.perf_counter()
    if a.tuning_file and a.no_tune and os.path.exists(a.tuning_file):
        K.lib().prk_tuning_load(a.tuning_file.encode())
    if not a.no_tune:
        tuning = K.autotune(pattern, radius, dtype, bufs[0], bufs[1], out, 2, stream)
        if a.tuning_file:
            K.check(K.lib().prk_tuning_save(a.tuning_file.encode()))
    else:
        K.check(K.lib().prk_tuning_get(pattern, radius, dtype, n, tuning))
    t_tune = time.perf_counter() - t_tune
    K.init(K.INIT_ANALYTIC, 0, dtype, bufs[0], out, stream)
    cur = 0
    for _ in range(a.warmup):
        K.sweep(pattern, radius, dtype, bufs[cur], bufs[cur ^ 1], out, stream)
        cur ^= 1
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    e0.record(stream)
    for _ in range(a.steps):
        K.sweep(pattern, radius, dtype, bufs[cur], bufs[cur ^ 1], out, stream)
        cur ^= 1
    e1.record(stream)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    ms = e0.elapsed_time(e1)
    sweeps = a.warmup + a.steps
    # Run-to-run spread after the timed region: five more timed runs of up
    # to 20 sweeps each (they count towards the validated sweep total).
    rep_n = min(a.steps, 20)
    rep_ms = []
    for _ in range(5):
        e0.record(stream)
        for _ in range(rep_n):
            K.sweep(pattern, radius, dtype, bufs[cur], bufs[cur ^ 1], out, stream)
            cur ^= 1
        e1.record(stream)
        e1.synchronize()
        rep_ms.append(e0.elapsed_time(e1) / rep_n)
    sweeps += 5 * rep_n
    norm = K.l1norm(dtype, radius, out, stream)
    eps = {0: 1e-8, 1: 1e-4, 2: 1e-2}[dtype]
    # BF16 (8-bit significand) cannot hold i + j + k beyond 256, so the
    # analytic closed form does not apply to it; its parity is checked
    # bitwise against the oracle's BF16 arithmetic in the tests instead.
    validated = None if dtype == 2 else abs(norm - 2.0 * sweeps) <= eps
    row_pitch = out.stride(0)
    del bufs, out
    torch.cuda.empty_cache()
    if a.copy_calibration:
        copy_calibration(torch, K, stream)

    bytes_iter = K.stencil_bytes_per_iter(n, radius, dtype)
    ms_step = ms / a.steps
    gbs = bytes_iter / (ms_step * 1e-3) / 1e9
    # the tiling that ran (tuned, or the table's entry with --no-tune)
    used = K.Tuning()
    K.check(K.lib().prk_tuning_get(pattern, radius, dtype, n, used))
    tiling_key = (used.variant, used.rows_per_wave, used.wgs_per_cu, used.order)
    traffic, profiled = traffic_from_profiles(a.workload, tiling_key)
    extra = {}
    if not a.no_extra and a.workload == "star2_f64_n30000":
        extra["fp32_storage"] = fp32_storage_section(a, torch, K, ms_step)
        extra["batched_config4"] = batch_section(a, torch, K)
    peak = {} if a.no_peak else measure_hbm_peak(torch, K)
    res = {
        "metric": METRIC,
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DT_NAME[dtype],
        "data": "synthetic (reference analytic init in[i][j]=i+j, out=0)",
        "config": {"workload": f"{'star' if pattern == 0 else 'grid'} R={radius} {DT_NAME[dtype]} "
                               f"n={n} fused sweep (out+=S(in), in_next=in+1)",
                   "pattern": "star" if pattern == 0 else "grid", "radius": radius, "n": n,
                   "iterations": a.steps, "tile": 32, "row_pitch_elements": row_pitch},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_tiling": " ".join(map(str, tiling_key)),
                     "traffic_profiled_tilings": profiled,
                     "algorithmic_bytes_per_launch": bytes_iter},
        "tuning": {**tuning.as_dict(), "autotuned": not a.no_tune, "tune_s": round(t_tune, 2)},
        "validated": validated if validated is None else bool(validated),
        "norm": norm,
        "mflops_reference_formula": round(K.nominal_flops(pattern, radius, n) / (ms_step * 1e-3) / 1e6, 1),
        "points_per_s": (n - 2 * radius) ** 2 / (ms_step * 1e-3),
        "wall_s_timed_region": round(t_wall, 4),
        "repeats": {"runs": 5, "sweeps_each": rep_n,
                    "gbs": [round(bytes_iter / (m * 1e-3) / 1e9, 1) for m in rep_ms],
                    "gbs_median": round(bytes_iter / (sorted(rep_ms)[2] * 1e-3) / 1e9, 1)},
    }
    if peak:
        res["hbm_measured"] = {k: (round(v, 1) if isinstance(v, float) else v) for k, v in peak.items()}
        res["frac_of_measured_peak"] = round(gbs / max(v for v in peak.values() if isinstance(v, float)), 4)
        res["ratio_to_sweep_mix"] = round(gbs / peak["sweep_mix"], 4)
    res.update(extra)
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_baseline(pattern, radius, n) if dtype == 0 else None
    print(json.dumps(res))
    bad = validated is False or extra.get("fp32_storage", {}).get("validated") is False
    return 1 if bad else 0


if __name__ == "__main__":
    raise SystemExit(main())
