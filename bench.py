#!/usr/bin/env python3
"""Benchmark of the HOME-LBM time step on one MI355X.

Default workload: D3Q27 flow past a static sphere, 256x128x128, FP32 moments,
Re = 200, seeded inlet noise. One step = the fused stream/collide/boundary
kernel plus the per-solid FP64 force reduction the reference performs inside
its step. Prints ONE JSON line (metric, MLUPS, roofline of the step kernel,
CPU-baseline of the oracle on the host cores).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload d3q27_sphere]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "D3Q27 MLUPS + % HBM peak; batched-env MLUPS with p50/p99 step latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--workload", default="d3q27_sphere", choices=["d3q27_sphere", "d2q9_karman"])
    p.add_argument("--dtype", default="f32", choices=["f32", "f16", "bf16", "f64"])
    p.add_argument("--kernel", default="auto", choices=["auto", "gather", "tiled"])
    p.add_argument("--cpu-seconds", type=float, default=10.0, help="bound on the CPU-baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def cpu_baseline(w, seconds: float):
    """Time the oracle's OpenMP FP32 step on the host cores (bounded sample)."""
    from oracle import oracle as orc  # the checker / baseline, never the product

    env = orc.Env(w.lattice, w.nx, w.ny, w.nz, nu=w.nu, bc_type=w.bc_type, bc_value=w.bc_value)
    env.mask = orc.voxelize(env, w.capsules)
    env.solids = w.poses
    m = orc.init_uniform(env, w.u_in, "f32")
    d = 2 if w.lattice == 9 else 3
    m[1:1 + d] += w.velocity_noise()
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    steps = 0
    t0 = time.perf_counter()
    cur = m
    while True:
        cur = orc.step(env, cur, "f32", omp=True)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 500:
            break
    return {"value": round(w.ncell * steps / el / 1e6, 3), "unit": "MLUPS", "cores": cores, "kind": "port",
            "sample": f"{steps} FP32 steps of the full {w.name} lattice, OpenMP oracle, {el:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from openhomelbm_amd import _native as N
    from openhomelbm_amd import inputs

    kernel = {"auto": N.HL_KERNEL_AUTO, "gather": N.HL_KERNEL_GATHER, "tiled": N.HL_KERNEL_TILED}[args.kernel]
    w = inputs.d3q27_sphere() if args.workload == "d3q27_sphere" else inputs.d2q9_karman()
    log(f"[bench] workload {w.name}: {w.description}; dtype {args.dtype}; kernel {args.kernel}")
    lat = inputs.build_lattice(w, args.dtype)
    nsolid = lat.nsolid
    K, W = args.steps, args.warmup
    hist = torch.zeros((max(K, W, 1), nsolid, 6), dtype=torch.float64, device=lat.device)
    mass0 = lat.mass()
    torch.cuda.synchronize()
    lat.step(W, kernel, force_history=hist)
    torch.cuda.synchronize()
    log(f"[bench] warmup {W} steps done")

    # timed region: K full steps (step kernel + force reduction), inputs resident in HBM
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    lat.step(K, kernel, force_history=hist)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = e0.elapsed_time(e1)
    ms_per_step = ms_total / K
    log(f"[bench] timed {K} steps: {ms_total:.2f} ms (events), {wall * 1e3:.2f} ms (wall)")

    # step kernel alone, same stream, for the roofline
    k0, k1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    KK = max(K // 2, 1)
    torch.cuda.synchronize()
    k0.record()
    lat.step(KK, kernel)
    k1.record()
    torch.cuda.synchronize()
    kernel_ms = k0.elapsed_time(k1) / KK
    mass1 = lat.mass()
    f_last = hist[K - 1, 0].cpu().numpy() if K > 0 else np.zeros(6)
    if not np.isfinite(lat.fields()[0].float().sum().item()):
        raise RuntimeError("non-finite density after the benchmark run")

    mlups = w.ncell * K / (ms_total * 1e-3) / 1e6
    b_cell = lat.bytes_per_cell_update
    traffic_alg = b_cell * w.ncell
    achieved = traffic_alg / (kernel_ms * 1e-3) / 1e9
    log(f"[bench] {mlups:.0f} MLUPS; step kernel {kernel_ms * 1e3:.1f} us -> {achieved:.0f} GB/s "
        f"({achieved / HBM_PEAK_GBS * 100:.1f}% of HBM peak); mass drift {(mass1 - mass0) / mass0:.3e}; "
        f"drag {f_last[0]:.5f}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, args.cpu_seconds)
        log(f"[bench] cpu baseline {cpu}")

    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded PCG64 inlet noise; static sphere mask)",
        "config": {"workload": w.name, "lattice": "D3Q27" if w.lattice == 27 else "D2Q9",
                   "grid": [w.nx, w.ny, w.nz], "nu": w.nu, "u_in": w.u_in[0], "kernel": args.kernel,
                   "step": "stream/collide/boundary kernel + FP64 per-solid force reduction",
                   "bytes_per_cell_update": b_cell},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel_us": round(kernel_ms * 1e3, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
