#!/usr/bin/env python3
"""Permeability solves/s of the SIMPLE-family outer iteration on one MI355X.

Workload (one step = one complete permeability solve): a 128^3 seeded
random-sequential-addition pack of non-overlapping spheres of radius 10 voxels
at porosity 0.75, flow along x, unit pressure drop, mu = rho = 1, voxel size 1,
SIMPLEC to an outer tolerance of 1e-6 on the relative mass residual, inner
tolerances 1e-10 (velocity, pressure) and 1e-6 (pressure correction), FP64.

The images are generated and uploaded (index scan included) before the timed
region; the timed region is K back-to-back solves on device-resident inputs.
Output: one JSON line (see the keys below); progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "permeability solves/s + p50/p99 latency, mixed vs FP64 GPU at |Δk|/k≤1e-4"
PEAK_HBM_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def workload(args):
    from ns3d_amd import generators
    n = args.size
    img, solid, overlap, nsph = generators.rsa_spheres((n, n, n), args.radius, 1.0 - args.porosity, args.seed)
    params = dict(dx=1.0, mu=1.0, rho=1.0, p0=1.0, gradp=1.0 / n, alpha=0.8, beta=1.0, vsolvertol=1e-10,
                  psolvertol=1e-10, pcsolvertol=1e-6, maxitsolver=5000, direction=1, border=0)
    desc = dict(workload=f"{n}^3 RSA sphere pack r={args.radius:g} porosity {1 - solid:.4f} (target "
                         f"{args.porosity}), {args.method}, flow x, simpletol {args.simpletol:g}",
                image=dict(shape=[n, n, n], radius=args.radius, seed=args.seed, porosity=round(1 - solid, 6),
                           rsa_overlap_voxels=overlap, spheres=nsph),
                method=args.method, simpletol=args.simpletol, inner_tol=[1e-10, 1e-10, 1e-6])
    return img, params, desc


def cpu_baseline(img, params, args, gpu_solver_iters):
    """The oracle (CPU restatement, OpenMP SpMV/sweeps, serial vector ops) on
    the same image: a bounded sample of outer iterations, extrapolated to a
    whole solve by the GPU run's BiCGSTAB iteration count."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    o = oracle_lib.Oracle(img, **{k: v for k, v in params.items() if k not in ("direction",)})
    t0 = time.perf_counter()
    rows, summ = o.solve(args.method, args.simpletol, args.cpu_iters, 4, 4)
    dt = time.perf_counter() - t0
    cpu_iters = sum(summ["solver_iterations"])
    per_iter = dt / max(cpu_iters, 1)
    est = per_iter * sum(gpu_solver_iters)
    threads = oracle_lib.lib().orc_num_threads()
    return dict(value=1.0 / est if est > 0 else None, unit="solves/s", cores=int(threads), kind="port",
                sample=f"{args.cpu_iters} outer iterations of the same {args.method} solve on the same image "
                       f"({cpu_iters} BiCGSTAB iterations, {dt:.1f} s), extrapolated to the GPU run's "
                       f"{sum(gpu_solver_iters)} BiCGSTAB iterations per solve",
                sample_seconds=round(dt, 3))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--radius", type=float, default=10.0)
    ap.add_argument("--porosity", type=float, default=0.75)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--method", default="SIMPLEC")
    ap.add_argument("--simpletol", type=float, default=1e-6)
    ap.add_argument("--maxit", type=int, default=3000)
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="no per-kernel events (profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU runs are not part of "
              f"this build", file=sys.stderr)
        sys.exit(2)

    from ns3d_amd import gpu, solver
    if gpu.device_count() < 1:
        raise SystemExit("no GPU visible")
    log("generating workload")
    img, params, desc = workload(args)
    run_kw = dict(params, method=args.method, simpletol=args.simpletol, maxitSIMPLE=args.maxit, N1=4, N2=4,
                  device=0)

    log(f"warmup: {args.warmup} solve(s)")
    for w in range(args.warmup):
        s = solver.Solver(img, **run_kw)
        rows, summ = s.run()
        log(f"  warmup {w}: {summ['iterations']} outer its, K={summ['K']:.10e}, {summ['seconds']:.2f} s")
        s.close()

    log(f"uploading {args.steps} image(s)")
    solvers = [solver.Solver(img, **run_kw) for _ in range(args.steps)]
    if not args.no_timing:
        for s in solvers:
            s.timing(True)
    gpu.lib()
    for s in solvers:
        gpu._check(gpu.lib().ns3d_sync(s.context()))
    lat, summs = [], []
    t_start = time.perf_counter()
    for q, s in enumerate(solvers):
        t0 = time.perf_counter()
        rows, summ = s.run()
        lat.append(time.perf_counter() - t0)
        summs.append(summ)
        log(f"  step {q}: {summ['iterations']} outer its, converged={summ['converged']}, K={summ['K']:.10e}, "
            f"{lat[-1]:.2f} s")
    wall = time.perf_counter() - t_start

    s0 = solvers[0]
    np_, nu, nv, nw = s0.np, s0.nu, s0.nv, s0.nw
    nnz = s0.nnz()
    outer = summs[0]["iterations"]
    roof = None
    kernels = {}
    if not args.no_timing:
        tot = {}
        for s in solvers:
            for key, (ms, launches, nbytes) in s.timing_read().items():
                a = tot.setdefault(key, [0.0, 0.0, 0.0])
                a[0] += ms; a[1] += launches; a[2] += nbytes
        for (kind, k), (ms, launches, nbytes) in tot.items():
            if launches:
                kernels[f"{kind}_{k}"] = dict(avg_us=round(1e3 * ms / launches, 3), launches=int(launches),
                                              GBs=round(nbytes / (ms * 1e-3) / 1e9, 1),
                                              frac=round(nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4))
        ms, launches, nbytes = tot[("pressure", "A")]
        if launches:
            achieved = nbytes / (ms * 1e-3) / 1e9
            roof = dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                        frac=round(achieved / PEAK_HBM_GBS, 4), traffic=None,
                        kernel="bicg_A_kernel (pressure system): p update + 7-point SpMV + r0.v",
                        bytes_per_launch=round(nbytes / launches), avg_launch_us=round(1e3 * ms / launches, 3))
    for s in solvers:
        s.close()

    value = args.steps / wall
    lat_sorted = sorted(lat)
    p50 = float(np.percentile(lat_sorted, 50))
    p99 = float(np.percentile(lat_sorted, 99))
    giters = summs[0]["solver_iterations"]
    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline: {args.cpu_iters} outer iterations of the oracle")
        cpu = cpu_baseline(img, params, args, giters)
        log(f"  CPU {cpu['value']:.5f} solves/s ({cpu['sample']})")
    updates = np_ * outer * args.steps / wall
    line = {
        "metric": METRIC, "value": round(value, 5), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * wall / args.steps, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "form": "plain FP64 GPU path",
        "config": dict(desc, n_p=np_, n_u=nu, n_v=nv, n_w=nw, nnz=dict(u=nnz[0], v=nnz[1], w=nnz[2], p=nnz[3]),
                       outer_iterations=outer, bicgstab_iterations=dict(u=giters[0], v=giters[1], w=giters[2],
                                                                         p=giters[3])),
        "latency_s": dict(p50=round(p50, 4), p99=round(p99, 4)),
        "fluid_cell_updates_per_s": round(updates, 1),
        "K": summs[0]["K"],
        "phases_s": {k: round(v, 4) for k, v in summs[0]["phases"].items()},
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
