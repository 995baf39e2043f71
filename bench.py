#!/usr/bin/env python3
"""Permeability solves/s of the SIMPLE-family outer iteration on one MI355X.

Workload (one step = one complete permeability solve): a 128^3 seeded
random-sequential-addition pack of non-overlapping spheres of radius 10 voxels
at porosity 0.75, flow along x, unit pressure drop, mu = rho = 1, voxel size 1,
SIMPLEC to an outer tolerance of 1e-6 on the relative mass residual, inner
tolerances 1e-10 (velocity, pressure) and 1e-6 (pressure correction).

Two forms are timed on the same image: the plain FP64 GPU path and the mixed
form (FP64 defect correction around FP32-storage BiCGSTAB sweeps).  The
headline value is the mixed form's solves/s when its permeability agrees with
the FP64 run's to |dK|/K <= 1e-4, else the FP64 form's.

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


def cpu_baseline(img, params, args, gpu_work):
    """The oracle (CPU restatement, OpenMP SpMV/sweeps, serial vector ops) on
    the same image: a bounded sample of outer iterations, extrapolated to a
    whole solve by the FP64 GPU run's executed BiCGSTAB iteration count (the
    work of one iteration is the same on both sides; zero right-hand sides,
    which Eigen reports as maxit iterations, execute none and count 0)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    o = oracle_lib.Oracle(img, **{k: v for k, v in params.items() if k not in ("direction",)})
    t0 = time.perf_counter()
    rows, summ = o.solve(args.method, args.simpletol, args.cpu_iters, 4, 4)
    dt = time.perf_counter() - t0
    cpu_work = summ["work"]
    per_iter = dt / max(cpu_work, 1)
    est = per_iter * sum(gpu_work)
    threads = oracle_lib.lib().orc_num_threads()
    return dict(value=1.0 / est if est > 0 else None, unit="solves/s", cores=int(threads), kind="port",
                sample=f"{args.cpu_iters} outer iterations of the same {args.method} solve on the same image "
                       f"({cpu_work} executed BiCGSTAB iterations, {dt:.1f} s), extrapolated to the FP64 GPU "
                       f"run's {sum(gpu_work)} executed BiCGSTAB iterations per solve",
                sample_seconds=round(dt, 3))


def measured_traffic(args, kernel):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy
    kernel) recorded with rocprofv3 on this workload, from
    profiles/hbm_traffic_128rsa.json; None for another workload."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_128rsa.json")
    default = (args.size, args.radius, args.porosity, args.seed) == (128, 10.0, 0.75, 20261203)
    if not default or not os.path.exists(path):
        return None, None
    k = json.load(open(path))["kernels"].get(kernel)
    if not k:
        return None, None
    return k["fetch_bytes"] + k["write_bytes"], "profiles/hbm_traffic_128rsa.json (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)"


def kernel_table(solvers):
    """Per-kernel averages over the timed solves from the in-library HIP
    events taken from the dispatches of one sampled BiCGSTAB iteration per
    queued chunk of iterations, on the solver's own stream."""
    tot = {}
    for s in solvers:
        for key, (ms, launches, nbytes) in s.timing_read().items():
            a = tot.setdefault(key, [0.0, 0.0, 0.0])
            a[0] += ms; a[1] += launches; a[2] += nbytes
    kernels = {}
    for (prec, kind, k), (ms, launches, nbytes) in sorted(tot.items()):
        if launches:
            kernels[f"{prec}_{kind}_{k}"] = dict(avg_us=round(1e3 * ms / launches, 3), launches=int(launches),
                                                 GBs=round(nbytes / (ms * 1e-3) / 1e9, 1),
                                                 frac=round(nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4))
    return tot, kernels


def run_form(img, run_kw, args, precision, label):
    """Warmup, then K timed back-to-back solves on pre-created solvers."""
    from ns3d_amd import gpu, solver
    kw = dict(run_kw, precision=precision, inner_tol=args.inner_tol)
    log(f"[{label}] warmup: {args.warmup} solve(s)")
    for w in range(args.warmup):
        s = solver.Solver(img, **kw)
        rows, summ = s.run()
        log(f"  warmup {w}: {summ['iterations']} outer its, K={summ['K']:.10e}, {summ['seconds']:.2f} s")
        s.close()
    log(f"[{label}] uploading {args.steps} image(s)")
    solvers = [solver.Solver(img, **kw) for _ in range(args.steps)]
    if not args.no_timing:
        for s in solvers:
            s.timing(True)
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
    tot, kernels = ({}, {}) if args.no_timing else kernel_table(solvers)
    info = dict(nnz=solvers[0].nnz(), counts=(solvers[0].np, solvers[0].nu, solvers[0].nv, solvers[0].nw))
    for s in solvers:
        s.close()
    lat_sorted = sorted(lat)
    return dict(value=args.steps / wall, wall=wall, p50=float(np.percentile(lat_sorted, 50)),
                p99=float(np.percentile(lat_sorted, 99)), summs=summs, tot=tot, kernels=kernels, **info)


def copy_calibration(mib):
    """A plain copy kernel of known size: under rocprofv3 --pmc its FETCH_SIZE
    and WRITE_SIZE per dispatch give the factor from counter KiB to bytes."""
    import torch
    nbytes = mib << 20
    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    src.fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": nbytes, "bytes_written": nbytes, "copies": 5}), flush=True)


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
    ap.add_argument("--inner-tol", type=float, default=1e-4, help="mixed form: FP32 sweep tolerance")
    ap.add_argument("--forms", default="fp64,mixed")
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="no per-kernel events (profiler runs)")
    ap.add_argument("--progress", action="store_true", help="one stderr line per outer iteration")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (scales HBM byte counters under a profiler)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU runs are not part of "
              f"this build", file=sys.stderr)
        sys.exit(2)

    from ns3d_amd import gpu
    if gpu.device_count() < 1:
        raise SystemExit("no GPU visible")
    log("generating workload")
    img, params, desc = workload(args)
    run_kw = dict(params, method=args.method, simpletol=args.simpletol, maxitSIMPLE=args.maxit, N1=4, N2=4,
                  device=0, verbose=2 if args.progress else 0)
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    res = {}
    for f in forms:
        if f not in ("fp64", "mixed"):
            raise SystemExit(f"unknown form {f}")
        res[f] = run_form(img, run_kw, args, 1 if f == "mixed" else 0, f)

    base = res.get("fp64") or res[forms[0]]
    np_, nu, nv, nw = base["counts"]
    nnz = base["nnz"]
    s0 = base["summs"][0]
    roof = None
    if not args.no_timing and ("f64", "pressure", "A") in base["tot"]:
        ms, launches, nbytes = base["tot"][("f64", "pressure", "A")]
        if launches:
            achieved = nbytes / (ms * 1e-3) / 1e9
            traffic, tsrc = measured_traffic(args, "bicg_A_kernel<double, 2>")
            roof = dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                        frac=round(achieved / PEAK_HBM_GBS, 4), traffic=traffic,
                        kernel="bicg_A_kernel<double, 2> (pressure system): p update + 7-point SpMV + r0.v",
                        bytes_per_launch=round(nbytes / launches), avg_launch_us=round(1e3 * ms / launches, 3),
                        traffic_source=tsrc)

    forms_out = {}
    for f, r in res.items():
        sm = r["summs"][0]
        forms_out[f] = dict(solves_per_s=round(r["value"], 5), ms_per_solve=round(1e3 * r["wall"] / args.steps, 2),
                            latency_s=dict(p50=round(r["p50"], 4), p99=round(r["p99"], 4)), K=sm["K"],
                            outer_iterations=sm["iterations"], converged=sm["converged"],
                            bicgstab_iterations=dict(zip("uvwp", sm["solver_work"])),
                            phases_s={k: round(v, 4) for k, v in sm["phases"].items()})
    headline, form_name, dtype = base, "fp64", "f64"
    if "mixed" in res and "fp64" in res:
        dk = abs(res["mixed"]["summs"][0]["K"] / res["fp64"]["summs"][0]["K"] - 1)
        forms_out["mixed"]["dK_over_K_vs_fp64"] = dk
        forms_out["mixed"]["speedup_vs_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 4)
        if dk <= 1e-4 and res["mixed"]["summs"][0]["converged"]:
            headline, form_name, dtype = res["mixed"], "mixed", "f64 arithmetic, f32 inner-solve storage"
    elif forms == ["mixed"]:
        headline, form_name, dtype = res["mixed"], "mixed", "f64 arithmetic, f32 inner-solve storage"

    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline: {args.cpu_iters} outer iterations of the oracle")
        cpu = cpu_baseline(img, params, args, s0["solver_work"])
        log(f"  CPU {cpu['value']:.5f} solves/s ({cpu['sample']})")
    kernels = {}
    for r in res.values():
        kernels.update(r["kernels"])
    line = {
        "metric": METRIC, "value": round(headline["value"], 5), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * headline["wall"] / args.steps, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "form": form_name,
        "config": dict(desc, n_p=np_, n_u=nu, n_v=nv, n_w=nw, nnz=dict(zip("uvwp", nnz))),
        "latency_s": dict(p50=round(headline["p50"], 4), p99=round(headline["p99"], 4)),
        "forms": forms_out,
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
