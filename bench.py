#!/usr/bin/env python3
"""Permeability solves/s and latency of the SIMPLE-family outer iteration on
one MI355X: the served, batched mixed-precision form against the plain FP64
GPU path on the same requests.

Headline workload (the batching configuration at full size): 48 seeded
images (ns3d_amd.mix: edges per axis from {32, 48, 64, 96, 128, 160},
porosity uniform in [0.35, 0.90], generators split among RSA sphere packs
r U[4,12], Boolean spheres r U[3,15] and fibres r 3 L 40; images whose fluid
does not percolate along x are redrawn), SIMPLEC, flow along x, unit
pressure drop, mu = 1, voxel size 1, rho = 1e-3 (pore Reynolds ~1e-4: the
Stokes limit, where the outer iteration's fixed point does not depend on
rounding), outer tolerance 1e-6 on the relative mass residual (at most 1000
outer iterations), inner tolerances 1e-10 (velocity, pressure) and 1e-6
(pressure correction).

  plain_fp64    every image once, one request at a time on its own context
                (the drop-in path), contexts created before the timed solves;
  served_mixed  the headline: a continuous stream of requests (request r is
                image r mod 48) through the native server -- FP64 defect
                correction around FP32-storage BiCGSTAB, batched on one solve
                engine, requests admitted by projected device footprint and
                a cap on fluid cells in flight, retired and backfilled as they
                converge.  A step is --per-step completed requests; --warmup
                steps are untimed, then --steps steps are timed (from the
                completion ending the warm-up to the one ending the last
                step).  Each request's upload, index scan and teardown are
                inside (that is the service).

speedup_vs_plain_fp64 compares the stream's rate with the plain path's rate
on exactly the requests the timed region completed.  dK_gate compares every
completed mixed request with the plain FP64 K of its image at |dK|/K <= 1e-4.

Second section ("single"): one 128^3 RSA pack (r = 10, porosity 0.75), plain
FP64 and mixed, with per-kernel timings from dispatch events -- the roofline
of the dominant FP64 kernel.  The CPU baseline is the oracle on the box's
host cores over a bounded sample of the mix.

Sections outside the default run: "precision" (the 384^3 Boolean image run
to convergence in the mixed form, Stokes limit and Forchheimer points, each
against plain FP64 at a horizon), "serving" (a bounded sample of the serving
workload's recipe at saturating and half load) and "cpu" (the CPU baseline's
whole runs).

Output: one JSON line; progress on stderr.
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
TOLS = dict(vsolvertol=1e-10, psolvertol=1e-10, pcsolvertol=1e-6, maxitsolver=5000)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pct(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else None


# ---------------------------------------------------------------- the mix
def mix_params(img, args, precision):
    nx = img.shape[0]
    return dict(dx=1.0, mu=1.0, rho=args.rho, p0=1.0, gradp=1.0 / nx, alpha=0.8, beta=1.0, direction=1, border=0,
                method="SIMPLEC", simpletol=args.simpletol, maxitSIMPLE=args.mix_maxit, N1=4, N2=4, device=0,
                precision=precision, inner_tol=args.inner_tol, mixed_min_rows=args.mixed_min_rows,
                mg_min_rows=args.mg_min_rows, **TOLS)


def make_server(args):
    from ns3d_amd import solver
    srv = solver.Server(device=0, budget_bytes=args.budget_gb * 1e9, max_slots=args.slots, chunk=args.chunk,
                        fallback=solver.FALLBACK_FAILED)
    srv.set_schedule(args.chunk, args.lanes, not args.no_graphs)
    srv.set_policy(largest_first=False, max_rows=args.max_rows)
    return srv


def stream(reqs, args, precision, label, warm_steps, steps):
    """A continuous stream of permeability requests through one native
    server: request r is image r mod len(reqs); the server is kept fed with
    args.backlog queued requests beyond the ones it runs, so requests join
    the running batch as others retire.  A step is args.per_step completed
    requests; the first warm_steps steps are untimed, the next `steps` are
    the timed region (from the completion that ends the warm-up to the one
    that ends the last step).  Requests still in flight after that finish
    untimed."""
    srv = make_server(args)
    need = (warm_steps + steps) * args.per_step
    done, pending = [], {}
    nsub = 0
    try:
        t0 = time.perf_counter()
        while len(done) < need:
            st = srv.stats()
            while st["queued"] < args.backlog and len(done) + len(pending) < need + args.backlog:
                img = reqs[nsub % len(reqs)][0]
                pending[srv.submit(img, **mix_params(img, args, precision))] = nsub
                nsub += 1
                st = srv.stats()
            for t in list(pending):
                r = srv.poll(t)
                if r is not None:
                    done.append((pending.pop(t), r, time.perf_counter() - t0))
                    if len(done) % args.per_step == 0:
                        k = len(done) // args.per_step
                        log(f"  [{label}] step {k - warm_steps if k > warm_steps else 'warm-up ' + str(k)} done "
                            f"at {done[-1][2]:.1f} s")
            time.sleep(0.002)
        est = srv.engine_stats()
        sst = srv.stats()
    finally:
        srv.close()  # finishes what is still in flight (untimed)
    w = warm_steps * args.per_step
    t_start = done[w - 1][2] if w > 0 else 0.0
    t_end = done[need - 1][2]
    timed = done[w:need]
    lat = [r["t_end"] - r["t_submit"] for _, r, _ in timed]
    wall = t_end - t_start
    log(f"  [{label}] {len(timed)} requests timed in {wall:.2f} s ({len(timed) / wall:.3f} solves/s), latency p50 "
        f"{pct(lat, 50):.2f} s p99 {pct(lat, 99):.2f} s, peak running {sst['peak_running']:.0f}, "
        f"fallbacks {sst['fallbacks']:.0f}, submitted {nsub}")
    return dict(wall=wall, timed=timed, all=done, latency=lat, stats=sst, engine=est, submitted=nsub)


def plain_pass(reqs, args):
    """The drop-in path: contexts created up front, then one solve at a time;
    per-request seconds and results."""
    from ns3d_amd import solver
    solvers = [solver.Solver(img, **mix_params(img, args, 0)) for img, _ in reqs]
    out, secs = [], []
    t0 = time.perf_counter()
    for q, s in enumerate(solvers):
        t1 = time.perf_counter()
        out.append(s.run()[1])
        secs.append(time.perf_counter() - t1)
        log(f"    plain {q}: {reqs[q][1]['shape']} {out[-1]['iterations']} outer its, {secs[-1]:.2f} s")
    wall = time.perf_counter() - t0
    for s in solvers:
        s.close()
    log(f"  [plain_fp64] {len(reqs)} requests in {wall:.2f} s ({len(reqs) / wall:.3f} solves/s), "
        f"{sum(int(r['converged']) for r in out)} converged")
    return dict(wall=wall, results=out, seconds=secs)


def engine_roofline(est, prec="f32", k="A"):
    ms, launches, nbytes = est[f"{prec}_{k}"]
    if not launches or ms <= 0:
        return None
    ach = nbytes / (ms * 1e-3) / 1e9
    return dict(avg_launch_us=round(1e3 * ms / launches, 2), sampled_launches=int(launches),
                bytes_per_launch=round(nbytes / launches), achieved=round(ach, 1), frac=round(ach / PEAK_HBM_GBS, 4))


def dk_gate(completed, plain, tol=1e-4):
    """|dK|/K of every completed mixed request against the plain FP64 run of
    its image, gated at tol.  A request that converges in FP64 but not in the
    mixed form fails the gate; one that stalls at the outer-iteration cap in
    both is listed (its K is not a fixed point to compare)."""
    out = dict(tolerance=tol, requests=0, within=0, over=[], unconverged_mixed=[], stalled_both=[])
    for r, m, _ in completed:
        q = r % len(plain)
        p = plain[q]
        if not p["converged"]:
            if not m["converged"]:
                out["stalled_both"].append(q)
            continue
        out["requests"] += 1
        if not m["converged"]:
            out["unconverged_mixed"].append(q)
            continue
        d = abs(m["K"] / p["K"] - 1)
        if d <= tol:
            out["within"] += 1
        else:
            out["over"].append(dict(request=r, image=q, dK_over_K=d))
    out["stalled_both"] = sorted(set(out["stalled_both"]))
    out["passed"] = out["within"] == out["requests"]
    return out


def engine_chunk(est):
    """Whole engine chunks (all solves in flight, all engine streams, from
    the first A start to the last C stop, launch gaps included): the
    algorithmic bytes of the iterations executed over the span."""
    ms, n, nbytes = est.get("chunk", (0, 0, 0))
    if not n or ms <= 0:
        return None
    ach = nbytes / (ms * 1e-3) / 1e9
    return dict(avg_span_us=round(1e3 * ms / n, 2), sampled=int(n), bytes_per_chunk=round(nbytes / n),
                achieved=round(ach, 1), frac=round(ach / PEAK_HBM_GBS, 4))


def mix_section(args):
    from ns3d_amd import mix
    edges = tuple(int(e) for e in args.edges.split(","))
    log(f"generating the request mix ({args.requests} requests, edges {args.edges}, seed {args.mix_seed})")
    reqs = mix.request_mix(args.requests, args.mix_seed, edges)
    fluid = sum(d["fluid"] for _, d in reqs)
    if args.mix_forms == "stream":
        # profiler runs: the timed stream only
        stream(reqs, args, 1, "served_mixed", args.warmup, args.steps)
        return None
    tuned = None
    if args.autotune:
        from ns3d_amd import autotune
        # a held-out sample of the same recipe (another seed), each request
        # capped at tune_maxit outer iterations
        tune_seed = args.mix_seed + 7919
        log(f"autotuning the engine schedule on {args.tune_requests} held-out requests (seed {tune_seed}), "
            f"each capped at {args.tune_maxit} outer iterations")
        sample = []
        for img, _ in mix.request_mix(args.tune_requests, tune_seed, edges):
            p = mix_params(img, args, 1)
            p["maxitSIMPLE"] = args.tune_maxit
            sample.append((img, p))
        tuned = autotune.tune(sample, candidates=[(c, l, True) for c in (4, 8) for l in (2, 3)] + [(8, 3, False)],
                              log=log, budget_gb=args.budget_gb, slots=args.slots)
        args.chunk, args.lanes, args.no_graphs = tuned["chunk"], tuned["lanes"], not tuned["graphs"]
        log(f"  chosen: chunk {args.chunk}, lanes {args.lanes}, graphs {int(not args.no_graphs)}")
    log("plain FP64 path, every request of the mix once")
    plain = plain_pass(reqs, args)
    log(f"served mixed stream: {args.warmup} warm-up and {args.steps} timed steps of {args.per_step} completed "
        f"requests")
    smx = stream(reqs, args, 1, "served_mixed", args.warmup, args.steps)
    n = len(smx["timed"])
    value = n / smx["wall"]
    # the plain path's rate on exactly the requests the timed region completed
    plain_secs = sum(plain["seconds"][r % len(reqs)] for r, _, _ in smx["timed"])
    plain_rate = n / plain_secs
    gate = dk_gate(smx["all"], plain["results"])
    dks = [abs(m["K"] / plain["results"][r % len(reqs)]["K"] - 1) for r, m, _ in smx["all"]
           if m["converged"] and plain["results"][r % len(reqs)]["converged"]]
    # fluid-cell updates: pressure rows x outer iterations of the completed solves, per second
    fcu_plain = sum(float(r.get("np", 0)) * r["iterations"] for r in plain["results"]) / plain["wall"]
    fcu_served = sum(float(m.get("np", 0)) * m["iterations"] for _, m, _ in smx["timed"]) / smx["wall"]
    forms = {
        "plain_fp64": dict(solves_per_s=round(len(reqs) / plain["wall"], 4), ms_per_pass=round(1e3 * plain["wall"], 1),
                           fluid_cell_updates_per_s=round(fcu_plain, 1),
                           solves_per_s_on_timed_requests=round(plain_rate, 4),
                           converged=int(sum(r["converged"] for r in plain["results"])),
                           outer_iterations=int(sum(r["iterations"] for r in plain["results"])),
                           latency_s=dict(p50=round(pct(plain["seconds"], 50), 3),
                                          p99=round(pct(plain["seconds"], 99), 3))),
        "served_mixed": dict(solves_per_s=round(value, 4), timed_requests=n,
                             fluid_cell_updates_per_s=round(fcu_served, 1),
                             latency_s=dict(p50=round(pct(smx["latency"], 50), 3), p99=round(pct(smx["latency"], 99), 3)),
                             max_dK_over_K=max(dks) if dks else None,
                             median_dK_over_K=float(np.median(dks)) if dks else None,
                             fallbacks=int(sum(m["fallback"] for _, m, _ in smx["all"])),
                             peak_running=smx["stats"]["peak_running"],
                             peak_device_bytes=smx["stats"]["peak_bytes_in_use"], submitted=smx["submitted"],
                             engine_A_f32=engine_roofline(smx["engine"], "f32", "A")),
    }
    if args.poisson:
        # latency under load: Poisson arrivals at a fraction of the stream's rate
        rate = args.poisson_frac * value
        rng = np.random.default_rng(args.mix_seed)
        arrivals = np.concatenate([[0.0], np.cumsum(rng.exponential(1.0 / rate, size=len(reqs) - 1))])
        srv = make_server(args)
        try:
            t0 = time.perf_counter()
            tickets = []
            for q, (img, _) in enumerate(reqs):
                dt = arrivals[q] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                tickets.append(srv.submit(img, **mix_params(img, args, 1)))
            res = [srv.wait(t) for t in tickets]
            pwall = time.perf_counter() - t0
        finally:
            srv.close()
        plat = [r["t_end"] - r["t_submit"] for r in res]
        log(f"  [poisson {rate:.3f}/s] latency p50 {pct(plat, 50):.2f} s p99 {pct(plat, 99):.2f} s")
        forms["served_mixed_poisson"] = dict(arrival_rate=round(rate, 4), solves_per_s=round(len(reqs) / pwall, 4),
                                             latency_s=dict(p50=round(pct(plat, 50), 3), p99=round(pct(plat, 99), 3)))
    desc = dict(workload=f"batching configuration: stream of {len(reqs)} seeded images (edges {args.edges} per axis, "
                         f"porosity U[0.35,0.90], RSA r U[4,12] / Boolean r U[3,15] / fibres r 3 L 40), SIMPLEC, "
                         f"flow x, Stokes limit (mu 1, rho {args.rho:g}, unit pressure drop), simpletol "
                         f"{args.simpletol:g}, at most {args.mix_maxit} outer iterations; a step = "
                         f"{args.per_step} completed requests",
                requests=len(reqs), fluid_cells=int(fluid), mix_seed=args.mix_seed, slots=args.slots,
                backlog=args.backlog, per_step=args.per_step,
                schedule=dict(chunk=args.chunk, lanes=args.lanes, graphs=not args.no_graphs,
                              autotuned=tuned is not None, max_rows_in_flight=args.max_rows),
                images=[dict(shape=d["shape"], generator=d["generator"], porosity=d["porosity"]) for _, d in reqs])
    return dict(forms=forms, desc=desc, reqs=reqs, plain=plain, value=value, plain_rate=plain_rate,
                ms_per_step=1e3 * smx["wall"] / args.steps, gate=gate, engine=smx["engine"], autotune=tuned,
                max_dk=max(dks) if dks else None)


# ---------------------------------------------------------------- single image
def single_workload(args):
    from ns3d_amd import generators
    n = args.size
    params = dict(dx=1.0, mu=1.0, rho=args.rho, p0=1.0, gradp=1.0 / n, alpha=0.8, beta=1.0, direction=1,
                  border=0, **TOLS)
    if getattr(args, "generator", "rsa") == "boolean":
        # the precision configuration's image: overlapping spheres, radii U[rmin, rmax]
        img = generators.boolean_spheres((n, n, n), args.rmin, args.rmax, args.porosity, args.seed)[0]
        por = 1.0 - float(img.mean())
        desc = dict(workload=f"{n}^3 Boolean spheres r=U[{args.rmin:g},{args.rmax:g}] porosity {por:.4f}, SIMPLEC, "
                             f"flow x", image=dict(shape=[n, n, n], rmin=args.rmin, rmax=args.rmax, seed=args.seed,
                                                   porosity=round(por, 6)))
        return img, params, desc
    img, solid, overlap, nsph = generators.rsa_spheres((n, n, n), args.radius, 1.0 - args.porosity, args.seed)
    desc = dict(workload=f"{n}^3 RSA sphere pack r={args.radius:g} porosity {1 - solid:.4f}, SIMPLEC, flow x",
                image=dict(shape=[n, n, n], radius=args.radius, seed=args.seed, porosity=round(1 - solid, 6),
                           rsa_overlap_voxels=overlap, spheres=nsph))
    return img, params, desc


def kernel_table(solvers):
    """Per-kernel averages from the in-library events taken from the
    dispatches of one sampled BiCGSTAB iteration per queued chunk."""
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


def single_form(img, params, args, precision, label, steps, warmup):
    from ns3d_amd import gpu, solver
    kw = dict(params, method="SIMPLEC", simpletol=args.simpletol, maxitSIMPLE=args.maxit, N1=4, N2=4, device=0,
              precision=precision, inner_tol=args.inner_tol, mg_min_rows=args.mg_min_rows,
              verbose=2 if args.progress else 0)
    for w in range(warmup):
        s = solver.Solver(img, **kw)
        summ = s.run()[1]
        log(f"  [{label}] warmup: {summ['iterations']} outer its, {summ['seconds']:.2f} s")
        s.close()
    solvers = [solver.Solver(img, **kw) for _ in range(steps)]
    if not args.no_timing:
        for s in solvers:
            s.timing(True)
    for s in solvers:
        gpu._check(gpu.lib().ns3d_sync(s.context()))
    summs, lat = [], []
    t0 = time.perf_counter()
    for s in solvers:
        t1 = time.perf_counter()
        summs.append(s.run()[1])
        lat.append(time.perf_counter() - t1)
        log(f"  [{label}] {summs[-1]['iterations']} outer its, K={summs[-1]['K']:.10e}, {lat[-1]:.2f} s")
    wall = time.perf_counter() - t0
    tot, kernels = ({}, {}) if args.no_timing else kernel_table(solvers)
    if not args.no_timing:
        # the multigrid sweeps' finest-level kernels (mixed form)
        for s in solvers:
            for name, (ms, launches, nbytes) in s.multigrid_timing().items():
                if launches:
                    kernels[f"mg_{name}"] = dict(avg_us=round(1e3 * ms / launches, 3), launches=int(launches),
                                                 GBs=round(nbytes / (ms * 1e-3) / 1e9, 1),
                                                 frac=round(nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                                                 bytes_per_launch=round(nbytes / launches))
    info = dict(nnz=solvers[0].nnz(), counts=(solvers[0].np, solvers[0].nu, solvers[0].nv, solvers[0].nw))
    for s in solvers:
        s.close()
    return dict(value=steps / wall, wall=wall, lat=lat, summs=summs, tot=tot, kernels=kernels, **info)


def measured_traffic(args, kernel):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy
    kernel) recorded with rocprofv3 on the 128^3 image, from
    profiles/hbm_traffic_128rsa.json; None for another image."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_128rsa.json")
    default = (getattr(args, "generator", "rsa"), args.size, args.radius, args.porosity, args.seed) == \
        ("rsa", 128, 10.0, 0.75, 20261203)
    if not default or not os.path.exists(path):
        return None, None
    k = json.load(open(path))["kernels"].get(kernel)
    if not k:
        return None, None
    return k["fetch_bytes"] + k["write_bytes"], "profiles/hbm_traffic_128rsa.json (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)"


def mix_traffic(args):
    """HBM bytes per launch of the batched A kernel in the served mixed pass
    (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy kernel, averaged over
    its working launches), from profiles/hbm_traffic_mix.json; None for
    another mix."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_mix.json")
    if not os.path.exists(path):
        return None, None
    d = json.load(open(path))
    if (d.get("requests"), d.get("mix_seed"), d.get("edges"), d.get("mix_maxit")) != \
            (args.requests, args.mix_seed, args.edges, args.mix_maxit):
        return None, None
    k = d["kernels"].get("bicg_A_kernel<float>")
    if not k:
        return None, None
    return k["fetch_bytes"] + k["write_bytes"], "profiles/hbm_traffic_mix.json (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)"


def single_section(args):
    img, params, desc = single_workload(args)
    log(f"single image: {desc['workload']}")
    f64 = single_form(img, params, args, 0, "single fp64", args.single_steps, args.single_warmup)
    mx = single_form(img, params, args, 1, "single mixed", args.single_steps, 0)
    roof = None
    if ("f64", "pressure", "A") in f64["tot"]:
        ms, launches, nbytes = f64["tot"][("f64", "pressure", "A")]
        if launches:
            ach = nbytes / (ms * 1e-3) / 1e9
            traffic, tsrc = measured_traffic(args, "bicg_A_kernel<double>")
            roof = dict(bound="hbm", achieved=round(ach, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                        frac=round(ach / PEAK_HBM_GBS, 4), traffic=traffic,
                        kernel="bicg_A_kernel<double> (pressure system, packed symmetric storage): p update + "
                               f"7-point SpMV + r0.v, {args.size}^3 image, plain FP64 path",
                        bytes_per_launch=round(nbytes / launches), avg_launch_us=round(1e3 * ms / launches, 3),
                        traffic_source=tsrc)
    kernels = dict(f64["kernels"])
    kernels.update(mx["kernels"])
    s0 = f64["summs"][0]
    out = dict(desc, n_p=f64["counts"][0], nnz=dict(zip("uvwp", f64["nnz"])),
               fp64=dict(solves_per_s=round(f64["value"], 5), ms_per_solve=round(1e3 * f64["wall"] / len(f64["lat"]), 1),
                         K=s0["K"], outer_iterations=s0["iterations"],
                         bicgstab_iterations=dict(zip("uvwp", s0["solver_work"])),
                         phases_s={k: round(v, 4) for k, v in s0["phases"].items()}),
               mixed=dict(solves_per_s=round(mx["value"], 5), ms_per_solve=round(1e3 * mx["wall"] / len(mx["lat"]), 1),
                          K=mx["summs"][0]["K"], dK_over_K_vs_fp64=abs(mx["summs"][0]["K"] / s0["K"] - 1),
                          outer_iterations=mx["summs"][0]["iterations"]),
               fluid_cell_updates_per_s_fp64=round(f64["counts"][0] * s0["iterations"] / (f64["wall"] / len(f64["lat"])), 1),
               fluid_cell_updates_per_s_mixed=round(mx["counts"][0] * mx["summs"][0]["iterations"] /
                                                    (mx["wall"] / len(mx["lat"])), 1))
    return dict(out=out, roof=roof, kernels=kernels, img=img, params=params, work=s0["solver_work"])


# ---------------------------------------------------------------- CPU baseline
def cpu_baseline_mix(mixres, args):
    """The oracle (CPU restatement, OpenMP SpMV and sweeps) on the largest
    request of the mix for a bounded number of outer iterations; its seconds per row x
    executed BiCGSTAB iteration, times the plain FP64 GPU run's rows x
    executed iterations summed over the mix, gives the CPU time of the whole
    mix."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    reqs, plain = mixres["reqs"], mixres["plain"]["results"]
    q = max(range(len(reqs)), key=lambda k: reqs[k][1]["fluid"])
    img = reqs[q][0]
    p = mix_params(img, args, 0)
    o = oracle_lib.Oracle(img, **{k: p[k] for k in ("dx", "mu", "rho", "p0", "gradp", "alpha", "beta",
                                                       "vsolvertol", "psolvertol", "pcsolvertol", "maxitsolver")})
    t0 = time.perf_counter()
    summ = o.solve("SIMPLEC", args.simpletol, args.cpu_iters, 4, 4)[1]
    dt = time.perf_counter() - t0
    rate = dt / max(1.0, float(o.np) * summ["work"])
    units = sum(r["np"] * w for r in plain for w in [sum(r["solver_work"])])
    est = rate * units
    threads = oracle_lib.lib().orc_num_threads()
    return dict(value=round(len(reqs) / est, 6) if est > 0 else None, unit="solves/s", cores=int(threads),
                kind="port",
                sample=f"{args.cpu_iters} outer iterations of request {q} ({reqs[q][1]['shape']}, "
                       f"{summ['work']} executed BiCGSTAB iterations, {dt:.1f} s), scaled per row x executed "
                       f"iteration to the plain FP64 GPU run's {units:.3e} rows x iterations over the mix",
                sample_seconds=round(dt, 3))


def cpu_runs(args):
    """The CPU-baseline plan's whole runs, each beside the GPU paths on the
    same case: (1) the parity configuration's 64^3 RSA pack (r = 8, solid
    fraction 0.40 target) to convergence at simpletol 1e-6 with the tight
    inner tolerances, the oracle on every host thread; (2) the reference's
    own timed case, tests/compare_algorithms.ini on the 64^3 sphere lattice
    with SIMPLEC (README.md:177-183: 5.435 s, 48 outer iterations, K
    3.28386e-12 on 4 threads), the oracle on 4 threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import oracle_lib
    import make_config_parity as MC
    from ns3d_amd import generators, solver
    out = {}
    host_threads = oracle_lib.lib().orc_num_threads()
    raw, rec = MC.image("c1")
    prm = dict(MC.params(raw.shape[0]))
    c1 = dict(case="64^3 RSA r=8 porosity %.4f, SIMPLEC, flow x, simpletol 1e-6, inner 1e-10/1e-10/1e-6"
              % rec["achieved_porosity"])
    for label, kw in (("gpu_fp64", dict(precision=0)), ("gpu_mixed", dict(precision=1))):
        t0 = time.perf_counter()
        s = solver.solve(raw, method="SIMPLEC", simpletol=1e-6, maxitSIMPLE=3000, N1=4, N2=4, device=0, **prm,
                         **kw)[1]
        c1[label] = dict(seconds=round(time.perf_counter() - t0, 3), iterations=s["iterations"], K=s["K"],
                         converged=s["converged"])
        log(f"  C1 {label}: {c1[label]}")
    o = oracle_lib.Oracle(raw, **prm)
    s = o.solve("SIMPLEC", 1e-6, 3000, 4, 4)[1]
    c1["cpu_oracle"] = dict(seconds=round(s["seconds"], 3), iterations=s["iterations"], K=s["K"], threads=host_threads)
    log(f"  C1 cpu: {c1['cpu_oracle']}")
    out["c1_converged"] = c1
    lat = generators.sphere_lattice(64, 8, 3)
    ini = dict(dx=2e-6, mu=1e-3, rho=1e3, p0=1e5, gradp=1e2, alpha=0.8, beta=1.0, vsolvertol=1e-5, psolvertol=1e-6,
               pcsolvertol=1e-3, deltaRMR=1e-4, maxitsolver=500)
    rd = dict(case="tests/compare_algorithms.ini, 64^3 sphere lattice, SIMPLEC (README.md:177-183)",
              readme=dict(seconds=5.435, iterations=48, K=3.28386e-12, threads=4))
    oracle_lib.lib().orc_set_threads(4)
    o = oracle_lib.Oracle(lat, **ini)
    s = o.solve("SIMPLEC", 1e-4, 500, 2, 2)[1]
    rd["cpu_oracle_4_threads"] = dict(seconds=round(s["seconds"], 3), iterations=s["iterations"], K=s["K"])
    oracle_lib.lib().orc_set_threads(host_threads)
    t0 = time.perf_counter()
    g = solver.solve(lat, method="SIMPLEC", simpletol=1e-4, maxitSIMPLE=500, N1=2, N2=2, device=0, **ini)[1]
    rd["gpu_fp64"] = dict(seconds=round(time.perf_counter() - t0, 3), iterations=g["iterations"], K=g["K"])
    log(f"  README case: {rd}")
    out["readme_64_simplec"] = rd
    return out


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def field_errors(mixed, fp64):
    return {k: rel_l2(mixed[k], fp64[k]) for k in "puvw"}


def connected_p_rows(raw, direction=1):
    """Mask over the pressure rows (C order of the prepared image) of the
    fluid 6-connected to the inlet or outlet plane."""
    import scipy.ndimage as ndi
    from ns3d_amd import gpu
    img = gpu.prepare_image(raw, direction, 0)
    lab, _ = ndi.label(img == 0)
    keep = np.setdiff1d(np.union1d(np.unique(lab[0]), np.unique(lab[-1])), [0])
    with gpu.NavierStokesGPU(img) as g:
        cells = g.index()[0] >= 0
    return np.isin(lab[cells], keep)


def precision_runs(args):
    """The precision configuration run to convergence in the mixed form: one
    384^3 Boolean image (radii U[6, 20], porosity 0.70, flow along x).

    Run 1, Stokes limit (rho = --rho): the mixed form to simpletol.  The
    plain FP64 path runs to a horizon of --prec-horizon outer iterations (its
    converged solve of this image takes hours) and is compared with the mixed
    form at the same iteration: |dK|/K from the RES rows, the fields p, u, v,
    w by relative L2 against a mixed run capped at the horizon.

    Run 2, Forchheimer points: rho = 1 and the gradients that give pore
    Reynolds numbers of about --forch-re under Darcy's law with run 1's K
    (Re = rho (v / phi) d / mu, v the Darcy velocity, d = rmin + rmax the
    mean sphere diameter).  The mixed form runs every point to simpletol,
    all on one engine; plain FP64 and mixed capped at --forch-horizon are
    compared as in run 1; -grad p = mu / k1 v + rho / k2 v^2 is fitted to
    the converged points.  Field errors: p, u, v, w over their node sets, and
    p over the fluid connected to the inlet or outlet (p_connected)."""
    from ns3d_amd import forchheimer, generators, solver
    n = args.size
    img = generators.boolean_spheres((n, n, n), args.rmin, args.rmax, args.porosity, args.seed)[0]
    phi = 1.0 - float(img.mean())
    d = args.rmin + args.rmax
    log(f"precision configuration: {n}^3 Boolean spheres r=U[{args.rmin:g},{args.rmax:g}], porosity {phi:.4f}")
    base = dict(dx=1.0, mu=1.0, p0=1.0, gradp=1.0 / n, alpha=0.8, beta=1.0, direction=1, border=0, method="SIMPLEC",
                simpletol=args.simpletol, N1=4, N2=4, device=0, inner_tol=args.inner_tol,
                mg_min_rows=args.mg_min_rows, verbose=2 if args.progress else 0, **TOLS)
    out = dict(image=dict(shape=[n, n, n], generator="boolean", rmin=args.rmin, rmax=args.rmax, seed=args.seed,
                          porosity=round(phi, 6)),
               simpletol=args.simpletol, inner_tol=args.inner_tol)
    parts = set(args.prec_parts.split(","))

    conn = []

    def horizon_pair(params, h, label):
        rows64, s64 = solver.solve(img, **dict(params, precision=0, maxitSIMPLE=h))
        log(f"  [{label} fp64 x{h}] K {s64['K']:.10e}, {s64['seconds']:.1f} s")
        rowsm, sm = solver.solve(img, **dict(params, precision=1, maxitSIMPLE=h))
        log(f"  [{label} mixed x{h}] K {sm['K']:.10e}, {sm['seconds']:.1f} s")
        err = field_errors(sm["fields"], s64["fields"])
        if not conn:
            conn.append(connected_p_rows(img))
        # pressure in pockets cut off from the inlet and outlet is fixed only up to a constant
        err["p_connected"] = rel_l2(sm["fields"]["p"][conn[0]], s64["fields"]["p"][conn[0]])
        return dict(outer_iterations=h, K_fp64=s64["K"], K_mixed=sm["K"], dK_over_K=abs(sm["K"] / s64["K"] - 1),
                    field_rel_l2=err,
                    seconds_fp64=round(s64["seconds"], 2), seconds_mixed=round(sm["seconds"], 2),
                    pressure_iterations_fp64=s64["solver_work"][3], pressure_iterations_mixed=sm["solver_work"][3])

    k_stokes = None
    if "stokes" in parts:
        prm = dict(base, rho=args.rho, maxitSIMPLE=args.maxit)
        s = solver.Solver(img, **dict(prm, precision=1))
        try:
            rows, sm = s.run()
        finally:
            s.close()
        log(f"  [stokes mixed] {sm['iterations']} outer its, converged {sm['converged']}, K {sm['K']:.10e}, "
            f"{sm['seconds']:.1f} s")
        k_stokes = sm["K"]
        run1 = dict(rho=args.rho, gradp=1.0 / n, mixed=dict(
            converged=int(sm["converged"]), outer_iterations=int(sm["iterations"]), K=sm["K"],
            seconds=round(sm["seconds"], 2), RMR=float(rows[-1, 3]), solver_work=sm["solver_work"],
            pore_Re=args.rho * (sm["v_moy"] / phi) * d / 1.0))
        h = min(args.prec_horizon, int(sm["iterations"]))
        run1["horizon"] = horizon_pair(prm, h, "stokes")
        # the converged mixed K against its own value at the horizon: how far the horizon is from the end
        run1["horizon"]["K_mixed_from_rows"] = float(rows[h - 1, 4])
        out["stokes"] = run1
    elif "stokes_horizon" in parts:
        out["stokes"] = dict(rho=args.rho, gradp=1.0 / n,
                             horizon=horizon_pair(dict(base, rho=args.rho), args.prec_horizon, "stokes"))
    if "forch" in parts:
        if k_stokes is None:
            k_stokes = args.forch_k
        if not k_stokes:
            raise SystemExit("the Forchheimer points need run 1's K (--prec-parts stokes,forch or --forch-k)")
        res = [float(x) for x in args.forch_re.split(",")]
        gradps = [r * phi / (k_stokes * d) for r in res]  # mu = rho = 1
        prm = dict(base, rho=1.0, maxitSIMPLE=args.maxit, precision=1)
        points = []
        for r, g in zip(res, gradps):
            s = solver.Solver(img, **dict(prm, gradp=g))
            try:
                sm = s.run()[1]
            finally:
                s.close()
            v = float(sm["v_moy"])
            points.append(dict(target_pore_Re=r, gradp=g, v_moy=v, K=float(sm["K"]), K_apparent=v / g,
                               pore_Re=v / phi * d, converged=int(sm["converged"]), iterations=int(sm["iterations"]),
                               seconds=round(sm["seconds"], 2)))
            p = points[-1]
            log(f"  [forchheimer Re~{r:g}] gradp {g:.4e}: {p['iterations']} outer its, converged {p['converged']}, "
                f"K_app {p['K_apparent']:.6e}, pore Re {p['pore_Re']:.2f}, {p['seconds']:.1f} s")
        good = [p for p in points if p["converged"] and p["v_moy"] > 0]
        fit = forchheimer.fit([p["gradp"] for p in good], [p["v_moy"] for p in good], 1.0, 1.0) \
            if len(good) >= 2 else None
        run2 = dict(rho=1.0, K_stokes=k_stokes, mean_diameter=d, points=points, fit=fit, horizons=[])
        for r, g in zip(res, gradps):
            hp = horizon_pair(dict(base, rho=1.0, gradp=g), args.forch_horizon, f"forchheimer Re~{r:g}")
            hp["target_pore_Re"] = r
            run2["horizons"].append(hp)
        out["forchheimer"] = run2
    return out


def serving_runs(args):
    """The serving workload on a bounded sample: the first --serve-requests
    requests of the serving recipe (ns3d_amd.mix.serving_mix: cubic edge
    log-uniform in [48, 512], Boolean spheres r U[6, 20], porosity U[0.30,
    0.90], random flow direction and method) whose edge is at most
    --serve-max-edge, Stokes limit (rho --rho), simpletol --simpletol.

    (1) Each request alone on the plain FP64 path: its K and seconds.
    (2) The mixed form served with FP64 fallback on non-convergence
    (FALLBACK_UNCONVERGED: an unconverged or failed mixed run is continued
    in FP64 from its fields): all requests arriving at once (saturating
    load), then Poisson arrivals at the fractions --serve-loads of that
    burst's rate.  Each load reports solves/s (completed requests over the
    time from the first arrival to the last completion), p50/p99 latency
    (arrival to result) and |dK|/K of every result against its FP64 K,
    gated at 1e-4 like the headline."""
    from ns3d_amd import mix, solver
    reqs = mix.serving_mix(args.serve_requests, args.mix_seed, max_edge=args.serve_max_edge)
    skipped = reqs[-1][1]["skipped_before"]
    log(f"serving sample: {len(reqs)} requests of edge <= {args.serve_max_edge} ({skipped} larger ones of the "
        f"recipe skipped), {sum(d['fluid'] for _, d in reqs)} fluid cells")

    def params(img, d, precision):
        p = mix_params(img, args, precision)
        # the pressure drop is unit along the request's own flow axis
        p.update(method=d["method"], direction=d["direction"], gradp=1.0 / img.shape[d["direction"] - 1])
        return p

    plain = []
    for q, (img, d) in enumerate(reqs):
        s = solver.Solver(img, **params(img, d, 0))
        try:
            t1 = time.perf_counter()
            sm = s.run()[1]
            sec = time.perf_counter() - t1
        finally:
            s.close()
        plain.append(dict(K=sm["K"], converged=int(sm["converged"]), iterations=int(sm["iterations"]), seconds=sec))
        log(f"    plain {q}: {d['shape'][0]}^3 {d['method']} dir {d['direction']} {sm['iterations']} outer its, "
            f"{sec:.2f} s")

    def serve(arrivals, label):
        srv = solver.Server(device=0, budget_bytes=args.budget_gb * 1e9, max_slots=args.slots, chunk=args.chunk,
                            fallback=solver.FALLBACK_UNCONVERGED)
        try:
            srv.set_schedule(args.chunk, args.lanes, not args.no_graphs)
            srv.set_policy(largest_first=False, max_rows=args.max_rows)
            t0 = time.perf_counter()
            tickets = []
            for q, (img, d) in enumerate(reqs):
                dt = arrivals[q] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                tickets.append(srv.submit(img, **params(img, d, 1)))
            res = [srv.wait(t) for t in tickets]
            wall = time.perf_counter() - t0
            st = srv.stats()
        finally:
            srv.close()
        lat = [r["t_end"] - r["t_submit"] for r in res]
        # the headline's strict gate, request q against its own plain FP64 run
        gate = dk_gate([(q, r, 0.0) for q, r in enumerate(res)], plain)
        dks = [abs(r["K"] / p["K"] - 1) for r, p in zip(res, plain) if r["converged"] and p["converged"]]
        out = dict(solves_per_s=round(len(res) / wall, 4), wall_s=round(wall, 2),
                   latency_s=dict(p50=round(pct(lat, 50), 3), p99=round(pct(lat, 99), 3)), gate=gate,
                   max_dK_over_K=max(dks) if dks else None, fallbacks=int(st["fallbacks"]),
                   peak_running=int(st["peak_running"]))
        log(f"  [{label}] {len(res)} requests in {wall:.2f} s ({out['solves_per_s']} solves/s), latency p50 "
            f"{out['latency_s']['p50']} s p99 {out['latency_s']['p99']} s, max dK/K {out['max_dK_over_K']}, "
            f"fallbacks {out['fallbacks']}")
        return out

    loads = {}
    loads["burst"] = serve([0.0] * len(reqs), "burst")
    rng = np.random.default_rng(args.mix_seed)
    for f in (float(x) for x in args.serve_loads.split(",")):
        rate = f * loads["burst"]["solves_per_s"]
        arr = np.concatenate([[0.0], np.cumsum(rng.exponential(1.0 / rate, size=len(reqs) - 1))])
        loads[f"poisson_{f:g}"] = dict(arrival_rate=round(rate, 4), **serve(arr, f"poisson {rate:.3f}/s"))
    plain_wall = sum(p["seconds"] for p in plain)
    return dict(workload=dict(recipe="serving: cubic edge log-uniform [48,512], Boolean spheres r U[6,20], porosity "
                                     "U[0.30,0.90], random direction and method; Stokes limit; mixed with FP64 "
                                     "fallback on non-convergence",
                              requests=len(reqs), max_edge=args.serve_max_edge, skipped_larger=skipped,
                              fluid_cells=int(sum(d["fluid"] for _, d in reqs)),
                              images=[{k: d[k] for k in ("recipe_index", "shape", "porosity", "direction", "method")}
                                      for _, d in reqs]),
                plain_fp64=dict(solves_per_s=round(len(reqs) / plain_wall, 4), seconds=round(plain_wall, 2),
                                converged=sum(p["converged"] for p in plain)),
                loads=loads, speedup_burst_vs_plain_fp64=round(loads["burst"]["solves_per_s"] * plain_wall / len(reqs), 3))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24,
                    help="timed steps of the served mixed stream (24 x 2 requests: one cycle of the 48 images)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps of the stream before the timed ones")
    ap.add_argument("--per-step", type=int, default=2, help="completed requests per step")
    ap.add_argument("--backlog", type=int, default=2, help="stream: queued requests kept beyond the running ones")
    ap.add_argument("--section", default="all", choices=["all", "mix", "single", "cpu", "precision", "serving"],
                    help="precision: the 384^3 precision configuration run to convergence; serving: a bounded "
                         "sample of the serving workload (neither in the default run)")
    ap.add_argument("--serve-requests", type=int, default=20, help="serving: requests in the sample")
    ap.add_argument("--serve-max-edge", type=int, default=128, help="serving: largest edge in the sample")
    ap.add_argument("--serve-loads", default="0.5", help="serving: Poisson rates as fractions of the burst rate")
    ap.add_argument("--prec-parts", default="stokes,forch",
                    help="precision: stokes (run 1), stokes_horizon (its FP64 / mixed comparison only), forch (run 2)")
    ap.add_argument("--prec-horizon", type=int, default=30,
                    help="precision: outer iterations of the FP64 / mixed comparison of run 1")
    ap.add_argument("--forch-re", default="1,3,10", help="precision: target pore Reynolds numbers of run 2")
    ap.add_argument("--forch-horizon", type=int, default=10,
                    help="precision: outer iterations of the FP64 / mixed comparison of each point of run 2")
    ap.add_argument("--forch-k", type=float, default=0.0, help="precision: run 1's K when run 2 runs alone")
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--mix-seed", type=int, default=20261203)
    ap.add_argument("--edges", default="32,48,64,96,128,160")
    ap.add_argument("--rho", type=float, default=1e-3,
                    help="density: 1e-3 keeps the pore Reynolds number ~1e-4 (Stokes limit)")
    ap.add_argument("--slots", type=int, default=32)
    ap.add_argument("--chunk", type=int, default=8)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--poisson", action="store_true", help="also serve the mix once as Poisson arrivals")
    ap.add_argument("--poisson-frac", type=float, default=0.5)
    ap.add_argument("--mix-forms", default="all", choices=["all", "stream"],
                    help="stream: only the timed served mixed stream, no JSON line (for profilers)")
    ap.add_argument("--tune-requests", type=int, default=8)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--radius", type=float, default=10.0)
    ap.add_argument("--porosity", type=float, default=0.75)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--generator", default="rsa", choices=["rsa", "boolean"],
                    help="single image: RSA pack (--radius) or Boolean spheres (--rmin, --rmax)")
    ap.add_argument("--rmin", type=float, default=6.0)
    ap.add_argument("--rmax", type=float, default=20.0)
    ap.add_argument("--single-steps", type=int, default=1)
    ap.add_argument("--single-warmup", type=int, default=1, help="untimed FP64 solves before the timed ones")
    ap.add_argument("--simpletol", type=float, default=1e-6)
    ap.add_argument("--maxit", type=int, default=3000, help="outer-iteration cap, single image")
    ap.add_argument("--mix-maxit", type=int, default=1000,
                    help="outer-iteration cap of the mix (its requests converge in <= 440; one stalls at "
                         "RMR ~4e-6 in every precision and runs to the cap)")
    ap.add_argument("--inner-tol", type=float, default=1e-4, help="mixed form: FP32 sweep tolerance")
    ap.add_argument("--mixed-min-rows", type=int, default=0,
                    help="mixed form: systems with fewer rows are solved in FP64")
    ap.add_argument("--mg-min-rows", type=int, default=200000,
                    help="mixed form: pressure systems from this many rows on are swept by multigrid CG (0: never)")
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="no per-kernel events (profiler runs)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="served forms: launch the engine's chunks kernel by kernel, not as HIP graphs")
    ap.add_argument("--lanes", type=int, default=3, help="served forms: engine streams")
    ap.add_argument("--max-rows", type=float, default=2e6,
                    help="served forms: cap on the fluid cells of the running requests (0: none)")
    ap.add_argument("--tune-maxit", type=int, default=40, help="autotuner: outer-iteration cap of its trials")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="pick chunk / lanes / graphs with ns3d_amd.autotune before the mix (full runs)")
    ap.add_argument("--progress", action="store_true", help="one stderr line per outer iteration (single)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the library's plain device copy of MIB MiB (counter calibration) and print its GB/s")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); it has no multi-GPU mode: requests "
              f"would shard across GPUs by image, one server per GPU", file=sys.stderr)
        sys.exit(2)
    from ns3d_amd import gpu
    if gpu.device_count() < 1:
        raise SystemExit("no GPU visible")
    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        print(json.dumps({"copy_bytes": nbytes, "copy_GBs": round(gpu.copy_bandwidth(nbytes, 10), 1)}), flush=True)
        return
    # the achievable streaming rate the roofline fractions are read against
    copy_gbs = gpu.copy_bandwidth(1 << 30, 20)
    log(f"device copy (1 GiB read + written): {copy_gbs:.0f} GB/s")

    if args.section == "serving":
        print(json.dumps({"serving_config": serving_runs(args), "device_copy_GBs": round(copy_gbs, 1)}), flush=True)
        return
    if args.section == "precision":
        print(json.dumps({"precision_config": precision_runs(args), "device_copy_GBs": round(copy_gbs, 1)}),
              flush=True)
        return
    if args.section == "cpu":
        # the CPU-baseline plan's whole runs (not part of the default run)
        print(json.dumps({"cpu_runs": cpu_runs(args), "device_copy_GBs": round(copy_gbs, 1)}), flush=True)
        return
    mixres = mix_section(args) if args.section in ("all", "mix") else None
    if args.section == "mix" and mixres is None:
        return  # partial (profiling) run: no JSON line
    single = single_section(args) if args.section in ("all", "single") else None

    cpu = None
    if not args.no_cpu and mixres is not None:
        log(f"CPU baseline: oracle, {args.cpu_iters} outer iterations of one request")
        cpu = cpu_baseline_mix(mixres, args)
        log(f"  CPU {cpu['value']} solves/s ({cpu['sample']})")

    if mixres is not None:
        f = mixres["forms"]
        value = f["served_mixed"]["solves_per_s"]
        line = {
            "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(mixres["ms_per_step"], 1), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64 arithmetic, f32 inner-solve storage",
            "data": "synthetic", "form": "served_mixed",
            "config": mixres["desc"],
            "latency_s": f["served_mixed"]["latency_s"],
            "speedup_vs_plain_fp64": round(mixres["value"] / mixres["plain_rate"], 3),
            "max_dK_over_K_vs_fp64": mixres["max_dk"],
            "dK_gate": mixres["gate"],
            "autotune": mixres["autotune"],
            "forms": f,
        }
    else:
        s = single["out"]
        line = {"metric": METRIC, "value": s["mixed"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
                "steps": args.single_steps, "warmup": args.single_warmup, "ms_per_step": s["mixed"]["ms_per_solve"],
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f64 arithmetic, f32 inner-solve storage", "data": "synthetic", "form": "single_mixed",
                "config": {"workload": s["workload"]}}
    if mixres is not None:
        # dominant kernel of the headline form: the batched A kernel (FP32
        # storage) over all requests in flight, sampled by the engine
        r = engine_roofline(mixres["engine"], "f32", "A")
        traffic, tsrc = mix_traffic(args)
        it = engine_chunk(mixres["engine"])
        line["roofline"] = None if r is None else dict(
            bound="hbm", achieved=r["achieved"], peak=PEAK_HBM_GBS, unit="GB/s", frac=r["frac"], traffic=traffic,
            kernel="bicg_A_kernel<float> batched over the requests in flight (served mixed form, one of three "
                   "concurrent streams): p update + 7-point SpMV + r0.v", bytes_per_launch=r["bytes_per_launch"],
            avg_launch_us=r["avg_launch_us"], sampled_launches=r["sampled_launches"], traffic_source=tsrc,
            engine_chunks=it)
        line["served_mixed_engine"] = {k: v for k, v in mixres["engine"].items()}
    if single is not None:
        if mixres is None:
            line["roofline"] = single["roof"]
        else:
            line["roofline_single_fp64"] = single["roof"]
        # the mixed form's pressure sweeps: the multigrid kernel with the
        # most device time in the single-image solve
        mg = {k: v for k, v in single["kernels"].items() if k.startswith("mg_")}
        if mg:
            k, v = max(mg.items(), key=lambda kv: kv[1]["avg_us"] * kv[1]["launches"])
            line["roofline_multigrid_single"] = dict(
                bound="hbm", kernel=f"{k} (finest level of the multigrid CG sweep, {args.size}^3 image, mixed form)",
                achieved=v["GBs"], peak=PEAK_HBM_GBS, unit="GB/s", frac=v["frac"], avg_launch_us=v["avg_us"],
                bytes_per_launch=v["bytes_per_launch"], sampled_launches=v["launches"])
        line["single_image"] = single["out"]
        line["kernels"] = single["kernels"]
    line["cpu_baseline"] = cpu
    line["device_copy_GBs"] = round(copy_gbs, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
