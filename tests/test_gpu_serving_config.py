"""The serving workload's requests served as a Poisson stream (GPU).

Six requests of the serving recipe (ns3d_amd.mix.serving_mix: cubic edge
log-uniform in [48, 512], kept here to edges <= 64; Boolean spheres r U[6,
20], porosity U[0.30, 0.90]; each with its own flow direction and method),
Stokes limit, arriving at Poisson times at a server that runs the mixed form
with FP64 fallback on non-convergence.  Each result must be the request's
own mixed solve alone on a context (bit for bit: same kernels, tiles and
summation order) unless it fell back, and within |dK|/K 1e-4 of its plain
FP64 solve.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

from ns3d_amd import mix
from ns3d_amd import solver

pytestmark = pytest.mark.gpu

PRM = dict(dx=1.0, mu=1.0, rho=1e-3, p0=1.0, alpha=0.8, beta=1.0, vsolvertol=1e-10, psolvertol=1e-10,
           pcsolvertol=1e-6, maxitsolver=5000, simpletol=1e-6, maxitSIMPLE=1000, N1=4, N2=4, device=0,
           inner_tol=1e-4)


def test_serving_recipe_poisson_stream():
    reqs = mix.serving_mix(6, max_edge=64)
    params = [dict(PRM, method=d["method"], direction=d["direction"], gradp=1.0 / d["shape"][0]) for _, d in reqs]
    print("[serving] requests: " + ", ".join(f"{d['shape'][0]}^3 {d['method']} dir {d['direction']} "
                                             f"phi {d['porosity']}" for _, d in reqs), flush=True)
    arrivals = np.concatenate([[0.0], np.cumsum(np.random.default_rng(7).exponential(0.25, size=len(reqs) - 1))])
    srv = solver.Server(device=0, budget_bytes=64e9, max_slots=8, chunk=8, fallback=solver.FALLBACK_UNCONVERGED)
    try:
        t0 = time.perf_counter()
        tickets = []
        for q, ((img, _), p) in enumerate(zip(reqs, params)):
            dt = arrivals[q] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets.append(srv.submit(img, **dict(p, precision=1)))
        served = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    assert st["done"] == len(reqs) and st["failed"] == 0
    for q, ((img, d), p, b) in enumerate(zip(reqs, params, served)):
        s64 = solver.solve(img, **dict(p, precision=0))[1]
        dk = abs(b["K"] / s64["K"] - 1)
        print(f"[serving] request {q}: served K {b['K']!r} (fallback {b['fallback']}, {b['iterations']} its), "
              f"FP64 K {s64['K']!r} ({s64['iterations']} its), |dK|/K {dk:.2e}", flush=True)
        if b["fallback"] == 0:
            a = solver.solve(img, **dict(p, precision=1))[1]
            assert a["K"] == b["K"] and a["iterations"] == b["iterations"], q
            assert a["solver_iterations"] == b["solver_iterations"], q
        if s64["converged"]:
            assert b["converged"] == 1 and dk <= 1e-4, q
        assert b["t_submit"] <= b["t_start"] <= b["t_end"]
