"""Serving (GPU): requests taken as they arrive by the native server.

Each served result must be the same alone and in the mix -- bit for bit,
since the engine runs a request's solves with the same kernels, tiles and
summation order as its own context would -- while requests overlap on the
device, admission respects the device-memory budget, and a mixed-precision
request that does not converge is re-run in FP64."""
from __future__ import annotations

import time

import numpy as np
import pytest

from ns3d_amd import generators
from ns3d_amd import gpu as G
from ns3d_amd import solver

pytestmark = pytest.mark.gpu

PRM = dict(dx=1.0, mu=1.0, rho=1.0, p0=1.0, alpha=0.8, beta=1.0, vsolvertol=1e-8, psolvertol=1e-8,
           pcsolvertol=1e-5, maxitsolver=3000, simpletol=1e-5, maxitSIMPLE=300, N1=4, N2=4, device=0)


def requests():
    out = []
    gens = [lambda s: generators.rsa_spheres((22, 20, 18), 3.0, 0.3, s)[0],
            lambda s: generators.boolean_spheres((18, 24, 20), 2.0, 4.0, 0.65, s)[0],
            lambda s: generators.fibres((26, 18, 20), 2.0, 10.0, 0.75, s)[0]]
    methods = ["SIMPLEC", "SIMPLER", "SIMPLERR", "IDEAL"]
    for q in range(8):
        img = gens[q % 3](100 + q)
        d = 1 + q % 3
        nx = img.shape[d - 1]
        out.append((img, dict(PRM, method=methods[q % 4], direction=d, gradp=1.0 / nx, precision=q % 2,
                              inner_tol=1e-4)))
    return out


def test_served_results_equal_solo_with_overlap():
    reqs = requests()
    solo = [solver.solve(img, **p)[1] for img, p in reqs]
    srv = solver.Server(device=0, budget_bytes=64e9, max_slots=4, chunk=4)
    try:
        tickets = []
        for q, (img, p) in enumerate(reqs):
            tickets.append(srv.submit(img, **p))
            if q % 3 == 2:
                time.sleep(0.05)  # later requests arrive while earlier ones run
        served = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    for q, (a, b) in enumerate(zip(solo, served)):
        assert a["converged"] == 1 and b["converged"] == 1 and b["fallback"] == 0, q
        assert a["K"] == b["K"], (q, a["K"], b["K"])
        assert a["iterations"] == b["iterations"], q
        assert a["solver_iterations"] == b["solver_iterations"], q
        assert b["t_submit"] <= b["t_start"] <= b["t_end"]
    assert st["done"] == len(reqs) and st["failed"] == 0
    assert st["peak_running"] >= 2, st  # requests did overlap on the device
    starts = sorted(b["t_start"] for b in served)
    ends = sorted(b["t_end"] for b in served)
    assert starts[1] < ends[0]


def test_admission_respects_budget():
    reqs = requests()[:6]
    fp = [G.footprint(img.shape, int((img == 0).sum()), p["precision"]) for img, p in reqs]
    budget = 2.2 * max(fp)  # room for two of the largest at a time
    srv = solver.Server(device=0, budget_bytes=budget, max_slots=6, chunk=4)
    try:
        tickets = [srv.submit(img, **p) for img, p in reqs]
        served = [srv.wait(t) for t in tickets]
        st = srv.stats()
        # a request larger than the whole budget is refused at submission
        big = np.zeros((64, 64, 64), np.uint8)
        with pytest.raises(G.NS3DError):
            srv.submit(big, **dict(PRM, gradp=1.0 / 64))
    finally:
        srv.close()
    assert all(b["converged"] == 1 for b in served)
    assert st["peak_bytes_in_use"] <= budget and st["peak_running"] >= 1, st
    assert st["bytes_in_use"] == 0


def test_rows_cap_and_largest_first():
    """With the fluid cells in flight capped at the largest request's, one
    request runs at a time (in arrival order the last-submitted largest
    request would start last; largest-first starts it first), and every
    result is still the one the request gets alone."""
    reqs = requests()[:4]
    rows = [int((img == 0).sum()) for img, _ in reqs]
    solo = [solver.solve(img, **p)[1]["K"] for img, p in reqs]
    order = []
    for largest in (False, True):
        srv = solver.Server(device=0, budget_bytes=16e9, max_slots=4, chunk=4)
        try:
            srv.set_policy(largest_first=largest, max_rows=max(rows))
            tickets = [srv.submit(img, **p) for img, p in reqs]
            out = [srv.wait(t) for t in tickets]
            st = srv.stats()
        finally:
            srv.close()
        assert st["peak_running"] == 1, st
        assert [r["K"] for r in out] == solo
        order.append(sorted(range(4), key=lambda q: out[q]["t_start"]))
    assert order[0] == [0, 1, 2, 3]
    assert order[1][0] == int(np.argmax(rows)) or order[1][0] == 0  # the first may start before the others queue
    with pytest.raises(G.NS3DError):
        srv2 = solver.Server(device=0, budget_bytes=1e9, max_slots=1, chunk=4)
        try:
            srv2.set_policy(largest_first=True, max_rows=-1.0)
        finally:
            srv2.close()


def test_footprint_bounds_context_allocation():
    eng = G.Engine(0, 8)
    try:
        for shape, seed, mixed, attached in (((20, 18, 16), 1, False, False), ((24, 22, 20), 2, True, False),
                                             ((16, 30, 12), 3, True, True), ((22, 20, 18), 4, False, True)):
            img = generators.rsa_spheres(shape, 3.0, 0.3, seed)[0]
            g = G.NavierStokesGPU(img)
            if attached:
                eng.attach(g.h)
            if mixed:
                g.set_precision(True, 1e-4)
            used = g.device_bytes()
            bound = G.footprint(img.shape, int((img == 0).sum()), mixed, engine=attached)
            if attached:
                eng.detach(g.h)
            g.close()
            assert used <= bound <= 1.5 * used + 8e6, (shape, mixed, attached, used, bound)
    finally:
        eng.close()


def test_mixed_request_falls_back_to_fp64():
    img = generators.rsa_spheres((18, 16, 14), 3.0, 0.3, 9)[0]
    p = dict(PRM, method="SIMPLEC", gradp=1.0 / 18, precision=1, maxitSIMPLE=3)
    srv = solver.Server(device=0, budget_bytes=16e9, max_slots=2, chunk=4, fallback=solver.FALLBACK_UNCONVERGED)
    try:
        r = srv.wait(srv.submit(img, **p))
        st = srv.stats()
    finally:
        srv.close()
    # FP64 continuation from the mixed run's fields with a quarter of the
    # outer-iteration cap (at least 10): 3 + 10 outer iterations in all
    assert r["fallback"] == 1 and st["fallbacks"] == 1
    assert r["iterations"] == 3 + 10
    # the default mode only replaces failed runs: an unconverged mixed
    # result stands
    srv = solver.Server(device=0, budget_bytes=16e9, max_slots=2, chunk=4)
    try:
        r = srv.wait(srv.submit(img, **p))
        st = srv.stats()
    finally:
        srv.close()
    assert r["fallback"] == 0 and st["fallbacks"] == 0
    assert r["iterations"] == 3 and r["converged"] == 0


def test_warm_start_from_converged_fields():
    """A run started from converged fields (the in-memory form of the
    pressurefield / vxfield inputs) converges in far fewer outer iterations
    than the cold run (the driver restarts its relaxation history, so not at
    once) to the same permeability."""
    img = generators.rsa_spheres((22, 18, 16), 3.0, 0.3, 12)[0]
    p = dict(PRM, method="SIMPLEC", gradp=1.0 / 22)
    rows, first = solver.solve(img, **p)
    f = first["fields"]
    s = solver.Solver(img, **p)
    try:
        s.warm_start(f["p"], f["u"], f["v"], f["w"])
        rows2, again = s.run()
    finally:
        s.close()
    assert first["converged"] == 1 and again["converged"] == 1
    assert again["iterations"] < first["iterations"] / 2, (again["iterations"], first["iterations"])
    assert abs(again["K"] / first["K"] - 1) < 1e-4
