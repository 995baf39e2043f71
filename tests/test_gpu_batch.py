"""Batched solves on one engine (GPU).

A batched solve runs the same kernels over the same tiles, with the same
partial-sum order, as the context's own path, so each result must be the
same alone and in the mix bit for bit: permeability, every RES row, the
fields and the solver iteration counts.  The mix holds images of uneven
size, all four algorithms, three flow directions and both precisions, so
solves of different systems, layouts and storage types share launches and
join and leave the batch at different times."""
from __future__ import annotations

import numpy as np
import pytest

from ns3d_amd import generators
from ns3d_amd import gpu as G
from ns3d_amd import solver

pytestmark = pytest.mark.gpu

PRM = dict(dx=1.0, mu=1.0, rho=1.0, p0=1.0, alpha=0.8, beta=1.0, vsolvertol=1e-8, psolvertol=1e-8,
           pcsolvertol=1e-5, maxitsolver=3000, simpletol=1e-5, maxitSIMPLE=300, N1=4, N2=4, device=0)


def mix():
    imgs = [generators.rsa_spheres((24, 22, 20), 3.5, 0.35, 41)[0],
            generators.boolean_spheres((20, 26, 18), 2.0, 4.0, 0.6, 42)[0],
            generators.fibres((28, 20, 22), 2.0, 12.0, 0.75, 43)[0],
            generators.rsa_spheres((16, 16, 30), 3.0, 0.3, 44)[0],
            generators.boolean_spheres((30, 18, 18), 2.0, 5.0, 0.7, 45)[0]]
    cases = [("SIMPLEC", 1, 0), ("SIMPLER", 1, 1), ("IDEAL", 2, 0), ("SIMPLERR", 3, 1), ("SIMPLEC", 1, 1)]
    params = []
    for img, (m, d, prec) in zip(imgs, cases):
        nx = img.shape[{1: 0, 2: 1, 3: 2}[d]]
        params.append(dict(PRM, method=m, direction=d, gradp=1.0 / nx, precision=prec, inner_tol=1e-4))
    return imgs, params


def test_batch_results_equal_solo_bit_for_bit():
    imgs, params = mix()
    solo = [solver.solve(img, **p) for img, p in zip(imgs, params)]
    batch = solver.solve_batch(imgs, params, chunk=4)
    for i, ((r0, s0), (r1, s1)) in enumerate(zip(solo, batch)):
        assert s0["converged"] == 1 and s1["converged"] == 1, i
        assert s0["K"] == s1["K"], (i, s0["K"], s1["K"])
        assert s0["iterations"] == s1["iterations"], i
        assert s0["solver_iterations"] == s1["solver_iterations"], i
        assert np.array_equal(r0, r1), i
        for f in "puvw":
            assert np.array_equal(s0["fields"][f], s1["fields"][f]), (i, f)
    st = batch[0][1]["engine"]
    assert st["solves"] > 0 and st["max_active"] >= 2, st


def test_graph_rounds_equal_direct_rounds():
    """Chunks submitted as HIP graphs -- each graph replayed round after
    round, the tables rewritten and the host's own work in between -- give
    the same bits as the same chunks launched kernel by kernel."""
    imgs, params = mix()
    direct = solver.solve_batch(imgs, params, chunk=4, graphs=False)
    graph = solver.solve_batch(imgs, params, chunk=4, graphs=True)
    for i, ((r0, s0), (r1, s1)) in enumerate(zip(direct, graph)):
        assert s0["K"] == s1["K"], (i, s0["K"], s1["K"])
        assert s0["solver_iterations"] == s1["solver_iterations"], i
        assert np.array_equal(r0, r1), i
        for f in "puvw":
            assert np.array_equal(s0["fields"][f], s1["fields"][f]), (i, f)
    assert direct[0][1]["engine"]["chunks"] > 0 and graph[0][1]["engine"]["chunks"] > 0


def test_schedules_give_the_same_bits():
    """The engine's schedule (chunk length, streams, graphs: what the
    autotuner chooses) moves only time: one stream with short chunks
    launched kernel by kernel and three streams with long graph chunks give
    the same results bit for bit; an out-of-range schedule is refused."""
    imgs, params = mix()
    out = []
    for chunk, lanes, graphs in ((3, 1, False), (12, 3, True)):
        eng = G.Engine(device=0, chunk=chunk, graphs=graphs, lanes=lanes)
        try:
            assert G.Engine.schedule_of(eng.h) == dict(chunk=chunk, lanes=lanes, graphs=graphs)
            out.append(solver.solve_batch(imgs, params, engine=eng))
        finally:
            eng.close()
    for i, ((r0, s0), (r1, s1)) in enumerate(zip(*out)):
        assert s0["K"] == s1["K"] and np.array_equal(r0, r1), i
        for f in "puvw":
            assert np.array_equal(s0["fields"][f], s1["fields"][f]), (i, f)
    with pytest.raises(G.NS3DError):
        G.Engine(device=0, chunk=4, lanes=5)


def test_lone_context_solves_on_its_own_stream():
    """A context alone on its engine solves on the stream its work runs on,
    leaving the engine idle; with a second context attached, the same run goes
    through the engine.  Both give the solo result bit for bit (mixed form:
    FP64 defect correction around FP32 inner solves)."""
    img = generators.rsa_spheres((18, 16, 14), 3.0, 0.3, 7)[0]
    p = dict(PRM, method="SIMPLEC", gradp=1.0 / 18, precision=1, inner_tol=1e-4)
    ref_rows, ref = solver.solve(img, **p)
    eng = G.Engine(device=0, chunk=4)
    try:
        s = solver.Solver(img, **p)
        s.attach(eng)
        rows, summ = s.run()
        s.close()
        assert eng.stats()["solves"] == 0
        assert summ["K"] == ref["K"] and np.array_equal(rows, ref_rows)
        other = solver.Solver(img, **p)
        other.attach(eng)
        s = solver.Solver(img, **p)
        s.attach(eng)
        rows, summ = s.run()
        s.close()
        other.close()
        assert eng.stats()["solves"] > 0
        assert summ["K"] == ref["K"] and np.array_equal(rows, ref_rows)
        assert summ["solver_iterations"] == ref["solver_iterations"]
    finally:
        eng.close()


def test_engine_lifecycle_and_detach():
    """A context detached from an engine solves on its own stream again; an
    engine with nothing attached shuts down cleanly; bad arguments fail."""
    img = generators.rsa_spheres((18, 16, 14), 3.0, 0.3, 7)[0]
    p = dict(PRM, method="SIMPLEC", gradp=1.0 / 18)
    ref = solver.solve(img, **p)[1]["K"]
    eng = G.Engine(device=0, chunk=2)
    s = solver.Solver(img, **p)
    s.attach(eng)
    s.attach(None)
    assert s.run()[1]["K"] == ref
    s.close()
    eng.close()
    h = G.C.c_void_p()
    assert G.lib().ns3d_engine_create(0, 0, G.C.byref(h)) != 0
    assert G.lib().ns3d_engine_create(0, 100000, G.C.byref(h)) != 0
