"""Autotuner for the solve engine's schedule: the chunk length (iterations
queued per round), the number of streams the solves are spread over, and
whether chunks are submitted as HIP graphs.

It times a sample of permeability requests through the native server under
each candidate schedule and keeps the fastest.  None of these parameters
changes any result (the engine's launches compute the same bits in any
schedule), so the choice moves only time.  The tile shape is not tuned here:
it fixes the order of the BiCGSTAB partial sums, so it is a build constant
(NS3D_TILE_ROWS), chosen by the same kind of timing offline.
"""
from __future__ import annotations

import json
import time

DEFAULT_CANDIDATES = [(c, l, True) for c in (4, 6, 8, 12) for l in (1, 2, 3)] + [(8, 2, False)]


def sample(count: int = 12, seed: int = 7, edges=(32, 48, 64), maxit: int = 40, **params):
    """A request sample: `count` seeded images of the bench mix's kinds (RSA,
    Boolean, fibres) and their parameters, each request capped at `maxit`
    outer iterations so that a trial takes about a second."""
    from . import mix
    reqs = mix.request_mix(count, seed, tuple(edges))
    out = []
    for img, _ in reqs:
        p = dict(dx=1.0, mu=1.0, rho=1.0, p0=1.0, gradp=1.0 / img.shape[0], alpha=0.8, beta=1.0, direction=1,
                 border=0, method="SIMPLEC", simpletol=1e-6, maxitSIMPLE=maxit, precision=1, inner_tol=1e-4,
                 vsolvertol=1e-10, psolvertol=1e-10, pcsolvertol=1e-6, maxitsolver=5000, device=0)
        p.update(params)
        out.append((img, p))
    return out


def run_pass(requests, chunk: int, lanes: int, graphs: bool, budget_gb: float = 200.0, slots: int = 32) -> float:
    """Wall seconds of one burst of the requests through a server with that
    schedule."""
    from . import solver
    srv = solver.Server(device=0, budget_bytes=budget_gb * 1e9, max_slots=slots, chunk=chunk)
    try:
        srv.set_schedule(chunk, lanes, graphs)
        t0 = time.perf_counter()
        tickets = [srv.submit(img, **p) for img, p in requests]
        for t in tickets:
            srv.wait(t)
        return time.perf_counter() - t0
    finally:
        srv.close()


def tune(requests, candidates=None, repeats: int = 1, log=None, **kw) -> dict:
    """Times every candidate (chunk, lanes, graphs) on the requests (after one
    untimed pass that loads and warms everything) and returns the fastest
    with all trials: {"chunk", "lanes", "graphs", "seconds", "trials"}."""
    candidates = list(candidates or DEFAULT_CANDIDATES)
    run_pass(requests, *candidates[0], **kw)
    trials = []
    for cand in candidates:
        secs = min(run_pass(requests, *cand, **kw) for _ in range(max(1, repeats)))
        trials.append(dict(chunk=cand[0], lanes=cand[1], graphs=bool(cand[2]), seconds=round(secs, 4)))
        if log:
            log(f"  schedule chunk {cand[0]:2d} lanes {cand[1]} graphs {int(cand[2])}: {secs:.3f} s")
    best = min(trials, key=lambda t: t["seconds"])
    return dict(best, trials=trials)


def main(argv=None) -> int:
    """python -m ns3d_amd.autotune [--requests N] [--maxit M] [--out schedule.json]"""
    import argparse
    import sys
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--requests", type=int, default=12)
    ap.add_argument("--maxit", type=int, default=40)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--repeats", type=int, default=1)
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    res = tune(sample(a.requests, a.seed, maxit=a.maxit), repeats=a.repeats,
               log=lambda m: print(m, file=sys.stderr, flush=True))
    text = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text + "\n")
    print(text)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
