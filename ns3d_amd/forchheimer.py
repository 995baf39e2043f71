"""Forchheimer sweep: one image at several imposed pressure gradients, and a
least-squares fit of the Darcy--Forchheimer relation

    -grad<p> = mu / k1 * <v> + rho / k2 * <v>^2

(doc/NS3D_methods.tex:520-531, which states the relation and leaves the
identification of k1 and k2 to the user; NS3D has no fitting code).

Each point is an ordinary permeability solve of the drop-in drivers with its
own `gradp`; <v> is the solve's `v_moy`, the mean outlet velocity that K is
computed from (K = mu <v> / gradp, NSBase.cxx:1886-1956).  The points are
independent, so by default they run together on one solve engine
(solver.solve_batch): their BiCGSTAB solves share launches.
"""
from __future__ import annotations

import math

import numpy as np


def fit(gradp, v, mu: float, rho: float) -> dict:
    """Least squares of gradp = a v + b v^2 over the points (no constant
    term: no flow without a gradient).  Returns a, b, k1 = mu / a,
    k2 = rho / b (inf when b <= 0: no measurable inertial term), the
    relative RMS residual of the fit and the number of points."""
    g = np.asarray(gradp, dtype=np.float64)
    x = np.asarray(v, dtype=np.float64)
    if g.shape != x.shape or g.ndim != 1 or len(g) < 2:
        raise ValueError("fit needs at least two (gradp, v) points of the same length")
    if not (np.all(np.isfinite(g)) and np.all(np.isfinite(x))):
        raise ValueError("fit needs finite points")
    A = np.column_stack([x, x * x])
    (a, b), *_ = np.linalg.lstsq(A, g, rcond=None)
    res = g - A @ np.array([a, b])
    rel = float(np.sqrt(np.mean(res * res)) / np.sqrt(np.mean(g * g)))
    k1 = mu / a if a > 0 else math.inf
    k2 = rho / b if b > 0 else math.inf
    return dict(a=float(a), b=float(b), k1=float(k1), k2=float(k2), rel_residual=rel, points=int(len(g)))


def sweep(img: np.ndarray, gradps, batch: bool = True, **params) -> dict:
    """Solves img at every gradient of gradps (same parameters otherwise) and
    fits the Forchheimer relation to the converged points.  Returns the
    points (gradp, v_moy, K, apparent K = mu v / gradp, Reynolds number
    rho v dx / mu per voxel length, converged, outer iterations) and the
    fit (see fit())."""
    from . import solver
    gradps = [float(g) for g in gradps]
    if len(gradps) < 2 or any(not (g > 0) for g in gradps):
        raise ValueError("sweep needs at least two positive gradients")
    if "gradp" in params:
        raise ValueError("gradp is set per point by the sweep")
    plist = [dict(params, gradp=g) for g in gradps]
    if batch:
        out = [s for _, s in solver.solve_batch([img] * len(gradps), plist)]
    else:
        out = [solver.solve(img, **p)[1] for p in plist]
    mu = float(params.get("mu", solver.default_params().mu))
    rho = float(params.get("rho", solver.default_params().rho))
    dx = float(params.get("dx", solver.default_params().dx))
    points = []
    for g, s in zip(gradps, out):
        v = float(s["v_moy"])
        points.append(dict(gradp=g, v_moy=v, K=float(s["K"]), K_apparent=mu * v / g, Re=rho * v * dx / mu,
                           converged=int(s["converged"]), iterations=int(s["iterations"])))
    good = [p for p in points if p["converged"] and p["v_moy"] > 0]
    result = dict(points=points)
    result["fit"] = fit([p["gradp"] for p in good], [p["v_moy"] for p in good], mu, rho) if len(good) >= 2 else None
    return result


def main(argv=None) -> int:
    """python -m ns3d_amd.forchheimer image.vti --gradp 1e2,1e3,1e4 [key=value ...]"""
    import argparse
    import json
    from . import solver
    from . import vti
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("image")
    ap.add_argument("--gradp", required=True, help="comma-separated pressure gradients")
    ap.add_argument("--sequential", action="store_true", help="one solve after the other (no engine)")
    ap.add_argument("params", nargs="*", help="solver parameters as key=value (the .ini keys)")
    a = ap.parse_args(argv)
    kw = {}
    for item in a.params:
        k, _, v = item.partition("=")
        cur = getattr(solver.default_params(), k)
        kw[k] = type(cur)(float(v)) if isinstance(cur, (int, float)) else v
    img = np.ascontiguousarray(vti.read_vti(a.image), dtype=np.uint8)  # binarised by the solver (porevalue)
    res = sweep(img, [float(x) for x in a.gradp.split(",")], batch=not a.sequential, **kw)
    print(json.dumps(res, indent=1))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
