"""Oracle fixtures of the parity configurations at full size.

    python tests/golden/make_config_parity.py [case ...]   (all cases by default)

Images (seeded, ns3d_amd.generators, seed 20261203):
  c1      64^3 random sequential addition of spheres r = 8, solid fraction
          target 0.40 (RSA jams near 0.38: the allowed overlap grows by one
          voxel after 20000 consecutive rejections; achieved porosity and
          overlap are recorded);
  c2_075  128^3 RSA r = 10, porosity 0.75 (plain RSA reaches it);
  c2_045  128^3 RSA r = 10, porosity 0.45 (RSA densified the same way: the
          overlap grows to 4 voxels).
Physics: unit pressure drop (p0 = 1, gradp = 1/nx), mu = rho = dx = 1,
alpha = 0.8, beta = 1; inner tolerances 1e-10 (velocity, pressure) and 1e-6
(pressure correction), maxitsolver 5000; simpletol 1e-6; N1 = N2 = 4.

The outer iteration's stop at simpletol 1e-6 is decided by rounding on
these images (the oracle run on the same image mirrored in y ends 3.4e-5
away in K, at another iteration), so each case is run to a horizon instead:
the first outer iteration (it > 1) with RMR < 1e-3 is completed and the run
stops there as if capped (at most MAX_OUTER iterations).  The GPU test runs
the same case with maxitSIMPLE = that count and compares RES rows, K and the
fields.

Per case the fixture holds: the RES rows, the final K and iteration count,
the node counts and sha256 of the four index arrays, and per field (p, u,
v, w) its exact sum and sum of squares and a CountSketch (tests/helpers.py)
of 1024 buckets.  The oracle is the checker here; nothing else is run.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, os.path.dirname(TESTS))
sys.path.insert(0, TESTS)

from ns3d_amd import generators  # noqa: E402

OUT = os.path.join(HERE, "config_parity")
SEED = 20261203
HORIZON_RMR = 1e-3
MAX_OUTER = 400
SKETCH_M = 1024

IMAGES = {
    "c1": dict(edge=64, radius=8.0, porosity=0.60),
    "c2_075": dict(edge=128, radius=10.0, porosity=0.75),
    "c2_045": dict(edge=128, radius=10.0, porosity=0.45),
}

CASES = {
    "c1_simplec_x": ("c1", "SIMPLEC", 1),
    "c2_075_simplec_x": ("c2_075", "SIMPLEC", 1),
    "c2_075_simpler_x": ("c2_075", "SIMPLER", 1),
    "c2_075_simplerr_x": ("c2_075", "SIMPLERR", 1),
    "c2_075_ideal_x": ("c2_075", "IDEAL", 1),
    "c2_075_simplec_y": ("c2_075", "SIMPLEC", 2),
    "c2_075_simplec_z": ("c2_075", "SIMPLEC", 3),
    "c2_045_simplec_x": ("c2_045", "SIMPLEC", 1),
    "c2_045_simpler_x": ("c2_045", "SIMPLER", 1),
    "c2_045_simplerr_x": ("c2_045", "SIMPLERR", 1),
    "c2_045_ideal_x": ("c2_045", "IDEAL", 1),
    "c2_045_simplec_y": ("c2_045", "SIMPLEC", 2),
    "c2_045_simplec_z": ("c2_045", "SIMPLEC", 3),
}

SIMPLETOL = 1e-6
N1 = N2 = 4


def image(name):
    """(raw image, recipe with achieved porosity)."""
    r = IMAGES[name]
    img, solid, overlap, count = generators.rsa_spheres((r["edge"],) * 3, r["radius"], 1.0 - r["porosity"], SEED)
    rec = dict(r, seed=SEED, achieved_porosity=1.0 - solid, overlap=overlap, spheres=count,
               sha256=hashlib.sha256(img.tobytes()).hexdigest())
    return img, rec


def params(nx):
    return dict(dx=1.0, mu=1.0, rho=1.0, p0=1.0, gradp=1.0 / nx, alpha=0.8, beta=1.0, vsolvertol=1e-10,
                psolvertol=1e-10, pcsolvertol=1e-6, maxitsolver=5000)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_case(case):
    import oracle_lib as O
    from helpers import count_sketch
    imname, method, direction = CASES[case]
    raw, rec = image(imname)
    img = O.prepare_image(raw, direction, 0)
    prm = params(img.shape[0])
    o = O.Oracle(img, **prm)
    o.set_horizon(HORIZON_RMR)
    t0 = time.time()
    rows, summ = o.solve(method, SIMPLETOL, MAX_OUTER, N1=N1, N2=N2)
    sec = time.time() - t0
    pg, ug, vg, wg = o.index()
    fields = dict(p=o.get("p"), u=o.get("u"), v=o.get("v"), w=o.get("w"))
    meta = dict(case=case, image=imname, recipe=rec, method=method, direction=direction, dims=list(img.shape),
                params=prm, simpletol=SIMPLETOL, N1=N1, N2=N2, horizon_rmr=HORIZON_RMR, max_outer=MAX_OUTER,
                horizon_it=o.horizon_it(), final_it=summ["iterations"], K=summ["K"],
                counts=[o.np, o.nu, o.nv, o.nw], index_sha256=[sha(pg), sha(ug), sha(vg), sha(wg)],
                solver_iterations=summ["solver_iterations"], oracle_seconds=round(sec, 1),
                oracle_threads=int(O.lib().orc_num_threads()), sketch_m=SKETCH_M,
                field_sum={k: float(v.sum()) for k, v in fields.items()},
                field_sumsq={k: float((v * v).sum()) for k, v in fields.items()})
    np.savez_compressed(os.path.join(OUT, case + ".npz"), rows=rows,
                        **{"sketch_" + k: count_sketch(v, SKETCH_M) for k, v in fields.items()})
    with open(os.path.join(OUT, case + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(f"{case}: {method} dir {direction} horizon_it {meta['horizon_it']} final_it {meta['final_it']} "
          f"K {meta['K']!r} {sec:.0f} s", flush=True)


if __name__ == "__main__":
    os.makedirs(OUT, exist_ok=True)
    for c in sys.argv[1:] or list(CASES):
        run_case(c)
