"""Extract the golden vectors of the reference's own test fixtures.

Run in a container that holds the reference checkout (read-only):

    python tests/golden/make_golden.py /root/reference

It reads only data files of the reference (no code is imported or run):

* data/spheres_{64,128}.vti -> sha256 of the decoded voxel array (C order,
  shape (nx, ny, nz), 1 = solid) and the solid fraction, so the lattice
  generator can be checked without the 12 MB images;
* tests/results/INFO-1-spheres_128.log -> node counts, porosity, p0/p1,
  every solver report (component, iterations, status), every K line, the
  pressure-correction max/min/norm lines and the final divergence lines;
* tests/results/RES-1-spheres_128.csv -> the RES rows as printed;
* the INFO log's first 16 lines verbatim and the XML headers of the field
  .vti files (up to the appended data);
* tests/results/V{x,y,z}-1-spheres_128.vti -> the final face velocities on the
  node sets are summarised by per-component sums, sums of squares and the
  outlet-plane sum (enough to check a re-run's final fields and the outlet
  flux without storing 45 MB of doubles).

The parameters of the fixture run are those of tests/test_spheres.ini.
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from ns3d_amd import vti  # noqa: E402


def parse_info(path):
    out = {"solves": [], "K": [], "pc": [], "iterations": []}
    cur = None
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            m = re.match(r"Size (\d+),(\d+),(\d+)", line)
            if m:
                out["size"] = [int(x) for x in m.groups()]
            m = re.match(r"Porosity: (\S+)", line)
            if m:
                out["porosity"] = float(m.group(1))
            for key, name in (("np", "pressure"), ("nu", "u velocity"), ("nv", "v velocity"),
                              ("nw", "w velocity")):
                m = re.match(r"Nodes for the %s field: (\d+)" % name, line)
                if m:
                    out[key] = int(m.group(1))
            m = re.match(r"p0, p1 \(Pa\)= (\S+),(\S+)", line)
            if m:
                out["p0"], out["p1"] = float(m.group(1)), float(m.group(2))
            m = re.match(r"Iteration (\d+)", line)
            if m:
                cur = int(m.group(1))
                out["iterations"].append(cur)
            m = re.match(r"Solving for component (\w) in (\d+) iterations \(status (\d+)\)", line)
            if m:
                out["solves"].append({"it": cur, "comp": m.group(1), "iterations": int(m.group(2)),
                                      "status": int(m.group(3))})
            m = re.match(r"Solving for pressure correction : (\d+) iterations \(status (\d+)\)", line)
            if m:
                out["solves"].append({"it": cur, "comp": "pc", "iterations": int(m.group(1)),
                                      "status": int(m.group(2))})
            m = re.match(r"(Initial )?Permeability K \(m²\): (\S+)", line)
            if m:
                out["K"].append({"it": cur, "initial": bool(m.group(1)), "K": float(m.group(2))})
            m = re.match(r"Correction de pression maximale et minimale: (\S+), (\S+)", line)
            if m:
                out["pc"].append({"it": cur, "max": float(m.group(1)), "min": float(m.group(2))})
            m = re.match(r"Norme de la correction de pression : (\S+)", line)
            if m:
                out["pc"][-1]["norm"] = float(m.group(1))
            m = re.match(r"Norm of the velocity field divergence: (\S+)", line)
            if m:
                out["final_normdiv"] = float(m.group(1))
            m = re.match(r"Relative divergence of the velocity field: (\S+)", line)
            if m:
                out["final_reldiv"] = float(m.group(1))
    return out


def parse_res(path):
    with open(path) as f:
        header = f.readline().strip()
        rows = [line.strip() for line in f if line.strip()]
    return header, rows


def main(ref):
    gold = {}
    lat = {}
    for name in ("spheres_64", "spheres_128"):
        a = vti.read_vti(os.path.join(ref, "data", name + ".vti"))
        lat[name] = {"shape": list(a.shape), "sha256": hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest(),
                     "solid_fraction": float(a.mean()), "solid_voxels": int(a.sum())}
    lat["spheres_64"]["generator"] = [64, 8, 3]
    lat["spheres_128"]["generator"] = [128, 12, 5]
    gold["lattice"] = lat

    res = os.path.join(ref, "tests", "results")
    gold["info"] = parse_info(os.path.join(res, "INFO-1-spheres_128.log"))
    # the log's first lines as written (image summary, setup, first solves)
    with open(os.path.join(res, "INFO-1-spheres_128.log")) as f:
        gold["info_head"] = [next(f).rstrip("\n") for _ in range(16)]
    # the .vti headers of the field files (XML up to the appended data)
    heads = {}
    for name in ("P", "Vx", "Vy", "Vz"):
        with open(os.path.join(res, "%s-1-spheres_128.vti" % name), "rb") as f:
            txt = f.read(2048).decode("latin-1")
        heads[name] = txt[:txt.index("<AppendedData")]
    gold["vti_heads"] = heads
    gold["res_header"], gold["res_rows"] = parse_res(os.path.join(res, "RES-1-spheres_128.csv"))
    gold["params"] = dict(input="spheres_128", dx=2e-6, dy=-1, dz=-1, mu=1e-3, rho=1e3, p0=1e5,
                          gradp=1e2, alpha=0.8, beta=1.0, vsolvertol=1e-5, psolvertol=1e-6,
                          pcsolvertol=1e-3, simpletol=1e-4, deltaRMR=1e-4, maxitsolver=500,
                          maxitSIMPLE=30, N1=2, N2=2, porevalue=0, direction=1, border=0,
                          method="SIMPLEC")

    fields = {}
    for comp in ("x", "y", "z"):
        for tag in ("", "-SAVE-1"):
            a = vti.read_vti(os.path.join(res, "V%s-1-spheres_128%s.vti" % (comp, tag)))
            d = {"shape": list(a.shape), "sum": float(a.sum(dtype=np.float64)),
                 "sumsq": float((a * a).sum(dtype=np.float64)), "nonzero": int(np.count_nonzero(a))}
            if comp == "x":
                d["outlet_sum"] = float(a[-1].sum(dtype=np.float64))
            fields["V%s%s" % (comp, tag)] = d
    for tag in ("", "-SAVE-1"):
        a = vti.read_vti(os.path.join(res, "P-1-spheres_128%s.vti" % tag))
        fields["P" + tag] = {"shape": list(a.shape), "sum": float(a.sum(dtype=np.float64)),
                             "nonzero": int(np.count_nonzero(a))}
    gold["fields"] = fields

    with open(os.path.join(HERE, "spheres128_simplec.json"), "w") as f:
        json.dump(gold, f, indent=1)
    print("wrote", os.path.join(HERE, "spheres128_simplec.json"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
