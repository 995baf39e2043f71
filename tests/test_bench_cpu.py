"""bench.py's host-side contract, without a GPU: more than one GPU is refused
with a message, and the single-image workloads are the ones their flags
name (the 128^3 RSA pack by default, the precision configuration's Boolean
spheres on request)."""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_single_workloads():
    a = argparse.Namespace(size=24, radius=3.0, porosity=0.75, seed=5, generator="rsa", rmin=6.0, rmax=20.0,
                           rho=1e-3)
    img, params, desc = bench.single_workload(a)
    assert img.shape == (24, 24, 24) and img.dtype == np.uint8
    assert desc["workload"].startswith("24^3 RSA") and params["gradp"] == 1.0 / 24 and params["rho"] == 1e-3
    b = argparse.Namespace(size=48, radius=3.0, porosity=0.70, seed=5, generator="boolean", rmin=3.0, rmax=6.0,
                           rho=1e-3)
    img, params, desc = bench.single_workload(b)
    assert img.shape == (48, 48, 48) and desc["workload"].startswith("48^3 Boolean")
    assert abs((1.0 - img.mean()) - 0.70) < 0.02
    assert desc["image"]["rmin"] == 3.0 and desc["image"]["rmax"] == 6.0
    # the recorded 128^3 traffic applies to the default RSA image only
    assert bench.measured_traffic(b, "bicg_A_kernel<double>") == (None, None)


def test_dk_gate_is_strict():
    """Every completed request is held to 1e-4 against the FP64 K of its image;
    a request that converges in FP64 but not mixed fails; one that stalls in
    both is listed and not compared."""
    plain = [dict(K=1.0, converged=1), dict(K=2.0, converged=1), dict(K=3.0, converged=0)]
    ok = [(0, dict(K=1.0 + 5e-5, converged=1), 0.0), (1, dict(K=2.0, converged=1), 0.0),
          (2, dict(K=3.1, converged=0), 0.0), (3, dict(K=1.0 - 9.9e-5, converged=1), 0.0)]
    g = bench.dk_gate(ok, plain)
    assert g["passed"] and g["requests"] == 3 and g["within"] == 3 and g["stalled_both"] == [2]
    over = ok + [(4, dict(K=2.0 * (1 + 1.01e-4), converged=1), 0.0)]
    g = bench.dk_gate(over, plain)
    assert not g["passed"] and g["over"][0]["image"] == 1
    unconv = ok + [(4, dict(K=2.0, converged=0), 0.0)]
    g = bench.dk_gate(unconv, plain)
    assert not g["passed"] and g["unconverged_mixed"] == [1]
