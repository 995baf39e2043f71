"""The precision and batching configurations at full size (GPU).

Precision: one 384^3 image (seeded Boolean model of overlapping spheres,
radii U[6, 20], porosity 0.70, flow along x, unit pressure drop, mu = rho =
1), 39.6M pressure rows.  A pressure-correction solve of the mixed form
(multigrid-preconditioned FP32 sweeps inside FP64 defect correction) must
meet the FP64 true-residual contract, checked on the host against the
assembled matrix; the first outer iterations of the plain FP64 and the
mixed SIMPLEC runs must agree on K to |dK|/K <= 1e-4 (the metric's gate).

Batching: the 48 seeded images of the batching recipe (edges from {32, 48,
64, 96, 128, 160} per axis, porosity U[0.35, 0.90], RSA / Boolean / fibre
generators; ns3d_amd.mix), SIMPLEC, mixed form, all solved together on one
engine for a capped number of outer iterations.  The largest and the
smallest image must give the same bits as alone, and the mixed form the
FP64 K to 1e-4 at the same outer iteration.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as O
from ns3d_amd import generators, mix
from ns3d_amd import gpu as G
from ns3d_amd import solver

pytestmark = pytest.mark.gpu

SEED = 20261203


def spmv(diag, off, nbr, x):
    y = diag * x
    for k in range(6):
        y += off[k] * x[nbr[k]]
    return y


@pytest.fixture(scope="module")
def image384():
    img = generators.boolean_spheres((384, 384, 384), 6.0, 20.0, 0.70, SEED)[0]
    print(f"[configs] 384^3 Boolean image, porosity {1 - img.mean():.4f}", flush=True)
    return img


def test_precision_config_pressure_solve_contract(image384):
    n = image384.shape[0]
    g = G.NavierStokesGPU(image384, dx=1.0, mu=1.0, rho=1.0, p0=1.0, gradp=1.0 / n)
    try:
        g.init_pressure_linear()
        g.assemble_momentum(0.8)
        vinfo = g.compute_velocity_field(0.8, False, 0.0, 1e-10, 5000)
        g.compute_hat(0.8, O.SET_VEL, O.SET_VEL)
        g.assemble_pressure(O.PC_SIMPLEC, 0, 1.0, O.SET_VEL)
        g.set_precision(True, 1e-4)
        g.set("pc", np.zeros(g.np))
        info = g.compute_pressure_correction(1e-6, 5000)
        levels = g.multigrid_levels()
        diag, off, nbr, _ = g.matrix(O.SYS_P)
        b = g.get("b")
        x = g.get("pc")
        err = np.linalg.norm(b - spmv(diag, off, nbr, x)) / np.linalg.norm(b)
        print(f"[configs] 384^3: np {g.np}, velocity solves {vinfo}, pressure correction {info}, levels {levels}, "
              f"true relative residual {err:.3e}", flush=True)
        assert len(levels) >= 5 and levels[0] == g.np
        assert info[1] == 0 and info[2] <= 1e-6 and err <= 1.01e-6, (info, err)
    finally:
        g.close()


def test_precision_config_first_iterations_fp64_vs_mixed(image384):
    n = image384.shape[0]
    prm = dict(dx=1.0, mu=1.0, rho=1.0, p0=1.0, gradp=1.0 / n, alpha=0.8, beta=1.0, vsolvertol=1e-10,
               psolvertol=1e-10, pcsolvertol=1e-6, maxitsolver=5000, simpletol=1e-6, maxitSIMPLE=2, method="SIMPLEC",
               device=0)
    rows64, s64 = solver.solve(image384, **prm)
    rowsm, sm = solver.solve(image384, precision=1, inner_tol=1e-4, **prm)
    dk = abs(sm["K"] / s64["K"] - 1)
    print(f"[configs] 384^3 two outer iterations: K fp64 {s64['K']:.10e} mixed {sm['K']:.10e} |dK|/K {dk:.2e}; "
          f"seconds {s64['seconds']:.1f} / {sm['seconds']:.1f}; pressure iterations {s64['solver_work'][3]} / "
          f"{sm['solver_work'][3]}", flush=True)
    assert s64["iterations"] == sm["iterations"] == 2
    assert np.isfinite(sm["K"]) and dk <= 1e-4


def test_batching_config_batch_equals_solo():
    reqs = mix.request_mix(48, SEED, (32, 48, 64, 96, 128, 160))
    imgs = [im for im, _ in reqs]
    prm = dict(dx=1.0, mu=1.0, rho=1e-3, p0=1.0, alpha=0.8, beta=1.0, vsolvertol=1e-10, psolvertol=1e-10,
               pcsolvertol=1e-6, maxitsolver=5000, simpletol=1e-6, maxitSIMPLE=12, N1=4, N2=4, method="SIMPLEC",
               device=0, precision=1, inner_tol=1e-4)
    params = [dict(prm, gradp=1.0 / im.shape[0]) for im in imgs]
    batch = solver.solve_batch(imgs, params)
    fluid = [d["fluid"] for _, d in reqs]
    for q in (int(np.argmax(fluid)), int(np.argmin(fluid))):
        rows, s = solver.solve(imgs[q], **params[q])
        brows, bs = batch[q]
        print(f"[configs] batching image {q} {imgs[q].shape}: K batch {bs['K']!r} solo {s['K']!r}", flush=True)
        assert bs["K"] == s["K"] and bs["iterations"] == s["iterations"], q
        assert bs["solver_iterations"] == s["solver_iterations"], q
        assert np.array_equal(brows, rows), q
        _, s64 = solver.solve(imgs[q], **dict(params[q], precision=0))
        dk = abs(s["K"] / s64["K"] - 1)
        print(f"[configs] batching image {q}: mixed vs FP64 at {s['iterations']} outer iterations |dK|/K {dk:.2e}",
              flush=True)
        assert dk <= 1e-4
    assert all(np.isfinite(s["K"]) for _, s in batch)
