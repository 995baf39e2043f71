"""Multigrid sweeps of the mixed form's pressure systems (GPU).

The reference solves every system with Eigen's Jacobi BiCGSTAB (NSBase.h:97-100)
and the plain FP64 form keeps that.  In the mixed form, the FP32 sweeps of
the pressure-type systems inside the FP64 defect correction are flexible CG
preconditioned by one aggregation-multigrid V-cycle.  What pins them:

* the hierarchy is the aggregation it claims to be: the rows of level l+1
  are the distinct 2x2x2 blocks of level l's cells, in C order, and the
  level's couplings equal the Galerkin product P^T A P of level l (P the
  piecewise-constant prolongation), formed here with a generic sparse
  product in FP64 from the level-l matrix the device holds (level 0: the
  assembled pressure matrix rounded to FP32);
* every pressure solve meets the FP64 true-residual contract |b - A x| <=
  tol |b|, checked on the host against the assembled matrix;
* whole permeability solves agree with the plain FP64 path to |dK|/K <=
  1e-6 and fields to 1e-6 relative L2 (both converge to the same fixed point
  at these tolerances; the defect correction makes each linear solve an
  FP64-accurate one).  The pressure is compared on the fluid connected to
  the inlet or the outlet: in a pocket connected to neither, the pressure
  equations are a pure Neumann block whose solution is fixed only up to a
  constant, so each solver leaves its own constant there (the velocity is
  zero in such a pocket and K does not see it);
* graph replays and kernel-by-kernel launches give the same bits, and a
  served request gives the same bits as the same request alone.
"""
from __future__ import annotations

import os
import time

import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as O
from helpers import rel_l2
from ns3d_amd import generators
from ns3d_amd import gpu as G
from ns3d_amd import solver

pytestmark = pytest.mark.gpu

PARAMS = dict(dx=1.0, dy=-1, dz=-1, mu=1.0, rho=1e-3, p0=1.0)
WHOLE = dict(dx=1.0, mu=1.0, rho=1e-3, p0=1.0, alpha=0.8, beta=1.0, vsolvertol=1e-10, psolvertol=1e-10,
             pcsolvertol=1e-6, maxitsolver=5000, simpletol=1e-6, maxitSIMPLE=600, N1=4, N2=4, device=0)


def spmv(diag, off, nbr, x):
    y = diag * x
    for k in range(6):
        y = y + off[k] * x[nbr[k]]
    return y


def image(shape=(34, 30, 28), seed=7, porosity=0.6):
    return generators.rsa_spheres(shape, 3.5, 1.0 - porosity, seed)[0]


def prepared(img):
    g = G.NavierStokesGPU(img, gradp=1.0 / img.shape[0], **PARAMS)
    g.init_pressure_linear()
    g.assemble_momentum(0.8)
    g.compute_velocity_field(0.8, False, 0.0, 1e-10, 5000)
    g.compute_hat(0.8, O.SET_VEL, O.SET_VEL)
    g.assemble_pressure(O.PC_SIMPLEC, 0, 1.0, O.SET_VEL)
    return g


def connected_rows(raw, direction=1):
    """Mask over the pressure rows (C order of the prepared image) of the
    fluid 6-connected to the inlet or outlet plane."""
    import scipy.ndimage as ndi
    img = G.prepare_image(raw, direction, 0)
    lab, _ = ndi.label(img == 0)
    keep = np.setdiff1d(np.union1d(np.unique(lab[0]), np.unique(lab[-1])), [0])
    with G.NavierStokesGPU(img) as g:
        pgrid = g.index()[0]
    cells = pgrid >= 0
    return np.isin(lab[cells], keep)


def sparse_of(diag, off, nbr):
    n = len(diag)
    rows = [np.arange(n)] + [np.arange(n)] * 6
    cols = [np.arange(n)] + [nbr[k] for k in range(6)]
    vals = [diag] + [off[k] for k in range(6)]
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    A.sum_duplicates()
    return A


def level_matrix(cells, coef, dims):
    """The symmetric 7-point matrix of a coarse level from its cells and its
    diagonal / E / N / T couplings (W, S, B are the neighbours' E, N, T)."""
    nx, ny, nz = dims
    n = len(cells)
    grid = -np.ones(nx * ny * nz, np.int64)
    grid[cells] = np.arange(n)
    i, j, k = cells // (ny * nz), (cells // nz) % ny, cells % nz
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [coef[0].astype(np.float64)]
    for s, (di, dj, dk) in zip((1, 2, 3), ((1, 0, 0), (0, 1, 0), (0, 0, 1))):
        ok = (i + di < nx) & (j + dj < ny) & (k + dk < nz)
        nb = np.full(n, -1)
        nb[ok] = grid[cells[ok] + di * ny * nz + dj * nz + dk]
        has = nb >= 0
        assert np.all(coef[s][~has] == 0), "a coupling without a neighbour"
        r, c, v = np.arange(n)[has], nb[has], coef[s][has].astype(np.float64)
        rows += [r, c]
        cols += [c, r]
        vals += [v, v]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def galerkin(A, cells, dims):
    """(coarse cells, coarse dims, P^T A P) for 2x2x2 aggregation."""
    nx, ny, nz = dims
    cd = ((nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2)
    i, j, k = cells // (ny * nz), (cells // nz) % ny, cells % nz
    cc = ((i >> 1) * cd[1] + (j >> 1)) * cd[2] + (k >> 1)
    ucells = np.unique(cc)
    agg = np.searchsorted(ucells, cc)
    P = sp.csr_matrix((np.ones(len(cells)), (np.arange(len(cells)), agg)), shape=(len(cells), len(ucells)))
    return ucells, cd, (P.T @ A @ P).tocsr()


def test_hierarchy_is_galerkin_aggregation():
    img = image()
    g = prepared(img)
    g.set_precision(True, 1e-4)
    g.set_multigrid(1000)
    info = g.compute_pressure_correction(1e-6, 5000)
    rows = g.multigrid_levels()
    print(f"[multigrid] levels {rows}, pressure solve {info}")
    assert len(rows) >= 3 and rows[0] == g.np and rows[-1] <= 512
    assert all(a > b for a, b in zip(rows, rows[1:]))
    diag, off, nbr, _ = g.matrix(O.SYS_P)
    # level 0 as the device holds it: the FP32 pressure matrix
    A = sparse_of(diag.astype(np.float32).astype(np.float64), off.astype(np.float32).astype(np.float64), nbr)
    pgrid = g.index()[0].reshape(-1)
    cells = np.nonzero(pgrid >= 0)[0]
    assert np.array_equal(pgrid[cells], np.arange(g.np))
    dims = img.shape
    for lev in range(1, len(rows)):
        ucells, dims, Ac = galerkin(A, cells, dims)
        dcell, dcoef = g.multigrid_level(lev)
        assert np.array_equal(dcell, ucells), lev
        Ad = level_matrix(dcell.astype(np.int64), dcoef, dims)
        # the device sums in FP64 and rounds to FP32 once
        scale = abs(Ac).max()
        err = abs(Ad - Ac).max() / scale
        assert err <= 2e-7, (lev, err)
        A, cells = Ad, dcell.astype(np.int64)


@pytest.mark.parametrize("tol", [1e-5, 1e-8, 1e-11])
def test_multigrid_pressure_solve_true_residual(tol):
    img = image(seed=11)
    g = prepared(img)
    diag, off, nbr, _ = g.matrix(O.SYS_P)
    b = g.get("b")
    g.set("pc", np.zeros(g.np))
    g.set_precision(True, 1e-4)
    g.set_multigrid(1000)
    info = g.compute_pressure_correction(tol, 20000)
    assert len(g.multigrid_levels()) >= 3
    x = g.get("pc")
    err = np.linalg.norm(b - spmv(diag, off, nbr, x)) / np.linalg.norm(b)
    print(f"[multigrid] tol {tol:g}: info {info}, true relative residual {err:.3e}")
    assert info[1] == 0 and info[2] <= tol and err <= 1.01 * tol, (info, err)
    # far fewer sweeps than the Jacobi BiCGSTAB the plain form runs
    g.set("pc", np.zeros(g.np))
    g.set_precision(False)
    ref = g.compute_pressure_correction(tol, 20000)
    assert ref[1] == 0 and info[0] * 4 < ref[0], (info, ref)


@pytest.mark.parametrize("method", ["SIMPLEC", "SIMPLER"])
def test_whole_solve_with_multigrid_matches_fp64(method):
    raw = image((40, 34, 30), seed=5, porosity=0.55)
    prm = dict(WHOLE, gradp=1.0 / raw.shape[0], method=method)
    _, s64 = solver.solve(raw, **prm)
    _, smg = solver.solve(raw, precision=1, inner_tol=1e-4, mg_min_rows=1000, **prm)
    _, sjb = solver.solve(raw, precision=1, inner_tol=1e-4, mg_min_rows=0, **prm)
    assert s64["converged"] == 1 and smg["converged"] == 1
    dk = abs(smg["K"] / s64["K"] - 1)
    f64, fm = s64["fields"], smg["fields"]
    errs = {n: rel_l2(fm[n], f64[n]) for n in "uvw"}
    conn = connected_rows(raw)
    assert conn.sum() > 0.9 * len(conn)
    errs["p"] = rel_l2(fm["p"][conn], f64["p"][conn])
    print(f"[multigrid] {method}: |dK|/K {dk:.2e}, fields {errs}, outer its {s64['iterations']} / "
          f"{smg['iterations']}, pressure iterations fp64 {s64['solver_work'][3]} mixed+mg "
          f"{smg['solver_work'][3]} mixed jacobi {sjb['solver_work'][3]}")
    assert dk <= 1e-6
    assert all(e <= 1e-6 for e in errs.values()), errs
    assert smg["solver_work"][3] * 3 < sjb["solver_work"][3]


def test_graph_replay_equals_direct_launches():
    raw = image((36, 30, 30), seed=9)
    prm = dict(WHOLE, gradp=1.0 / raw.shape[0], method="SIMPLEC", maxitSIMPLE=40, precision=1, mg_min_rows=1000)
    out = []
    for graphs in ("1", "0", "1"):
        os.environ["NS3D_MG_GRAPHS"] = graphs
        try:
            rows, s = solver.solve(raw, **prm)
        finally:
            os.environ.pop("NS3D_MG_GRAPHS", None)
        out.append((rows, s))
    for rows, s in out[1:]:
        assert np.array_equal(rows, out[0][0])
        assert s["K"] == out[0][1]["K"] and s["solver_work"] == out[0][1]["solver_work"]
        for n in "puvw":
            assert np.array_equal(s["fields"][n], out[0][1]["fields"][n]), n


def test_served_multigrid_requests_equal_solo():
    imgs = [image((36, 30, 28), seed=21), image((30, 34, 32), seed=22, porosity=0.7),
            generators.boolean_spheres((28, 32, 30), 2.0, 4.0, 0.65, 23)[0]]
    reqs = [(im, dict(WHOLE, gradp=1.0 / im.shape[0], method=m, maxitSIMPLE=80, precision=1, mg_min_rows=1000))
            for im, m in zip(imgs, ("SIMPLEC", "SIMPLER", "IDEAL"))]
    solo = [solver.solve(im, **p)[1] for im, p in reqs]
    srv = solver.Server(device=0, budget_bytes=64e9, max_slots=3, chunk=4)
    try:
        tickets = []
        for im, p in reqs:
            tickets.append(srv.submit(im, **p))
            time.sleep(0.02)
        served = [srv.wait(t) for t in tickets]
    finally:
        srv.close()
    for q, (a, b) in enumerate(zip(solo, served)):
        assert b["fallback"] == 0, q
        assert a["K"] == b["K"] and a["iterations"] == b["iterations"], (q, a["K"], b["K"])
        assert a["solver_iterations"] == b["solver_iterations"], q


def test_footprint_bounds_multigrid_context():
    """The admission projection covers the hierarchy, built at the first
    pressure solve of a mixed context whose system reaches the default
    threshold."""
    img = generators.rsa_spheres((72, 66, 60), 4.0, 0.2, 31)[0]
    nfluid = int((img == 0).sum())
    g = prepared(img)
    try:
        g.set_precision(True, 1e-4)
        assert g.np >= 200000, g.np
        g.compute_pressure_correction(1e-6, 5000)
        assert len(g.multigrid_levels()) >= 4
        used = g.device_bytes()
        bound = G.footprint(img.shape, nfluid, True)
        print(f"[multigrid] footprint: used {used / 1e6:.1f} MB, bound {bound / 1e6:.1f} MB")
        assert used <= bound <= 1.5 * used + 8e6, (used, bound)
    finally:
        g.close()
