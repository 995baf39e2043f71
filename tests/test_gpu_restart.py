"""BiCGSTAB's restart path (GPU).

Eigen restarts BiCGSTAB when |r0 . r| < eps^2 |r0|^2 (BiCGSTAB.h, restated in
kernels_solver.hip): r0 = r = b - A x again, rho = |r0|^2, and Eigen's
iteration counter is reset on the first restart only.  With eps the machine
epsilon this almost never fires, so contexts created while NS3D_RESTART_EPS2
is set use that factor instead, which drives solves through the restart
kernels and through the engine's restart bookkeeping (a restart found in
round k is queued with round k+1 or k+2 of the pipelined schedule).

Checks:
* a pressure solve restarts and still meets its tolerance: the true FP64
  relative residual on the host, and the solution against a direct sparse
  solve of the same matrix;
* the stand-alone stencil solve with restarts converges to the direct
  solution as the oracle's BiCGSTAB (Eigen's threshold: no restart) does;
* solves with restarts give the same bits batched on the engine as alone,
  with graph replay and launched kernel by kernel, FP64 and mixed.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle_lib as O
from helpers import rel_l2, small_image
from ns3d_amd import generators
from ns3d_amd import gpu as G
from ns3d_amd import solver

pytestmark = pytest.mark.gpu

EPS2 = "1e-3"  # restart once |r0 . r| < 1e-3 |r0|^2


class restarts_forced:
    def __enter__(self):
        self.old = os.environ.get("NS3D_RESTART_EPS2")
        os.environ["NS3D_RESTART_EPS2"] = EPS2

    def __exit__(self, *a):
        if self.old is None:
            del os.environ["NS3D_RESTART_EPS2"]
        else:
            os.environ["NS3D_RESTART_EPS2"] = self.old


def sparse_of(diag, off, nbr):
    n = len(diag)
    rows = np.concatenate([np.arange(n)] * 7)
    cols = np.concatenate([np.arange(n)] + [nbr[k] for k in range(6)])
    vals = np.concatenate([diag] + [off[k] for k in range(6)])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def test_pressure_solve_with_restarts_meets_tolerance():
    img = generators.rsa_spheres((30, 26, 24), 3.0, 0.35, 17)[0]
    with restarts_forced():
        g = G.NavierStokesGPU(img, dx=1.0, mu=1.0, rho=1.0, p0=1.0, gradp=1.0 / 30)
    try:
        g.init_pressure_linear()
        g.assemble_momentum(0.8)
        g.compute_velocity_field(0.8, False, 0.0, 1e-10, 5000)
        g.assemble_pressure(O.PC_SIMPLEC, 0, 1.0, O.SET_VEL)
        diag, off, nbr, _ = g.matrix(O.SYS_P)
        b = g.get("b")
        g.set("pc", np.zeros(g.np))
        info = G.SolveInfo()
        G._check(G.lib().ns3d_compute_pressure_correction(g.h, 1e-10, 20000, C.byref(info)))
        x = g.get("pc")
        A = sparse_of(diag, off, nbr)
        err = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        ref = spl.spsolve(A.tocsc(), b)
        print(f"[restart] pressure solve: iterations {info.iterations}, restarts {info.restarts}, true residual "
              f"{err:.2e}, vs direct {rel_l2(x, ref):.2e}")
        assert info.restarts >= 1 and info.info == 0
        assert err <= 2e-10 and rel_l2(x, ref) < 1e-7
    finally:
        g.close()


def test_stencil_solve_with_restarts_matches_direct_and_oracle():
    rng = np.random.default_rng(5)
    img = small_image((14, 12, 11), seed=23)
    g = G.NavierStokesGPU(img, dx=1.0, mu=1.0, rho=1.0, p0=1.0, gradp=1.0 / 14)
    try:
        g.init_pressure_linear()
        g.assemble_momentum(0.8)
        g.assemble_pressure(O.PC_SIMPLEC, 0, 1.0, O.SET_VEL)
        diag, off, nbr, _ = g.matrix(O.SYS_P)
    finally:
        g.close()
    n = diag.size
    b = rng.standard_normal(n)
    ref = spl.spsolve(sparse_of(diag, off, nbr).tocsc(), b)
    with restarts_forced():
        x, info = G.bicgstab_stencil(diag, off, nbr, b, np.zeros(n), 1e-12, 5000)
    x0, info0 = G.bicgstab_stencil(diag, off, nbr, b, np.zeros(n), 1e-12, 5000)
    # the oracle's Eigen BiCGSTAB on the same matrix (no restart at Eigen's threshold)
    # ELL rows: the diagonal and the existing neighbours, in ascending column order
    rows_c = np.stack([np.arange(n)] + [nbr[k] for k in range(6)], axis=1)
    rows_v = np.stack([diag] + [off[k] for k in range(6)], axis=1)
    keep = np.ones_like(rows_c, bool)
    keep[:, 1:] = rows_c[:, 1:] != np.arange(n)[:, None]
    cnt = keep.sum(1).astype(np.int32)
    col = np.zeros((n, 7), np.int32)
    val = np.zeros((n, 7))
    for r in range(n):
        c, v = rows_c[r][keep[r]], rows_v[r][keep[r]]
        o = np.argsort(c)
        col[r, :len(c)], val[r, :len(c)] = c[o], v[o]
    xo, io = O.bicgstab_ell(cnt, col, val, b, np.zeros(n), 1e-12, 5000)
    print(f"[restart] stencil: forced {info}, Eigen threshold {info0}, oracle {io}")
    assert info[1] == 0 and info0[1] == 0
    assert rel_l2(x, ref) < 1e-9 and rel_l2(x0, ref) < 1e-9 and rel_l2(xo, ref) < 1e-9
    assert info[0] > info0[0]  # the restarts cost iterations


@pytest.mark.parametrize("graphs", [True, False])
def test_batch_equals_solo_with_restarts(graphs):
    imgs = [generators.rsa_spheres((22, 20, 18), 3.0, 0.3, 31)[0],
            generators.boolean_spheres((18, 24, 20), 2.0, 4.0, 0.65, 32)[0],
            generators.fibres((26, 18, 20), 2.0, 10.0, 0.75, 33)[0]]
    prm = dict(dx=1.0, mu=1.0, rho=1.0, p0=1.0, alpha=0.8, beta=1.0, vsolvertol=1e-9, psolvertol=1e-9,
               pcsolvertol=1e-6, maxitsolver=4000, simpletol=1e-5, maxitSIMPLE=60, N1=4, N2=4, device=0)
    params = [dict(prm, gradp=1.0 / im.shape[0], method=m, precision=p)
              for im, m, p in zip(imgs, ("SIMPLEC", "SIMPLER", "IDEAL"), (0, 1, 0))]
    with restarts_forced():
        solo = [solver.solve(im, **p) for im, p in zip(imgs, params)]
        batch = solver.solve_batch(imgs, params, chunk=4, graphs=graphs)
    for q, ((r0, s0), (r1, s1)) in enumerate(zip(solo, batch)):
        print(f"[restart] image {q}: K {s0['K']!r} / {s1['K']!r}, solver iterations {s0['solver_iterations']}")
        assert s0["K"] == s1["K"] and s0["iterations"] == s1["iterations"], q
        assert s0["solver_iterations"] == s1["solver_iterations"], q
        assert np.array_equal(r0, r1), q
        for name in "puvw":
            assert np.array_equal(s0["fields"][name], s1["fields"][name]), (q, name)
