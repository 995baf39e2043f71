"""Plane-band order of the solver kernels (GPU).

Systems of at least NS3D_CHUNK_MIN_ROWS rows (4M by default: the 384^3 and
larger images) walk their rows as chunks of one plane of one j-band, band
by band per XCD, so that the one-plane-away gathers hit L2.  Here the
threshold is lowered so that small images take that order, and:
  * a whole solve of each method matches the oracle at the converged bar of
    the contiguous order (|dK|/K <= 1e-6, fields <= 1e-6 / 5e-6);
  * it agrees with the contiguous order to rounding;
  * batched on the engine it equals its own solo run bit for bit, in both
    precisions and with the graph path, as the contiguous order does.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle_lib as O
from helpers import connected_image, rel_l2
from ns3d_amd import generators
from ns3d_amd import gpu as G
from ns3d_amd import solver

pytestmark = pytest.mark.gpu

# Stokes limit (rho scaled down, pore Re ~ 1e-4): the outer iteration's fixed
# point is then reached without the rounding-decided relaxation of the
# convective case, so whole solves compare at the converged bar
TIGHT = dict(dx=1.0, mu=1.0, rho=1e-3, p0=1.0, alpha=0.8, beta=1.0, vsolvertol=1e-11, psolvertol=1e-11,
             pcsolvertol=1e-9, maxitsolver=5000)


class chunked:
    """Contexts created inside take the plane-band order (threshold lowered)."""

    def __init__(self, rows=1000):
        self.rows = str(rows)

    def __enter__(self):
        self.old = os.environ.get("NS3D_CHUNK_MIN_ROWS")
        os.environ["NS3D_CHUNK_MIN_ROWS"] = self.rows

    def __exit__(self, *a):
        if self.old is None:
            del os.environ["NS3D_CHUNK_MIN_ROWS"]
        else:
            os.environ["NS3D_CHUNK_MIN_ROWS"] = self.old


def image(seed=33, shape=(24, 22, 20)):
    return generators.rsa_spheres(shape, 3.5, 0.35, seed)[0]


@pytest.mark.parametrize("method", ["SIMPLEC", "SIMPLER", "SIMPLERR", "IDEAL"])
def test_chunked_solve_matches_oracle(method):
    img = image()
    prm = dict(TIGHT, gradp=1.0 / img.shape[0])
    with chunked():
        rows, summ = solver.solve(img, method=method, simpletol=1e-6, maxitSIMPLE=400, N1=4, N2=4, device=0, **prm)
    o = O.Oracle(img, **prm)
    orows, osumm = o.solve(method, 1e-6, 400, 4, 4)
    assert summ["converged"] == 1
    assert abs(summ["K"] / osumm["K"] - 1) <= 1e-6, (summ["K"], osumm["K"])
    f = summ["fields"]
    vel = rel_l2(np.concatenate([f[n] for n in "uvw"]), np.concatenate([o.get(n) for n in "uvw"]))
    assert vel <= 1e-6 and rel_l2(f["p"] - o.p1, o.get("p") - o.p1) <= 1e-6
    for n in "uvw":
        assert rel_l2(f[n], o.get(n)) <= 5e-6, n
    print(f"chunked {method}: |dK|/K {abs(summ['K'] / osumm['K'] - 1):.2e}, velocity {vel:.2e}")


def test_chunked_agrees_with_contiguous():
    img = image(34)
    prm = dict(TIGHT, gradp=1.0 / img.shape[0])
    _, s0 = solver.solve(img, method="SIMPLEC", simpletol=1e-6, maxitSIMPLE=400, device=0, **prm)
    with chunked():
        _, s1 = solver.solve(img, method="SIMPLEC", simpletol=1e-6, maxitSIMPLE=400, device=0, **prm)
    # the two orders differ in the summation order of the dot products only
    assert abs(s1["K"] / s0["K"] - 1) <= 1e-6
    f0, f1 = s0["fields"], s1["fields"]
    assert rel_l2(np.concatenate([f1[n] for n in "uvw"]), np.concatenate([f0[n] for n in "uvw"])) <= 1e-6
    assert rel_l2(f1["p"], f0["p"]) <= 1e-6


def test_chunked_matrix_stencil_solve_matches_dense():
    """One linear solve through the plane-band walk (the pressure system of a
    chunked context) against a dense solve of the exported matrix."""
    img = connected_image((20, 24, 18), 35)
    with chunked(500):
        g = G.NavierStokesGPU(img, dx=1.0, mu=1.0, rho=1.0, p0=1.0, gradp=1.0 / 20)
    with g:
        g.zero_fields()
        g.init_pressure_linear()
        g.assemble_momentum(0.8)
        g.compute_velocity_field(0.8, 0, 0.0, 1e-12, 5000)
        g.assemble_pressure(G.PC_SIMPLEC, 0, 1.0, G.SET_VEL)
        g.set("pc", np.zeros(g.np))
        info = g.compute_pressure_correction(1e-12, 5000)
        x = g.get("pc")
        diag, off, nbr = g.matrix(G.SYS_P)[:3]
        b = g.get("b")
    n = diag.size
    A = np.zeros((n, n))
    A[np.arange(n), np.arange(n)] = diag
    for d in range(6):
        m = nbr[d] != np.arange(n)
        A[np.nonzero(m)[0], nbr[d][m]] = off[d][m]
    xd = np.linalg.solve(A, b)
    assert info[0] > 0
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)


def test_chunked_batch_equals_solo_bit_for_bit():
    imgs = [image(36), generators.boolean_spheres((20, 26, 18), 2.0, 4.0, 0.6, 42)[0],
            generators.fibres((28, 20, 22), 2.0, 12.0, 0.75, 43)[0]]
    base = dict(dx=1.0, mu=1.0, rho=1.0, p0=1.0, alpha=0.8, beta=1.0, vsolvertol=1e-8, psolvertol=1e-8,
                pcsolvertol=1e-5, maxitsolver=3000, simpletol=1e-5, maxitSIMPLE=300, N1=4, N2=4, device=0)
    cases = [("SIMPLEC", 0), ("SIMPLER", 1), ("SIMPLEC", 1)]
    params = [dict(base, method=m, precision=q, inner_tol=1e-4, gradp=1.0 / im.shape[0])
              for im, (m, q) in zip(imgs, cases)]
    with chunked():
        solo = [solver.solve(im, **p) for im, p in zip(imgs, params)]
        for graphs in (True, False):
            batch = solver.solve_batch(imgs, params, chunk=4, graphs=graphs)
            for i, ((r0, s0), (r1, s1)) in enumerate(zip(solo, batch)):
                assert s0["converged"] == 1 and s0["K"] == s1["K"], (graphs, i)
                assert s0["solver_iterations"] == s1["solver_iterations"], (graphs, i)
                assert np.array_equal(r0, r1), (graphs, i)
                for f in "puvw":
                    assert np.array_equal(s0["fields"][f], s1["fields"][f]), (graphs, i, f)
