"""Mixed form: FP64 defect correction around FP32-storage BiCGSTAB sweeps
(GPU).  It has no counterpart in the reference (whose solves are FP64 Eigen
BiCGSTAB, NSBase.cxx:2285-2423); the bar is the solve contract itself: each
linear solve reaches the requested FP64 relative residual |b - A x| / |b|,
checked on the host in FP64 against the assembled matrix, and a whole
permeability solve agrees with the FP64 path to |dK|/K <= 1e-4."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as O
from helpers import rel_l2, small_image
from ns3d_amd import generators
from ns3d_amd import gpu as G
from ns3d_amd import solver

pytestmark = pytest.mark.gpu

PARAMS = dict(dx=1.3, dy=-1, dz=-1, mu=0.7, rho=2.1, p0=5.0, gradp=0.25)


def spmv(diag, off, nbr, x):
    """7-point product on the host; absent neighbours point at the row itself
    with a zero coefficient."""
    y = diag * x
    for k in range(6):
        y = y + off[k] * x[nbr[k]]
    return y


def prepared(seed, shape=(22, 19, 17)):
    img = small_image(shape, seed=seed)
    g = G.NavierStokesGPU(img, **PARAMS)
    rng = np.random.default_rng(seed)
    for name, n in (("u", g.nu), ("v", g.nv), ("w", g.nw)):
        g.set(name, 1e-3 * rng.standard_normal(n))
    g.init_pressure_linear()
    g.assemble_momentum(0.8)
    return g


@pytest.mark.parametrize("tol", [1e-6, 1e-10])
def test_mixed_velocity_solve_reaches_fp64_tolerance(tol):
    g = prepared(3)
    start = {k: g.get(k) for k in ("u", "v", "w")}
    ref = g.compute_velocity_field(0.8, False, 0.0, tol, 5000)
    x64 = {k: g.get(k) for k in ("u", "v", "w")}
    for k, x in start.items():
        g.set(k, x)
    g.set_precision(True, 1e-4)
    info = g.compute_velocity_field(0.8, False, 0.0, tol, 5000)
    for q, name in enumerate(("u", "v", "w")):
        assert ref[q][1] == 0 and info[q][1] == 0, (name, ref[q], info[q])
        assert info[q][2] <= tol, (name, info[q])
        diag, off, nbr, _ = g.matrix(q)
        x = g.get(name)
        # the right-hand side the solve saw: A x64 - r64, rebuilt from the FP64 solve
        b = spmv(diag, off, nbr, x64[name])
        err = np.linalg.norm(b - spmv(diag, off, nbr, x)) / np.linalg.norm(b)
        assert err <= tol * 1.01 + ref[q][2], (name, err)
        assert rel_l2(x, x64[name]) < 1e3 * tol, name


def test_mixed_pressure_solve_true_residual():
    g = prepared(5)
    g.compute_velocity_field(0.8, False, 0.0, 1e-10, 5000)
    g.compute_hat(0.8, O.SET_VEL, O.SET_VEL)
    g.assemble_pressure(O.PRESSURE_FULL, 1, 1.0, O.SET_HAT)
    diag, off, nbr, _ = g.matrix(O.SYS_P)
    b = g.get("b")
    p_start = g.get("p")
    for tol in (1e-4, 1e-8, 1e-11):
        g.set("p", p_start)
        g.set_precision(False)
        ref = g.compute_pressure_field(tol, 20000)
        p64 = g.get("p")
        g.set("p", p_start)
        g.set_precision(True, 1e-4)
        info = g.compute_pressure_field(tol, 20000)
        p = g.get("p")
        err = np.linalg.norm(b - spmv(diag, off, nbr, p)) / np.linalg.norm(b)
        assert info[1] == 0 and info[2] <= tol and err <= 1.01 * tol, (tol, info, err)
        err64 = np.linalg.norm(b - spmv(diag, off, nbr, p64)) / np.linalg.norm(b)
        assert ref[1] == 0 and err64 <= 1.01 * tol, (tol, ref, err64)


def test_mixed_zero_rhs_and_precision_switch():
    """A zero right-hand side gives x = 0 with iterations = maxit as in the
    FP64 path (NSBase.cxx:2285-2335 reads Eigen's report); switching back to
    FP64 restores the FP64 solver; bad modes are rejected."""
    img = np.zeros((6, 5, 4), np.uint8)
    g = G.NavierStokesGPU(img, **dict(PARAMS, p0=0.0, gradp=0.0))
    g.zero_fields()
    g.assemble_momentum(0.8)
    g.set_precision(True, 1e-3)
    info = g.compute_velocity_field(0.8, False, 0.0, 1e-8, 77)
    assert [i[0] for i in info] == [77, 77, 77]
    assert all(np.all(g.get(k) == 0) for k in ("u", "v", "w"))
    g.set_precision(False)
    info = g.compute_velocity_field(0.8, False, 0.0, 1e-8, 77)
    assert [i[0] for i in info] == [77, 77, 77]
    with pytest.raises(G.NS3DError):
        g.set_precision(True, 0.0)
    assert G.lib().ns3d_set_precision(g.h, 7, 1e-4) != 0


@pytest.mark.parametrize("method", ["SIMPLEC", "SIMPLER"])
def test_mixed_whole_solve_permeability(method):
    raw = generators.rsa_spheres((24, 22, 20), 3.5, 0.35, 41)[0]
    prm = dict(dx=1.0, mu=1.0, rho=1.0, p0=1.0, alpha=0.8, beta=1.0, vsolvertol=1e-10, psolvertol=1e-10,
               pcsolvertol=1e-6, maxitsolver=5000, gradp=1.0 / 24, simpletol=1e-6, maxitSIMPLE=400, N1=4, N2=4,
               device=0, method=method)
    rows64, s64 = solver.solve(raw, **prm)
    rowsm, sm = solver.solve(raw, precision=1, inner_tol=1e-4, **prm)
    assert s64["converged"] == 1 and sm["converged"] == 1
    dk = abs(sm["K"] / s64["K"] - 1)
    assert dk <= 1e-4, (sm["K"], s64["K"])
    f64, fm = s64["fields"], sm["fields"]
    vel = rel_l2(np.concatenate([fm[n] for n in "uvw"]), np.concatenate([f64[n] for n in "uvw"]))
    print(f"{method}: |dK|/K {dk:.2e}, velocity rel L2 {vel:.2e}, outer its fp64 {s64['iterations']} "
          f"mixed {sm['iterations']}")
    assert vel <= 1e-4
