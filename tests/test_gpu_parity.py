"""HIP path vs the CPU oracle, method by method on identical inputs (GPU).

Bar: index arrays and every assembled coefficient bit-identical (integer /
FP64 with the same expression trees); reductions and solves within the
tolerances stated per test.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as O
from helpers import (assert_same_matrix, bit_equal, edge_images, random_fields, rel_l2, small_image)
from ns3d_amd import generators
from ns3d_amd import gpu as G

pytestmark = pytest.mark.gpu

PARAMS = dict(dx=1.3, dy=-1, dz=-1, mu=0.7, rho=2.1, p0=5.0, gradp=0.25, alpha=0.8, beta=1.0)


def pair(img, **kw):
    prm = dict(PARAMS)
    prm.update(kw)
    o = O.Oracle(img, **prm)
    g = G.NavierStokesGPU(img, **prm)
    assert (g.np, g.nu, g.nv, g.nw) == (o.np, o.nu, o.nv, o.nw)
    assert g.p0 == o.p0 and g.p1 == o.p1
    return o, g


def set_both(o, g, fields):
    for k, x in fields.items():
        o.set(k, x)
        g.set(k, x)


def vel_fields(o, seed):
    return random_fields({"u": o.nu, "v": o.nv, "w": o.nw, "ut": o.nu, "vt": o.nv, "wt": o.nw,
                          "uh": o.nu, "vh": o.nv, "wh": o.nw, "p": o.np, "pc": o.np}, seed=seed)


# ---------------------------------------------------------------- index scan
@pytest.mark.parametrize("name", sorted(edge_images()))
def test_index_edge_cases_bit_exact(name):
    img = edge_images()[name]
    o = O.Oracle(img)
    g = G.NavierStokesGPU(img)
    for a, b in zip(o.index(), g.index()):
        assert np.array_equal(a, b), name


@pytest.mark.parametrize("case", ["lattice64", "lattice128", "rsa48", "dir2", "dir3", "border3"])
def test_index_arrays_bit_exact(case):
    if case == "lattice64":
        img = generators.sphere_lattice(64, 8, 3)
    elif case == "lattice128":
        img = generators.sphere_lattice(128, 12, 5)
    elif case == "rsa48":
        img = generators.rsa_spheres((48, 40, 36), 5, 0.35, 11)[0]
    else:
        base = small_image((23, 19, 17), seed=5)
        d = {"dir2": (2, 0), "dir3": (3, 0), "border3": (1, 3)}[case]
        img = G.prepare_image(base, *d)
        assert np.array_equal(img, O.prepare_image(base, *d))
    o = O.Oracle(img)
    g = G.NavierStokesGPU(img)
    for a, b in zip(o.index(), g.index()):
        assert np.array_equal(a, b), case
    if case == "lattice128":
        assert (g.np, g.nu, g.nv, g.nw) == (1582152, 1517536, 1484768, 1484768)


# ----------------------------------------------------------------- assembly
@pytest.mark.parametrize("seed,shape", [(1, (20, 17, 13)), (2, (9, 11, 1)), (3, (16, 16, 16))])
def test_momentum_assembly_bit_exact(seed, shape):
    img = small_image(shape, seed=seed)
    o, g = pair(img)
    set_both(o, g, vel_fields(o, seed))
    o.set_alpha_beta(0.7, 1.0)
    o.assemble_momentum()
    g.assemble_momentum(0.7)
    for sys in (O.SYS_U, O.SYS_V, O.SYS_W):
        if [o.nu, o.nv, o.nw][sys] == 0:
            continue
        assert_same_matrix(o.matrix(sys), g.matrix(sys), what=f"sys {sys}")
        cnt, col, val = o.matrix(sys)
        rs = g.matrix(sys)[3]
        # Eigen's row(r).sum(): stored entries, ascending column, from 0
        for r in range(0, cnt.size, max(1, cnt.size // 50)):
            acc = 0.0
            for e in range(cnt[r]):
                acc += val[r, e]
            assert rs[r] == acc


@pytest.mark.parametrize("kind,relax", [(O.PRESSURE_FULL, 1), (O.PRESSURE_FULL, 0), (O.PC_SIMPLE, 0),
                                        (O.PC_SIMPLEC, 0)])
def test_pressure_assembly_bit_exact(kind, relax):
    img = small_image((18, 15, 14), seed=4)
    o, g = pair(img)
    set_both(o, g, vel_fields(o, 4))
    o.set_alpha_beta(0.8, 0.7)
    o.assemble_momentum()
    g.assemble_momentum(0.8)
    vel_set = O.SET_HAT if kind == O.PRESSURE_FULL else O.SET_VEL
    o.assemble_pressure(kind, relax, vel_set)
    g.assemble_pressure(kind, relax, 0.7, vel_set)
    assert_same_matrix(o.matrix(O.SYS_P), g.matrix(O.SYS_P), what=f"p kind {kind}")
    assert bit_equal(o.get("b"), g.get("b"))


@pytest.mark.parametrize("uc_set,uold_set", [(O.SET_TEMP, O.SET_VEL), (O.SET_VEL, O.SET_TEMP),
                                             (O.SET_VEL, O.SET_VEL)])
def test_hat_bit_exact(uc_set, uold_set):
    img = small_image((17, 16, 15), seed=6)
    o, g = pair(img)
    set_both(o, g, vel_fields(o, 6))
    o.set_alpha_beta(0.75, 1.0)
    o.assemble_momentum()
    g.assemble_momentum(0.75)
    o.compute_hat(uc_set, uold_set)
    g.compute_hat(0.75, uc_set, uold_set)
    for name in ("uh", "vh", "wh"):
        assert bit_equal(o.get(name), g.get(name)), name


@pytest.mark.parametrize("kind,target", [(O.CORR_SIMPLE, O.SET_VEL), (O.CORR_SIMPLEC, O.SET_VEL),
                                         (O.CORR_SIMPLER, O.SET_TEMP), (O.CORR_SIMPLER, O.SET_VEL)])
def test_velocity_correction_bit_exact(kind, target):
    img = small_image((19, 14, 13), seed=8)
    o, g = pair(img)
    f = vel_fields(o, 8)
    set_both(o, g, f)
    o.assemble_momentum()
    g.assemble_momentum(PARAMS["alpha"])
    dv = 2e-4  # small enough that the clamp is active on some faces
    o.correct_velocity(kind, dv, target, o.p0, o.p1)
    g.correct_velocity(kind, dv, target, o.p0, o.p1)
    names = {O.SET_VEL: ("u", "v", "w"), O.SET_TEMP: ("ut", "vt", "wt")}[target]
    for name in names:
        assert bit_equal(o.get(name), g.get(name)), name


def test_pressure_correction_update():
    img = small_image((15, 14, 13), seed=9)
    o, g = pair(img)
    set_both(o, g, vel_fields(o, 9))
    a = o.correct_pressure()
    o.set_alpha_beta(0.8, 1.0)
    b = g.correct_pressure(1.0)
    assert bit_equal(o.get("p"), g.get("p"))
    assert a[0] == b[0] and a[1] == b[1]
    assert abs(a[2] / b[2] - 1) < 1e-14


@pytest.mark.parametrize("border", [0, 2])
def test_divergence_and_permeability(border):
    img = G.prepare_image(small_image((16, 15, 14), seed=10), 1, border)
    o, g = pair(img, border=border)
    set_both(o, g, vel_fields(o, 10))
    do, dg = o.compute_div(), g.compute_div()
    assert do[0] == dg[0] and do[3] == dg[3]
    assert abs(do[1] / dg[1] - 1) < 1e-13 and abs(do[2] / dg[2] - 1) < 1e-13
    po, pg = o.compute_permeability(), g.compute_permeability()
    assert np.allclose(po, pg, rtol=1e-13, atol=0)


def test_linear_pressure_profile():
    img = small_image((14, 12, 11), seed=12)
    o, g = pair(img)
    o.init_pressure_linear()
    g.init_pressure_linear()
    assert bit_equal(o.get("p"), g.get("p"))


# ---------------------------------------------------------------- BiCGSTAB
def test_bicgstab_stencil_vs_dense_and_oracle():
    rng = np.random.default_rng(3)
    img = small_image((12, 11, 10), seed=13)
    o, g = pair(img)
    set_both(o, g, vel_fields(o, 13))
    o.assemble_momentum()
    g.assemble_momentum(PARAMS["alpha"])
    o.assemble_pressure(O.PC_SIMPLEC, 0, O.SET_VEL)
    g.assemble_pressure(O.PC_SIMPLEC, 0, 1.0, O.SET_VEL)
    diag, off, nbr, _ = g.matrix(O.SYS_P)
    n = diag.size
    A = np.zeros((n, n))
    A[np.arange(n), np.arange(n)] = diag
    for d in range(6):
        m = nbr[d] != np.arange(n)
        A[np.arange(n)[m], nbr[d][m]] += off[d][m]
    b = rng.standard_normal(n)
    x, info = G.bicgstab_stencil(diag, off, nbr, b, np.zeros(n), 1e-12, 5000)
    ref = np.linalg.solve(A, b)
    assert info[1] == 0
    assert rel_l2(x, ref) < 1e-9
    cnt, col, val = o.matrix(O.SYS_P)
    xo, io = O.bicgstab_ell(cnt, col, val, b, np.zeros(n), 1e-12, 5000)
    assert abs(info[0] - io[0]) <= max(2, 0.05 * io[0]), (info, io)
    # zero right-hand side: x = 0, iterations = maxit (Eigen contract)
    x, info = G.bicgstab_stencil(diag, off, nbr, np.zeros(n), rng.standard_normal(n), 1e-5, 33)
    assert np.all(x == 0) and info == (33, 0, 1e-5)
    # iteration cap -> NoConvergence
    x, info = G.bicgstab_stencil(diag, off, nbr, b, np.zeros(n), 1e-15, 2)
    assert info[0] == 2 and info[1] == 2


def test_velocity_solve_matches_oracle():
    img = small_image((20, 18, 16), seed=14)
    o, g = pair(img)
    f = vel_fields(o, 14)
    set_both(o, g, {k: f[k] for k in ("u", "v", "w", "p")})
    o.assemble_momentum()
    g.assemble_momentum(PARAMS["alpha"])
    o.set_params(**dict(PARAMS, vsolvertol=1e-12, maxitsolver=2000))
    io = o.compute_velocity_field(True, 5e-4)
    ig = g.compute_velocity_field(PARAMS["alpha"], True, 5e-4, 1e-12, 2000)
    for q, name in enumerate(("u", "v", "w")):
        assert ig[q][1] == 0 and io[q][1] == 0
        assert rel_l2(g.get(name), o.get(name)) < 1e-10, name


def test_fixture_initial_permeability():
    """Initial solve of the reference's 128^3 SIMPLEC run: u in 11
    iterations, v and w through the zero-RHS path (500 iterations, status 0),
    K = 9.14196330389082e-13 (INFO-1-spheres_128.log:13-16)."""
    img = generators.sphere_lattice(128, 12, 5)
    g = G.NavierStokesGPU(img, dx=2e-6, mu=1e-3, rho=1e3, p0=1e5, gradp=1e2)
    g.init_pressure_linear()
    g.assemble_momentum(0.8)
    info = g.compute_velocity_field(0.8, False, 0.0, 1e-5, 500)
    assert [i[0] for i in info] == [11, 500, 500] and all(i[1] == 0 for i in info)
    K = g.compute_permeability()[2]
    assert abs(K / 9.14196330389082e-13 - 1) < 1e-9


@pytest.mark.parametrize("kind", [O.PC_SIMPLEC, O.PRESSURE_FULL])
def test_packed_symmetric_solve_matches_indexed_layout(kind):
    """The context solves its pressure system from packed neighbour distances
    and the upper coefficients only; the stand-alone entry point solves the
    same matrix from the full int32-indexed storage.  Same arithmetic, same
    reduction order: the results must agree bit for bit."""
    img = small_image((22, 19, 17), seed=8)
    o, g = pair(img)
    f = vel_fields(o, 8)
    for k in ("u", "v", "w", "uh", "vh", "wh", "p"):
        g.set(k, f[k])
    g.assemble_momentum(0.8)
    g.assemble_pressure(kind, 1, 0.7, O.SET_HAT if kind == O.PRESSURE_FULL else O.SET_VEL)
    diag, off, nbr, _ = g.matrix(O.SYS_P)
    n = g.np
    rows = np.arange(n)
    for a, bdir in ((0, 1), (2, 3), (4, 5)):  # E-W, N-S, T-B pairs are transposes
        has = nbr[bdir] != rows
        assert bit_equal(off[bdir][has], off[a][nbr[bdir][has]])
    b = g.get("b")
    x0 = 1e-3 * np.random.default_rng(8).standard_normal(n)
    if kind == O.PRESSURE_FULL:
        g.set("p", x0)
        info = g.compute_pressure_field(1e-10, 5000)
        x_ctx = g.get("p")
    else:
        g.set("pc", x0)
        info = g.compute_pressure_correction(1e-10, 5000)
        x_ctx = g.get("pc")
    x_ind, info_ind = G.bicgstab_stencil(diag, off, nbr, b, x0, 1e-10, 5000)
    assert info[0] == info_ind[0] and info[1] == info_ind[1] == 0
    assert bit_equal(x_ctx, x_ind)
