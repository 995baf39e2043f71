"""The CPU oracle pinned against the reference's own fixtures and against
mathematics (CPU only).

Golden data: tests/golden/spheres128_simplec.json, extracted by
tests/golden/make_golden.py from the reference's data/ and tests/results/
(the 128^3 sphere-lattice SIMPLEC run of tests/test_spheres.ini).
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from ns3d_amd import generators

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "spheres128_simplec.json")))
FIXTURE_PARAMS = dict(dx=2e-6, dy=-1, dz=-1, mu=1e-3, rho=1e3, p0=1e5, gradp=1e2, alpha=0.8, beta=1.0,
                      vsolvertol=1e-5, psolvertol=1e-6, pcsolvertol=1e-3, deltaRMR=1e-4, maxitsolver=500)


@pytest.fixture(scope="module")
def lattice128():
    return generators.sphere_lattice(128, 12, 5)


@pytest.mark.parametrize("name", ["spheres_64", "spheres_128"])
def test_lattice_generator_rebuilds_reference_images(name):
    g = GOLD["lattice"][name]
    img = generators.sphere_lattice(*g["generator"])
    assert list(img.shape) == g["shape"]
    assert int(img.sum()) == g["solid_voxels"]
    assert hashlib.sha256(np.ascontiguousarray(img).tobytes()).hexdigest() == g["sha256"]


def test_index_counts_match_fixture(lattice128):
    o = O.Oracle(lattice128, **FIXTURE_PARAMS)
    info = GOLD["info"]
    assert (o.np, o.nu, o.nv, o.nw) == (info["np"], info["nu"], info["nv"], info["nw"])
    # NSBase.cxx:182 prints float(np)/float(nx*ny*nz) with 15 digits
    assert float("%.15g" % (np.float32(o.np) / np.float32(128 ** 3))) == info["porosity"]
    assert o.p0 == info["p0"] and float("%.15g" % o.p1) == info["p1"]


def _index_invariants(img, p, u, v, w):
    """Properties every valid numbering has, stated without the counting rule."""
    nx, ny, nz = img.shape
    fl = img == 0
    # u node between two fluid cells, or on a fluid inlet/outlet face
    uex = np.zeros((nx + 1, ny, nz), bool)
    uex[0] = fl[0]
    uex[nx] = fl[nx - 1]
    uex[1:nx] = fl[1:] & fl[:-1]
    vex = np.zeros_like(fl)
    vex[:, 1:] = fl[:, 1:] & fl[:, :-1]
    wex = np.zeros_like(fl)
    wex[:, :, 1:] = fl[:, :, 1:] & fl[:, :, :-1]
    nb = np.zeros_like(fl)
    nb[1:] |= fl[:-1]
    nb[:-1] |= fl[1:]
    nb[:, 1:] |= fl[:, :-1]
    nb[:, :-1] |= fl[:, 1:]
    nb[:, :, 1:] |= fl[:, :, :-1]
    nb[:, :, :-1] |= fl[:, :, 1:]
    nb[0] = True
    nb[nx - 1] = True
    pex = fl & nb
    for idx, ex in ((p, pex), (u, uex), (v, vex), (w, wex)):
        assert np.array_equal(idx >= 0, ex)
        vals = idx[idx >= 0]
        assert np.array_equal(np.sort(vals), np.arange(vals.size))
    # p, v, w are numbered in lexicographic order
    for idx in (p, v, w):
        vals = idx[idx >= 0]
        assert np.all(np.diff(vals) == 1)
    # u: numbered in visit order of the cell that creates it (the outlet face
    # u(nx) belongs to cell nx-1, right after u(nx-1))
    key = np.full(u.shape, -1, np.int64)
    lin = np.arange(nx * ny * nz).reshape(nx, ny, nz) * 2
    key[:nx] = lin
    key[nx] = lin[nx - 1] + 1
    order = np.argsort(key[u >= 0], kind="stable")
    assert np.array_equal(u[u >= 0][order], np.arange(order.size))


@pytest.mark.parametrize("seed", [1, 2])
def test_index_arrays_satisfy_invariants(seed):
    img = generators.boolean_spheres((15, 13, 11), 1.5, 3.0, 0.55, seed)[0]
    o = O.Oracle(img)
    _index_invariants(img, *o.index())


def test_index_rotation_and_border():
    img = generators.boolean_spheres((10, 12, 9), 1.5, 3.0, 0.6, 3)[0]
    r2 = O.prepare_image(img, 2, 0)
    assert r2.shape == (12, 10, 9)
    # temp(j, nx-1-i, k) = img(i, j, k)  (NSBase.cxx:2468)
    assert np.array_equal(r2, img[::-1].transpose(1, 0, 2))
    r3 = O.prepare_image(img, 3, 0)
    assert np.array_equal(r3, np.ascontiguousarray(img.transpose(2, 1, 0)))
    b = O.prepare_image(img, 1, 3)
    assert b.shape == (16, 12, 9) and not b[:3].any() and not b[-3:].any()
    for im in (r2, r3, b):
        _index_invariants(im, *O.Oracle(im).index())


def _ell_system(n, rng, offscale=0.15):
    """Random 7-point system: sorted ELL with a dominant diagonal."""
    cnt = np.zeros(n, np.int32)
    col = np.zeros((n, 7), np.int32)
    val = np.zeros((n, 7))
    A = np.zeros((n, n))
    for r in range(n):
        cols = {r}
        for c in rng.choice(n, size=min(6, n - 1), replace=False):
            if c != r and len(cols) < 7:
                cols.add(int(c))
        cs = sorted(cols)
        vs = [(-2.0 - rng.random()) if c == r else offscale * rng.standard_normal() for c in cs]
        cnt[r] = len(cs)
        col[r, :len(cs)] = cs
        val[r, :len(cs)] = vs
        A[r, cs] = vs
    return cnt, col, val, A


def test_bicgstab_solves_small_systems_like_dense_lu():
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 40, 300):
        cnt, col, val, A = _ell_system(n, rng)
        b = rng.standard_normal(n)
        x, info = O.bicgstab_ell(cnt, col, val, b, np.zeros(n), 1e-13, 1000)
        ref = np.linalg.solve(A, b)
        assert info[1] == 0, info
        assert np.linalg.norm(x - ref) <= 1e-10 * np.linalg.norm(ref)
        # the recursive residual can run below the true one's rounding floor
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= max(info[2], 1e-14) * 1.0000001


def test_bicgstab_eigen_contract():
    rng = np.random.default_rng(1)
    cnt, col, val, A = _ell_system(50, rng)
    # zero right-hand side: x = 0, iterations = maxit, error = tol, Success
    x, info = O.bicgstab_ell(cnt, col, val, np.zeros(50), rng.standard_normal(50), 1e-5, 77)
    assert np.all(x == 0) and info[0] == 77 and info[1] == 0 and info[2] == 1e-5
    # exact guess: zero iterations
    b = rng.standard_normal(50)
    xs = np.linalg.solve(A, b)
    x, info = O.bicgstab_ell(cnt, col, val, A @ xs, xs, 1e-6, 100)
    assert info[0] == 0
    # iteration cap: NoConvergence
    x, info = O.bicgstab_ell(cnt, col, val, b, np.zeros(50), 1e-15, 1)
    assert info[0] == 1 and info[1] == 2
    # SpMV sums the stored entries in ascending column order from 0
    y = O.spmv_ell(cnt, col, val, xs)
    for r in (0, 7, 49):
        acc = 0.0
        for e in range(cnt[r]):
            acc += val[r, e] * xs[col[r, e]]
        assert y[r] == acc


def test_simplec_trajectory_matches_fixture(lattice128):
    """First outer iterations of the reference's own 128^3 SIMPLEC run: solver
    iteration counts, K to 15 digits, pressure-correction statistics and RES
    rows as printed (tests/results/INFO-1-spheres_128.log, RES-1-*.csv)."""
    o = O.Oracle(lattice128, **FIXTURE_PARAMS)
    rows, summ = o.solve("SIMPLEC", 1e-4, 3)
    info = GOLD["info"]
    ks = [k["K"] for k in info["K"] if k["it"] is not None and not k["initial"]]
    for r in rows:
        it = int(r[0])
        assert abs(r[4] / ks[it] - 1) < 1e-11, (it, r[4], ks[it])
        solves = {s["comp"]: s["iterations"] for s in info["solves"] if s["it"] == it}
        if it > 0:
            assert (int(r[10]), int(r[11]), int(r[12])) == (solves["u"], solves["v"], solves["w"])
        gold_row = GOLD["res_rows"][it].split(",")
        mine = ["%g" % x for x in r[1:10]]
        assert mine == [g.strip() for g in gold_row[1:10]], (it, mine, gold_row)
        if it > 0:
            assert int(r[13]) == int(gold_row[13])
    assert abs(summ["K"] / info["K"][3]["K"] - 1) < 1e-11


@pytest.mark.skipif(not os.path.isdir("/root/reference/tests/results"), reason="reference fixtures absent")
def test_divergence_of_fixture_fields(lattice128):
    """ComputeDiv and ComputePermeability on the reference's final fields
    reproduce the logged final divergence norm (INFO-1-spheres_128.log:259)."""
    from ns3d_amd import vti
    o = O.Oracle(lattice128, **FIXTURE_PARAMS)
    p, u, v, w = o.index()
    res = "/root/reference/tests/results/"
    fu = vti.read_vti(res + "Vx-1-spheres_128.vti")
    fv = vti.read_vti(res + "Vy-1-spheres_128.vti")
    fw = vti.read_vti(res + "Vz-1-spheres_128.vti")
    for name, idx, f in (("u", u, fu), ("v", v, fv), ("w", w, fw)):
        x = np.zeros(idx.max() + 1)
        x[idx[idx >= 0]] = f[idx >= 0]
        o.set(name, x)
    md, normdiv, vnorm, maxdiv = o.compute_div()
    assert abs(normdiv / GOLD["info"]["final_normdiv"] - 1) < 1e-12
    q, vmoy, K, pin, pout = o.compute_permeability()
    assert abs(q / (GOLD["fields"]["Vx"]["outlet_sum"] * 4e-12) - 1) < 1e-14


@pytest.mark.parametrize("kind,relax", [(O.PRESSURE_FULL, 1), (O.PRESSURE_FULL, 0), (O.PC_SIMPLE, 0),
                                        (O.PC_SIMPLEC, 0)])
def test_pressure_matrix_is_bit_symmetric(kind, relax):
    """Both rows of a face take the same -1/(a_f h^2) (NSBase.cxx:478-851:
    each off-diagonal is written from the face's momentum diagonal or row
    sum), so the pressure matrix equals its transpose bit for bit.  The GPU
    solver relies on this: it reads only the diagonal and the E, N, T
    coefficients of the pressure system and takes W, S, B from the
    neighbours' E, N, T."""
    import scipy.sparse as sp
    from helpers import random_fields, small_image
    img = small_image((18, 15, 14), seed=11)
    o = O.Oracle(img, dx=1.3, dy=-1, dz=-1, mu=0.7, rho=2.1, p0=5.0, gradp=0.25, alpha=0.8, beta=1.0)
    f = random_fields({"u": o.nu, "v": o.nv, "w": o.nw, "uh": o.nu, "vh": o.nv, "wh": o.nw, "p": o.np},
                      seed=11)
    for k, x in f.items():
        o.set(k, x)
    o.set_alpha_beta(0.8, 0.7)
    o.assemble_momentum()
    o.assemble_pressure(kind, relax, O.SET_HAT if kind == O.PRESSURE_FULL else O.SET_VEL)
    cnt, col, val = o.matrix(O.SYS_P)
    n = len(cnt)
    rows = np.repeat(np.arange(n), cnt)
    mask = np.arange(7)[None, :] < cnt[:, None]
    A = sp.csr_matrix((val[mask], (rows, col[mask])), shape=(n, n))
    A.sort_indices()
    At = A.T.tocsr()
    At.sort_indices()
    assert np.array_equal(A.indptr, At.indptr) and np.array_equal(A.indices, At.indices)
    assert np.array_equal(A.data.view(np.uint64) & ~np.uint64(1 << 63) == 0, At.data.view(np.uint64) & ~np.uint64(1 << 63) == 0)
    nz = A.data != 0
    assert np.array_equal(A.data[nz].view(np.uint64), At.data[nz].view(np.uint64))
    # and the fixed point is not trivial: off-diagonals exist in all three axes
    assert A.nnz > 4 * n
