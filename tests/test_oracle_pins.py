"""Independent pins of the oracle's SIMPLER-family functions and of the
outer loop's relaxation control (CPU only).

Each check is chosen so that a plausible slip (a sign, a dropped term, the
wrong operand, a branch taken at the wrong time) fails it, and none restates
the oracle's own formula:

* Compute_*_hat (NSBase.cxx:1548-1755) against the SpMV of the assembled
  momentum matrix: u_hat = uc - (A uc) / diag(A) + (1 - alpha) u_old;
* AssemblePressure (FULL, :478-610) + an exact sparse solve +
  CorrectVelocityFieldSIMPLER (:2265-2335) make the velocity
  divergence-free (ComputeDiv, pinned by the reference's final fields);
  the same for the SIMPLE (:734-851, :2082-2172) and SIMPLEC (:612-732,
  :2174-2263) pressure-correction pairs;
* the beta relaxation of FULL keeps the unrelaxed solution as its fixed
  point, and moves towards it from any other p_old;
* ComputePressureField's tighten loop (:2383-2423) on a guess that meets
  psolvertol but not 1e-9, and on an exact guess;
* RelaxControl (:196-301) on scripted residual histories, with alpha and
  beta worked out by hand from the rules, for the oracle and the product's
  host drivers alike;
* the README's 64^3 algorithm comparison (README.md:177-192,
  tests/compare_algorithms.ini) for all four drivers.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle_lib as O
from helpers import connected_image, random_fields
from ns3d_amd import generators, solver


def csr(cnt, col, val):
    n = cnt.size
    mask = np.arange(7)[None, :] < cnt[:, None]
    rows = np.repeat(np.arange(n), cnt)
    return sp.csr_matrix((val[mask], (rows, col[mask])), shape=(n, n))


def prepared(seed=3, shape=(14, 12, 11), **kw):
    img = connected_image(shape, seed)
    prm = dict(dx=1.0, mu=1.0, rho=1.0, p0=1.0, gradp=1.0 / shape[0], alpha=0.7, beta=1.0)
    prm.update(kw)
    o = O.Oracle(img, **prm)
    o.set_alpha_beta(prm["alpha"], prm["beta"])
    # realistic velocities with both signs: random fields then one momentum solve
    f = random_fields({"u": o.nu, "v": o.nv, "w": o.nw}, seed=seed)
    for k in "uvw":
        o.set(k, f[k] + (0.02 if k == "u" else 0.0))
    o.init_pressure_linear()
    o.assemble_momentum()
    return o


# ---------------------------------------------------------------- hats
@pytest.mark.parametrize("alpha", [1.0, 0.7])
def test_hat_equals_spmv_identity(alpha):
    o = prepared(alpha=alpha)
    rng = np.random.default_rng(11)
    for sys, name in ((O.SYS_U, "u"), (O.SYS_V, "v"), (O.SYS_W, "w")):
        n = {"u": o.nu, "v": o.nv, "w": o.nw}[name]
        uc = rng.standard_normal(n)
        uold = rng.standard_normal(n)
        o.set(name + "t", uc)
        o.set(name, uold)
        o.compute_hat(O.SET_TEMP, O.SET_VEL)
        A = csr(*o.matrix(sys))
        d = A.diagonal()
        expect = uc - O.spmv_ell(*o.matrix(sys), uc) / d + (1.0 - alpha) * uold
        got = o.get(name + "h")
        scale = np.abs(uc).max() + np.abs(uold).max()
        assert np.abs(got - expect).max() <= 1e-12 * scale, name
        assert np.abs(A.diagonal()).min() > 0


# ------------------------------------------------- pressure + correction
def solve_p(o):
    A = csr(*o.matrix(O.SYS_P))
    b = o.get("b")
    return spl.spsolve(A.tocsc(), b), A, b


def div_ratio(o, vel_set):
    mdv, normdiv, vnorm, maxdiv = o.compute_div(vel_set)
    vel = {O.SET_VEL: "uvw", O.SET_TEMP: ("ut", "vt", "wt")}[vel_set]
    scale = max(np.abs(o.get(n)).max() for n in vel)
    return mdv / scale, normdiv


def test_full_pressure_and_simpler_correction_conserve_mass():
    o = prepared()
    o.set("ut", o.get("u")); o.set("vt", o.get("v")); o.set("wt", o.get("w"))
    o.compute_hat(O.SET_VEL, O.SET_VEL)
    o.assemble_pressure(O.PRESSURE_FULL, False, O.SET_HAT)
    p, A, b = solve_p(o)
    assert np.linalg.norm(A @ p - b) <= 1e-12 * np.linalg.norm(b)
    o.set("p", p)
    o.correct_velocity(O.CORR_SIMPLER, 0.0, O.SET_TEMP, o.p0, o.p1)
    before, _ = div_ratio(o, O.SET_VEL)
    after, normdiv = div_ratio(o, O.SET_TEMP)
    assert before > 1e-3          # the uncorrected field is far from solenoidal
    assert after < 1e-11, after   # the corrected one is, to rounding
    # the correction left the velocity nodes without a pressure gradient alone:
    # u_temp - u_hat is exactly -grad(p)/a on every node
    assert np.isfinite(normdiv)


def test_full_pressure_inlet_outlet_terms_matter():
    """The same check with the boundary pressures swapped in the correction
    fails: the p0/p1 terms of the assembly and of the correction pair up."""
    o = prepared()
    o.compute_hat(O.SET_VEL, O.SET_VEL)
    o.assemble_pressure(O.PRESSURE_FULL, False, O.SET_HAT)
    p, _, _ = solve_p(o)
    o.set("p", p)
    o.correct_velocity(O.CORR_SIMPLER, 0.0, O.SET_TEMP, o.p1, o.p0)
    bad, _ = div_ratio(o, O.SET_TEMP)
    assert bad > 1e-3


@pytest.mark.parametrize("kind,corr", [(O.PC_SIMPLE, O.CORR_SIMPLE), (O.PC_SIMPLEC, O.CORR_SIMPLEC)])
def test_pressure_correction_pairs_conserve_mass(kind, corr):
    o = prepared(seed=5)
    o.assemble_pressure(kind, False, O.SET_VEL)
    pc, A, b = solve_p(o)
    before, _ = div_ratio(o, O.SET_VEL)
    o.set("pc", pc)
    o.correct_velocity(corr, 1e300, O.SET_VEL)   # no clamp
    after, _ = div_ratio(o, O.SET_VEL)
    assert before > 1e-3
    assert after < 1e-11, after


def test_pressure_correction_pairs_are_not_interchangeable():
    """Pairing the SIMPLE correction (diagonals) with the SIMPLEC system (row
    sums) leaves a divergence: the two a_f choices are really different."""
    o = prepared(seed=5)
    o.assemble_pressure(O.PC_SIMPLEC, False, O.SET_VEL)
    pc, _, _ = solve_p(o)
    o.set("pc", pc)
    o.correct_velocity(O.CORR_SIMPLE, 1e300, O.SET_VEL)
    bad, _ = div_ratio(o, O.SET_VEL)
    assert bad > 1e-4


@pytest.mark.parametrize("beta", [0.9, 0.5])
def test_full_pressure_relaxation_fixed_point(beta):
    o = prepared(seed=7, beta=beta)
    o.set_alpha_beta(0.7, beta)
    o.compute_hat(O.SET_VEL, O.SET_VEL)
    o.assemble_pressure(O.PRESSURE_FULL, False, O.SET_HAT)
    pstar, _, _ = solve_p(o)
    # p_old = p*: the relaxed system has p* as its solution
    o.set("p", pstar)
    o.assemble_pressure(O.PRESSURE_FULL, True, O.SET_HAT)
    p1, _, _ = solve_p(o)
    assert np.abs(p1 - pstar).max() <= 1e-10 * np.abs(pstar).max()
    # p_old elsewhere: the relaxed solution lies strictly between p_old and p*
    # in the diagonal's metric, and beta = 1 would jump to p* at once
    pold = pstar + 0.05 * np.random.default_rng(3).standard_normal(pstar.size)
    o.set("p", pold)
    o.assemble_pressure(O.PRESSURE_FULL, True, O.SET_HAT)
    pr, _, _ = solve_p(o)
    e_old = np.linalg.norm(pold - pstar)
    e_new = np.linalg.norm(pr - pstar)
    assert 0.0 < e_new < e_old
    assert np.linalg.norm(pr - pold) > 1e-3 * e_old


def test_pressure_field_tighten_loop():
    o = prepared(seed=9, psolvertol=1e-6, maxitsolver=2000)
    o.compute_hat(O.SET_VEL, O.SET_VEL)
    o.assemble_pressure(O.PRESSURE_FULL, False, O.SET_HAT)
    pstar, A, b = solve_p(o)
    # a guess whose relative residual is about 3e-9: the solves at 1e-6,
    # 1e-7 and 1e-8 stop at once, the one at 1e-9 iterates
    dirn = np.random.default_rng(1).standard_normal(pstar.size)
    r1 = np.linalg.norm(A @ dirn) / np.linalg.norm(b)
    guess = pstar + (3e-9 / r1) * dirn
    res0 = np.linalg.norm(A @ guess - b) / np.linalg.norm(b)
    assert 1e-9 < res0 < 1e-8
    o.set("p", guess)
    it, info, err = o.compute_pressure_field()
    assert it > 0 and info == 0
    assert err <= 1e-9
    got = o.get("p")
    assert np.linalg.norm(A @ got - b) / np.linalg.norm(b) < res0
    # an exact guess: every tightened solve takes 0 iterations down to 1e-15,
    # and the pressure is left as it was, bit for bit
    o.set("p", pstar)
    res_star = np.linalg.norm(A @ pstar - b) / np.linalg.norm(b)
    if res_star < 1e-15:
        it, info, err = o.compute_pressure_field()
        assert it == 0
        assert np.array_equal(o.get("p"), pstar)


# ---------------------------------------------------------- RelaxControl
def replays(rmr, beta_cst):
    ao, bo = O.relax_replay(rmr, beta_cst=beta_cst)
    ap, bp, ev = solver.relax_replay(rmr, beta_cst=beta_cst)
    assert np.array_equal(ao, ap) and np.array_equal(bo, bp)  # oracle and product agree bit for bit
    return ap, bp, ev


def test_relax_decreasing_history_changes_nothing():
    a, b, ev = replays(1e-2 * 0.9 ** np.arange(45), False)
    assert np.all(a == 0.8) and np.all(b == 1.0) and np.all(ev == 0)


def test_relax_flat_history():
    # constant residual: the history first exceeds 30 entries at iteration
    # 30 (the window is flat and not decreasing): alpha -= 0.0125 and 15
    # entries are dropped; 16 more iterations fill it again (iteration 46)
    a, b, ev = replays(np.full(50, 1e-3), False)
    assert np.all(a[:30] == 0.8) and np.all(b[:30] == 1.0)
    assert a[30] == pytest.approx(0.7875, abs=1e-15) and b[30] == pytest.approx(0.9875, abs=1e-15)
    assert np.all(a[30:46] == a[30])
    assert a[46] == pytest.approx(0.775, abs=1e-15) and b[46] == pytest.approx(0.975, abs=1e-15)
    assert list(np.nonzero(ev)[0]) == [30, 46] and set(ev[[30, 46]]) == {2}
    # with beta held constant (SIMPLEC, SIMPLER, IDEAL) only alpha moves
    a2, b2, _ = replays(np.full(50, 1e-3), True)
    assert np.array_equal(a2, a) and np.all(b2 == 1.0)


def test_relax_jump():
    # a residual above 2.5x the best so far: alpha and beta -= 0.05 once
    rmr = np.array([1e-3] * 5 + [3e-3] + [1e-3] * 3)
    a, b, ev = replays(rmr, False)
    assert np.all(a[:5] == 0.8) and a[5] == pytest.approx(0.75, abs=1e-15)
    assert b[5] == pytest.approx(0.95, abs=1e-15) and np.all(a[5:] == a[5])
    assert list(np.nonzero(ev)[0]) == [5] and ev[5] == 1
    # 2.4x is not a jump
    a, _, ev = replays(np.array([1e-3] * 5 + [2.4e-3] + [1e-3] * 3), False)
    assert np.all(a == 0.8) and np.all(ev == 0)


def test_relax_jumps_clamp_then_flat_window_has_no_clamp():
    # 16 jumps in a row (each raises the best residual by 10%, so 1.0 keeps
    # jumping): alpha 0.8 -> 0.75 -> ... -> 0.05 -> clamped at 0.001
    rmr = np.array([1e-6] * 3 + [1.0] * 16 + [1e-6] * 40)
    a, b, ev = replays(rmr, False)
    assert np.all(ev[3:19] & 1) and not np.any(ev[:3])
    assert a[17] == pytest.approx(0.05, abs=1e-12)
    assert a[18] == 0.001 and b[18] == pytest.approx(0.2, abs=1e-12)  # beta started at 1.0
    # then a flat history: the window holds only the small residuals from
    # iteration 48 on, and the flat branch has no lower clamp (NSBase.cxx:263)
    assert np.all(a[18:48] == 0.001) and not np.any(ev[19:48])
    assert ev[48] == 2
    assert a[48] == pytest.approx(0.001 - 0.0125, abs=1e-15)
    assert b[48] == pytest.approx(0.2 - 0.0125, abs=1e-12)


def test_relax_rising_window():
    rmr = 1e-3 * 1.02 ** np.arange(31)
    a, b, ev = replays(rmr, False)
    assert np.all(a[:30] == 0.8)
    assert ev[30] == 4 and a[30] == pytest.approx(0.75, abs=1e-15) and b[30] == pytest.approx(0.95, abs=1e-15)


# -------------------------------------------- README algorithm comparison
README_64 = {  # README.md:183-186: K (m^2), outer iterations (4 threads, Eigen)
    "SIMPLEC": (3.28386e-12, 48),
    "SIMPLER": (3.28382e-12, 42),
    "SIMPLERR": (3.28175e-12, 23),
    "IDEAL": (3.28212e-12, 19),
}
COMPARE_INI = dict(dx=2e-6, mu=1e-3, rho=1e3, p0=1e5, gradp=1e2, alpha=0.8, beta=1.0, vsolvertol=1e-5,
                   psolvertol=1e-6, pcsolvertol=1e-3, deltaRMR=1e-4, maxitsolver=500)


@pytest.fixture(scope="module")
def lattice64():
    return generators.sphere_lattice(64, 8, 3)


@pytest.mark.parametrize("method", list(README_64))
def test_readme_algorithm_comparison(lattice64, method):
    """tests/compare_algorithms.ini on data/spheres_64.vti (regenerated bit for
    bit): K must round to the README's 6 printed digits and the outer
    iteration count must be the README's."""
    K_ref, it_ref = README_64[method]
    o = O.Oracle(lattice64, **COMPARE_INI)
    rows, s = o.solve(method, 1e-4, 200, N1=2, N2=2)
    print(f"{method}: K {s['K']:.6e} (README {K_ref:.5e}), its {s['iterations']} (README {it_ref})")
    assert float(f"{s['K']:.5e}") == K_ref
    assert s["iterations"] == it_ref


# ------------------------------------------------ one fixed point for all
def test_four_methods_share_one_fixed_point():
    """All four couplings discretise the same equations, so they converge to
    one fixed point.  On this 16x14x12 image the outer residual's floor is
    ~1e-7 (SIMPLEC at simpletol 1e-12 never stops, and ends 4e-9 from its
    1e-10 stop in K), which bounds the field agreement to ~1e-6.  SIMPLERR
    reaches the point by iteration ~25 and then drifts (its pressure step
    relaxes beta and every solve takes one iteration), so it is taken at
    its first RMR < 1e-6."""
    img = connected_image((16, 14, 12), 21)
    prm = dict(dx=1.0, mu=1.0, rho=1.0, p0=1.0, gradp=1.0 / 16, alpha=0.8, beta=1.0, vsolvertol=1e-13,
               psolvertol=1e-13, pcsolvertol=1e-12, maxitsolver=5000)
    out = {}
    for m in ("SIMPLEC", "SIMPLER", "SIMPLERR", "IDEAL"):
        o = O.Oracle(img, **prm)
        if m == "SIMPLERR":
            o.set_horizon(1e-6)
        rows, s = o.solve(m, 1e-10, 3000, N1=4, N2=4)
        if m == "SIMPLERR":
            assert 0 < o.horizon_it() < 100
        else:
            assert s["iterations"] < 3000, m  # converged
        out[m] = (s["K"], {k: o.get(k) for k in "puvw"})
    K0, f0 = out["SIMPLEC"]
    for m, (K, f) in out.items():
        assert abs(K / K0 - 1) <= 1e-7, (m, K, K0)
        if m in ("SIMPLER", "IDEAL"):
            for k in "puvw":
                assert np.linalg.norm(f[k] - f0[k]) <= 5e-6 * np.linalg.norm(f0[k]), (m, k)
