"""ctypes binding of the CPU oracle (oracle/ns3d_oracle.cpp).

Test infrastructure: used only by tests, the smoke check and the CPU-baseline
leg of bench.py, as the checker.  It mirrors the product's C ABI method by
method so that the two can be driven in lockstep on the same inputs.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "build", "libns3d_oracle.so")

SYS_U, SYS_V, SYS_W, SYS_P = 0, 1, 2, 3
SET_VEL, SET_TEMP, SET_HAT = 0, 1, 2
PRESSURE_FULL, PC_SIMPLE, PC_SIMPLEC = 0, 1, 2
CORR_SIMPLE, CORR_SIMPLEC, CORR_SIMPLER = 0, 1, 2
METHODS = {"SIMPLEC": 0, "SIMPLER": 1, "SIMPLERR": 2, "IDEAL": 3}
VEC = {"u": 0, "v": 1, "w": 2, "ut": 3, "vt": 4, "wt": 5, "uh": 6, "vh": 7, "wh": 8,
       "p": 9, "pc": 10, "b": 11}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
        P = C.POINTER
        L.orc_create.restype = vp
        L.orc_create.argtypes = [P(C.c_uint8), i32, i32, i32]
        L.orc_destroy.argtypes = [vp]
        L.orc_counts.argtypes = [vp, P(C.c_int64)]
        L.orc_get_index.argtypes = [vp, P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_int32)]
        L.orc_set_params.argtypes = [vp, P(dbl), P(i32)]
        L.orc_get_p0p1.argtypes = [vp, P(dbl)]
        L.orc_set_alpha_beta.argtypes = [vp, dbl, dbl]
        L.orc_vec_len.restype = i64
        L.orc_vec_len.argtypes = [vp, i32]
        L.orc_set_vec.argtypes = [vp, i32, P(dbl)]
        L.orc_get_vec.argtypes = [vp, i32, P(dbl)]
        L.orc_init_pressure_linear.argtypes = [vp]
        L.orc_assemble_momentum.argtypes = [vp]
        L.orc_matrix_rows.restype = i64
        L.orc_matrix_rows.argtypes = [vp, i32]
        L.orc_get_matrix.argtypes = [vp, i32, P(C.c_int32), P(C.c_int32), P(dbl)]
        L.orc_compute_velocity_field.argtypes = [vp, i32, dbl, P(dbl)]
        L.orc_compute_hat.argtypes = [vp, i32, i32]
        L.orc_assemble_pressure.argtypes = [vp, i32, i32, i32]
        L.orc_compute_pressure_correction.argtypes = [vp, P(dbl)]
        L.orc_compute_pressure_field.argtypes = [vp, P(dbl)]
        L.orc_correct_velocity.argtypes = [vp, i32, dbl, i32, dbl, dbl]
        L.orc_correct_pressure.argtypes = [vp, P(dbl)]
        L.orc_compute_div.argtypes = [vp, i32, P(dbl)]
        L.orc_compute_permeability.argtypes = [vp, P(dbl)]
        L.orc_bicgstab_ell.argtypes = [i32, P(C.c_int32), P(C.c_int32), P(dbl), P(dbl), P(dbl),
                                       dbl, i32, P(dbl)]
        L.orc_spmv_ell.argtypes = [i32, P(C.c_int32), P(C.c_int32), P(dbl), P(dbl), P(dbl)]
        L.orc_solve.restype = i32
        L.orc_solve.argtypes = [vp, i32, dbl, i32, i32, i32, P(dbl), i32, P(dbl)]
        L.orc_prepare_image.argtypes = [P(C.c_uint8), i32, i32, i32, i32, i32, P(C.c_uint8), P(i32)]
        L.orc_num_threads.restype = i32
        L.orc_set_threads.argtypes = [i32]
        L.orc_set_threads.restype = None
        L.orc_relax_replay.argtypes = [P(dbl), i32, i32, dbl, dbl, dbl, P(dbl), P(dbl)]
        L.orc_set_horizon.argtypes = [vp, dbl]
        L.orc_horizon_it.restype = i32
        L.orc_horizon_it.argtypes = [vp]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def prepare_image(img: np.ndarray, direction: int = 1, border: int = 0) -> np.ndarray:
    """Rotation (NSBase.cxx:2451) and AddBorder (NSBase.cxx:2425) as the oracle does them."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    dims = np.zeros(3, dtype=np.int32)
    lib().orc_prepare_image(_p(img, C.c_uint8), *img.shape, direction, border, None, _p(dims, C.c_int))
    out = np.zeros(tuple(int(d) for d in dims), dtype=np.uint8)
    lib().orc_prepare_image(_p(img, C.c_uint8), *img.shape, direction, border, _p(out, C.c_uint8),
                            _p(dims, C.c_int))
    return out


DEFAULTS = dict(dx=1.0, dy=-1.0, dz=-1.0, mu=1e-3, rho=1e3, p0=1e5, gradp=1e2, alpha=0.8, beta=1.0,
                vsolvertol=1e-5, psolvertol=1e-6, pcsolvertol=1e-3, deltaRMR=1e-4,
                maxitsolver=1000, border=0, save_step=20)


class Oracle:
    """CPU restatement of NavierStokesBase on one (already rotated/padded) image."""

    def __init__(self, img: np.ndarray, **params):
        img = np.ascontiguousarray(img, dtype=np.uint8)
        self.shape = img.shape
        self.h = lib().orc_create(_p(img, C.c_uint8), *img.shape)
        c = np.zeros(4, dtype=np.int64)
        lib().orc_counts(self.h, _p(c, C.c_int64))
        self.np, self.nu, self.nv, self.nw = (int(x) for x in c)
        self.set_params(**params)

    def set_params(self, **kw):
        prm = dict(DEFAULTS)
        prm.update(kw)
        if prm["dy"] <= 0:
            prm["dy"] = prm["dx"]
        if prm["dz"] <= 0:
            prm["dz"] = prm["dx"]
        self.params = prm
        d = np.array([prm[k] for k in ("dx", "dy", "dz", "mu", "rho", "p0", "gradp", "alpha", "beta",
                                        "vsolvertol", "psolvertol", "pcsolvertol", "deltaRMR")],
                     dtype=np.float64)
        i = np.array([prm["maxitsolver"], prm["border"], prm["save_step"]], dtype=np.int32)
        lib().orc_set_params(self.h, _p(d, C.c_double), _p(i, C.c_int))
        pp = np.zeros(2)
        lib().orc_get_p0p1(self.h, _p(pp, C.c_double))
        self.p0, self.p1 = float(pp[0]), float(pp[1])

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def index(self):
        nx, ny, nz = self.shape
        p = np.empty((nx, ny, nz), np.int32)
        u = np.empty((nx + 1, ny, nz), np.int32)
        v = np.empty((nx, ny, nz), np.int32)
        w = np.empty((nx, ny, nz), np.int32)
        lib().orc_get_index(self.h, _p(p, C.c_int32), _p(u, C.c_int32), _p(v, C.c_int32),
                            _p(w, C.c_int32))
        return p, u, v, w

    def get(self, name):
        n = lib().orc_vec_len(self.h, VEC[name])
        x = np.empty(n, np.float64)
        lib().orc_get_vec(self.h, VEC[name], _p(x, C.c_double))
        return x

    def set(self, name, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = lib().orc_vec_len(self.h, VEC[name])
        assert x.size == n, (name, x.size, n)
        lib().orc_set_vec(self.h, VEC[name], _p(x, C.c_double))

    def set_alpha_beta(self, alpha, beta):
        lib().orc_set_alpha_beta(self.h, alpha, beta)

    def init_pressure_linear(self):
        lib().orc_init_pressure_linear(self.h)

    def assemble_momentum(self):
        lib().orc_assemble_momentum(self.h)

    def matrix(self, sys):
        n = lib().orc_matrix_rows(self.h, sys)
        cnt = np.empty(n, np.int32)
        col = np.empty((n, 7), np.int32)
        val = np.empty((n, 7), np.float64)
        lib().orc_get_matrix(self.h, sys, _p(cnt, C.c_int32), _p(col, C.c_int32), _p(val, C.c_double))
        return cnt, col, val

    def compute_velocity_field(self, limit, dvmax):
        info = np.zeros(9)
        lib().orc_compute_velocity_field(self.h, int(limit), float(dvmax), _p(info, C.c_double))
        return info.reshape(3, 3)

    def compute_hat(self, uc_set, uold_set):
        lib().orc_compute_hat(self.h, uc_set, uold_set)

    def assemble_pressure(self, kind, relax, vel_set):
        lib().orc_assemble_pressure(self.h, kind, int(relax), vel_set)

    def compute_pressure_correction(self):
        info = np.zeros(3)
        lib().orc_compute_pressure_correction(self.h, _p(info, C.c_double))
        return info

    def compute_pressure_field(self):
        info = np.zeros(3)
        lib().orc_compute_pressure_field(self.h, _p(info, C.c_double))
        return info

    def correct_velocity(self, kind, dvmax, target, p_0=0.0, p_1=0.0):
        lib().orc_correct_velocity(self.h, kind, float(dvmax), target, float(p_0), float(p_1))

    def correct_pressure(self):
        out = np.zeros(3)
        lib().orc_correct_pressure(self.h, _p(out, C.c_double))
        return out

    def compute_div(self, vel_set=SET_VEL):
        out = np.zeros(4)
        lib().orc_compute_div(self.h, vel_set, _p(out, C.c_double))
        return out

    def compute_permeability(self):
        out = np.zeros(5)
        lib().orc_compute_permeability(self.h, _p(out, C.c_double))
        return out

    def set_horizon(self, rmr):
        """Stop after the first outer iteration (it > 1) with RMR < rmr, as if
        capped there; horizon_it() gives that iteration (-1: not met)."""
        lib().orc_set_horizon(self.h, float(rmr))

    def horizon_it(self):
        return int(lib().orc_horizon_it(self.h))

    def solve(self, method="SIMPLEC", simpletol=1e-4, maxit=500, N1=2, N2=2, max_rows=100000):
        rows = np.zeros((max_rows, 14))
        summ = np.zeros(15)
        nr = lib().orc_solve(self.h, METHODS[method], simpletol, maxit, N1, N2, _p(rows, C.c_double),
                             max_rows, _p(summ, C.c_double))
        return rows[:nr].copy(), dict(iterations=int(summ[0]), K=float(summ[1]), seconds=float(summ[2]),
                                      phases=summ[3:10].copy(), solver_iterations=summ[10:14].astype(int).tolist(),
                                      work=int(summ[14]))


def bicgstab_ell(cnt, col, val, b, x0, tol, maxit):
    n = b.size
    cnt = np.ascontiguousarray(cnt, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, dtype=np.float64, copy=True)
    info = np.zeros(3)
    lib().orc_bicgstab_ell(n, _p(cnt, C.c_int32), _p(col, C.c_int32), _p(val, C.c_double),
                           _p(b, C.c_double), _p(x, C.c_double), tol, maxit, _p(info, C.c_double))
    return x, info


def spmv_ell(cnt, col, val, x):
    n = x.size
    cnt = np.ascontiguousarray(cnt, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(n)
    lib().orc_spmv_ell(n, _p(cnt, C.c_int32), _p(col, C.c_int32), _p(val, C.c_double), _p(x, C.c_double),
                       _p(y, C.c_double))
    return y


def relax_replay(rmr, beta_cst=True, alpha0=0.8, beta0=1.0, deltaRMR=1e-4):
    """The oracle's RelaxControl over a scripted RMR sequence: (alpha, beta)
    after each outer iteration."""
    rmr = np.ascontiguousarray(rmr, np.float64)
    a = np.zeros(rmr.size)
    b = np.zeros(rmr.size)
    lib().orc_relax_replay(_p(rmr, C.c_double), rmr.size, int(beta_cst), alpha0, beta0, deltaRMR,
                           _p(a, C.c_double), _p(b, C.c_double))
    return a, b
