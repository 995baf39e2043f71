"""ctypes binding of libns3d_gpu.so (include/ns3d.h).

``NavierStokesGPU`` mirrors the NavierStokesBase methods of the reference
(src/NSBase.h:63-594) over one device-resident image.  Importing this module
never falls back to a CPU implementation: if the HIP library is missing or no
GPU is visible, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# NS3D_GPU_LIBRARY may name another build of the library (e.g. one compiled
# with different tuning macros); the default is the in-tree build.
LIB_PATH = os.environ.get("NS3D_GPU_LIBRARY") or os.path.join(HERE, "libns3d_gpu.so")

SYS_U, SYS_V, SYS_W, SYS_P = 0, 1, 2, 3
SET_VEL, SET_TEMP, SET_HAT = 0, 1, 2
PRESSURE_FULL, PC_SIMPLE, PC_SIMPLEC = 0, 1, 2
CORR_SIMPLE, CORR_SIMPLEC, CORR_SIMPLER = 0, 1, 2
VEC = {"u": 0, "v": 1, "w": 2, "ut": 3, "vt": 4, "wt": 5, "uh": 6, "vh": 7, "wh": 8,
       "p": 9, "pc": 10, "b": 11}

# Every entry point declared in include/ns3d.h.
EXPORTS = [
    "ns3d_last_error", "ns3d_device_count", "ns3d_prepare_image", "ns3d_create", "ns3d_destroy",
    "ns3d_counts", "ns3d_get_index", "ns3d_set_params", "ns3d_zero_fields", "ns3d_init_pressure_linear",
    "ns3d_set_vector", "ns3d_get_vector", "ns3d_copy_velocity", "ns3d_assemble_momentum",
    "ns3d_compute_velocity_field", "ns3d_compute_hat", "ns3d_assemble_pressure",
    "ns3d_compute_pressure_correction", "ns3d_compute_pressure_field", "ns3d_correct_velocity",
    "ns3d_correct_pressure", "ns3d_compute_div", "ns3d_compute_permeability", "ns3d_get_matrix",
    "ns3d_bicgstab_stencil", "ns3d_timing", "ns3d_timing_read", "ns3d_matrix_nnz", "ns3d_set_precision",
    "ns3d_set_mixed_min_rows", "ns3d_set_multigrid", "ns3d_multigrid_levels", "ns3d_multigrid_level", "ns3d_multigrid_timing", "ns3d_copy_bandwidth", "ns3d_sync", "ns3d_bicgstab_stencil_noalloc", "ns3d_noalloc_max_rows",
    "ns3d_engine_create", "ns3d_engine_destroy", "ns3d_attach_engine", "ns3d_engine_stats",
    "ns3d_engine_set_graphs", "ns3d_engine_set_schedule", "ns3d_engine_get_schedule",
    "ns3d_footprint", "ns3d_ctx_bytes",
]


class NS3DError(RuntimeError):
    pass


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("info", C.c_int32), ("restarts", C.c_int32), ("error", C.c_double),
                ("work", C.c_int64)]

    def as_tuple(self):
        return (int(self.iterations), int(self.info), float(self.error))


class Params(C.Structure):
    _fields_ = [("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("mu", C.c_double),
                ("rho", C.c_double), ("p0", C.c_double), ("p1", C.c_double), ("border", C.c_int32),
                ("reserved", C.c_int32)]


_lib = None


def lib():
    """Loads libns3d_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NS3DError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
        P = C.POINTER
        L.ns3d_last_error.restype = C.c_char_p
        L.ns3d_device_count.restype = i32
        for name in EXPORTS:
            f = getattr(L, name)
            if name not in ("ns3d_last_error", "ns3d_device_count", "ns3d_destroy", "ns3d_engine_destroy"):
                f.restype = i32
        L.ns3d_destroy.restype = None
        L.ns3d_engine_destroy.restype = None
        L.ns3d_engine_create.argtypes = [i32, i32, P(vp)]
        L.ns3d_engine_destroy.argtypes = [vp]
        L.ns3d_attach_engine.argtypes = [vp, vp]
        L.ns3d_engine_stats.argtypes = [vp, P(dbl)]
        L.ns3d_engine_set_graphs.argtypes = [vp, i32]
        L.ns3d_engine_set_schedule.argtypes = [vp, i32, i32, i32]
        L.ns3d_engine_get_schedule.argtypes = [vp, P(i32)]
        L.ns3d_footprint.argtypes = [i32, i32, i32, i64, i32, i32, P(dbl)]
        L.ns3d_ctx_bytes.argtypes = [vp, P(dbl)]
        L.ns3d_prepare_image.argtypes = [P(C.c_uint8), i32, i32, i32, i32, i32, P(C.c_uint8), P(i32)]
        L.ns3d_create.argtypes = [P(C.c_uint8), i32, i32, i32, i32, P(vp)]
        L.ns3d_destroy.argtypes = [vp]
        L.ns3d_counts.argtypes = [vp, P(i64)]
        L.ns3d_get_index.argtypes = [vp, P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_int32)]
        L.ns3d_set_params.argtypes = [vp, P(Params)]
        L.ns3d_zero_fields.argtypes = [vp]
        L.ns3d_init_pressure_linear.argtypes = [vp]
        L.ns3d_set_vector.argtypes = [vp, i32, P(dbl), i64]
        L.ns3d_get_vector.argtypes = [vp, i32, P(dbl), i64]
        L.ns3d_copy_velocity.argtypes = [vp, i32, i32]
        L.ns3d_assemble_momentum.argtypes = [vp, dbl]
        L.ns3d_compute_velocity_field.argtypes = [vp, dbl, i32, dbl, dbl, i64, P(SolveInfo)]
        L.ns3d_compute_hat.argtypes = [vp, dbl, i32, i32]
        L.ns3d_assemble_pressure.argtypes = [vp, i32, i32, dbl, i32]
        L.ns3d_compute_pressure_correction.argtypes = [vp, dbl, i64, P(SolveInfo)]
        L.ns3d_compute_pressure_field.argtypes = [vp, dbl, i64, P(SolveInfo)]
        L.ns3d_correct_velocity.argtypes = [vp, i32, dbl, i32, dbl, dbl]
        L.ns3d_correct_pressure.argtypes = [vp, dbl, P(dbl)]
        L.ns3d_compute_div.argtypes = [vp, i32, P(dbl)]
        L.ns3d_compute_permeability.argtypes = [vp, P(dbl)]
        L.ns3d_get_matrix.argtypes = [vp, i32, P(dbl), P(dbl), P(C.c_int32), P(dbl)]
        L.ns3d_bicgstab_stencil.argtypes = [i32, P(dbl), P(dbl), P(C.c_int32), P(dbl), P(dbl), dbl, i64,
                                            P(SolveInfo)]
        L.ns3d_bicgstab_stencil_noalloc.argtypes = [i32, vp, vp, vp, vp, vp, dbl, i64, P(SolveInfo), vp]
        L.ns3d_noalloc_max_rows.argtypes = [P(i64)]
        L.ns3d_timing.argtypes = [vp, i32]
        L.ns3d_timing_read.argtypes = [vp, P(dbl)]
        L.ns3d_matrix_nnz.argtypes = [vp, P(i64)]
        L.ns3d_set_precision.argtypes = [vp, i32, dbl]
        L.ns3d_set_mixed_min_rows.argtypes = [vp, i64]
        L.ns3d_set_multigrid.argtypes = [vp, i64]
        L.ns3d_multigrid_levels.argtypes = [vp, P(i32), P(i64)]
        L.ns3d_multigrid_level.argtypes = [vp, i32, P(C.c_int32), P(C.c_float)]
        L.ns3d_copy_bandwidth.argtypes = [i32, i64, i32, P(dbl)]
        L.ns3d_multigrid_timing.argtypes = [vp, P(dbl)]
        L.ns3d_sync.argtypes = [vp]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _check(rc):
    if rc != 0:
        msg = lib().ns3d_last_error().decode(errors="replace")
        raise NS3DError(f"ns3d error {rc}: {msg}")


def device_count() -> int:
    return int(lib().ns3d_device_count())


def prepare_image(img: np.ndarray, direction: int = 1, border: int = 0) -> np.ndarray:
    """Rotation so the flow is along x (direction 2: y, 3: z) and a fluid
    border of ``border`` voxels on both x sides (NSBase.cxx:2425-2492)."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    dims = np.zeros(3, np.int32)
    _check(lib().ns3d_prepare_image(_p(img, C.c_uint8), *img.shape, direction, border, None, _p(dims, C.c_int)))
    out = np.empty(tuple(int(d) for d in dims), np.uint8)
    _check(lib().ns3d_prepare_image(_p(img, C.c_uint8), *img.shape, direction, border, _p(out, C.c_uint8),
                                    _p(dims, C.c_int)))
    return out


def p0_p1(p0: float, gradp: float, dx: float, nx: int):
    """Outlet pressure from the imposed gradient with the negative-p1 shift
    (NSBase.cxx:155-161)."""
    p1 = -gradp * (dx * nx) + p0
    if p1 < 0:
        p0 -= p1
        p1 = 0.0
    return p0, p1


class NavierStokesGPU:
    """One image on the GPU; methods mirror NavierStokesBase."""

    def __init__(self, img: np.ndarray, dx=1.0, dy=-1.0, dz=-1.0, mu=1e-3, rho=1e3, p0=1e5, gradp=1e2,
                 border=0, device=-1, **_ignored):
        img = np.ascontiguousarray(img, dtype=np.uint8)
        self.shape = img.shape
        h = C.c_void_p()
        _check(lib().ns3d_create(_p(img, C.c_uint8), *img.shape, device, C.byref(h)))
        self.h = h
        c = np.zeros(4, np.int64)
        _check(lib().ns3d_counts(self.h, _p(c, C.c_int64)))
        self.np, self.nu, self.nv, self.nw = (int(x) for x in c)
        dy = dx if dy <= 0 else dy
        dz = dx if dz <= 0 else dz
        self.p0, self.p1 = p0_p1(p0, gradp, dx, img.shape[0])
        self.prm = Params(dx, dy, dz, mu, rho, self.p0, self.p1, border, 0)
        _check(lib().ns3d_set_params(self.h, C.byref(self.prm)))

    def close(self):
        if getattr(self, "h", None):
            lib().ns3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def index(self):
        nx, ny, nz = self.shape
        p = np.empty((nx, ny, nz), np.int32)
        u = np.empty((nx + 1, ny, nz), np.int32)
        v = np.empty((nx, ny, nz), np.int32)
        w = np.empty((nx, ny, nz), np.int32)
        _check(lib().ns3d_get_index(self.h, _p(p, C.c_int32), _p(u, C.c_int32), _p(v, C.c_int32),
                                    _p(w, C.c_int32)))
        return p, u, v, w

    def _len(self, name):
        return {"u": self.nu, "v": self.nv, "w": self.nw}.get(name[0], self.np)

    def get(self, name):
        n = self._len(name)
        x = np.empty(n, np.float64)
        _check(lib().ns3d_get_vector(self.h, VEC[name], _p(x, C.c_double), n))
        return x

    def set(self, name, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        _check(lib().ns3d_set_vector(self.h, VEC[name], _p(x, C.c_double), x.size))

    def zero_fields(self):
        _check(lib().ns3d_zero_fields(self.h))

    def init_pressure_linear(self):
        _check(lib().ns3d_init_pressure_linear(self.h))

    def copy_velocity(self, dst, src):
        _check(lib().ns3d_copy_velocity(self.h, dst, src))

    def assemble_momentum(self, alpha):
        _check(lib().ns3d_assemble_momentum(self.h, alpha))

    def compute_velocity_field(self, alpha, limit, dvmax, tol, maxit):
        info = (SolveInfo * 3)()
        _check(lib().ns3d_compute_velocity_field(self.h, alpha, int(limit), float(dvmax), tol, maxit, info))
        return [info[q].as_tuple() for q in range(3)]

    def compute_hat(self, alpha, uc_set, uold_set):
        _check(lib().ns3d_compute_hat(self.h, alpha, uc_set, uold_set))

    def assemble_pressure(self, kind, relax, beta, vel_set):
        _check(lib().ns3d_assemble_pressure(self.h, kind, int(relax), beta, vel_set))

    def compute_pressure_correction(self, tol, maxit):
        info = SolveInfo()
        _check(lib().ns3d_compute_pressure_correction(self.h, tol, maxit, C.byref(info)))
        return info.as_tuple()

    def compute_pressure_field(self, tol, maxit):
        info = SolveInfo()
        _check(lib().ns3d_compute_pressure_field(self.h, tol, maxit, C.byref(info)))
        return info.as_tuple()

    def correct_velocity(self, kind, dvmax, target, p_0=0.0, p_1=0.0):
        _check(lib().ns3d_correct_velocity(self.h, kind, float(dvmax), target, float(p_0), float(p_1)))

    def correct_pressure(self, beta):
        out = np.zeros(3)
        _check(lib().ns3d_correct_pressure(self.h, beta, _p(out, C.c_double)))
        return out

    def compute_div(self, vel_set=SET_VEL):
        out = np.zeros(4)
        _check(lib().ns3d_compute_div(self.h, vel_set, _p(out, C.c_double)))
        return out

    def compute_permeability(self):
        out = np.zeros(5)
        _check(lib().ns3d_compute_permeability(self.h, _p(out, C.c_double)))
        return out

    def matrix(self, sys):
        n = [self.nu, self.nv, self.nw, self.np][sys]
        diag = np.empty(n)
        off = np.empty((6, n))
        nbr = np.empty((6, n), np.int32)
        rowsum = np.empty(n)
        _check(lib().ns3d_get_matrix(self.h, sys, _p(diag, C.c_double), _p(off, C.c_double),
                                     _p(nbr, C.c_int32), _p(rowsum, C.c_double)))
        return diag, off, nbr, (rowsum if sys != SYS_P else None)

    def timing(self, enable=True):
        _check(lib().ns3d_timing(self.h, int(enable)))

    def timing_read(self):
        """{(prec, kind, kernel): (ms, launches, bytes)} for prec in f64/f32,
        kind in momentum/pressure and kernel in A/B/C of the fused BiCGSTAB
        (sampled launches)."""
        out = np.zeros(36)
        _check(lib().ns3d_timing_read(self.h, _p(out, C.c_double)))
        res = {}
        for q, prec in enumerate(("f64", "f32")):
            for m, kind in enumerate(("momentum", "pressure")):
                for k, name in enumerate("ABC"):
                    o = 18 * q + 9 * m + 3 * k
                    res[(prec, kind, name)] = (float(out[o]), float(out[o + 1]), float(out[o + 2]))
        return res

    def device_bytes(self):
        out = C.c_double()
        _check(lib().ns3d_ctx_bytes(self.h, C.byref(out)))
        return out.value

    def set_precision(self, mixed: bool, inner_tol: float = 1e-4, min_rows: int = 0):
        """min_rows: in the mixed form, systems with fewer rows are solved in FP64."""
        _check(lib().ns3d_set_precision(self.h, int(bool(mixed)), inner_tol))
        _check(lib().ns3d_set_mixed_min_rows(self.h, int(min_rows)))

    def set_multigrid(self, min_rows: int):
        """Mixed form: pressure systems from min_rows rows on are swept by
        multigrid-preconditioned CG (0: never)."""
        _check(lib().ns3d_set_multigrid(self.h, int(min_rows)))

    def multigrid_levels(self):
        """Row counts of the pressure system's multigrid levels ([] until built)."""
        lev = C.c_int32(0)
        rows = (C.c_int64 * 12)()
        _check(lib().ns3d_multigrid_levels(self.h, C.byref(lev), rows))
        return [int(rows[i]) for i in range(lev.value)]

    def multigrid_level(self, level):
        """(cells, coef[4][n]) of hierarchy level `level` >= 1 as on the device."""
        n = self.multigrid_levels()[level]
        cell = np.empty(n, np.int32)
        coef = np.empty((4, n), np.float32)
        _check(lib().ns3d_multigrid_level(self.h, int(level), cell.ctypes.data_as(C.POINTER(C.c_int32)),
                                          coef.ctypes.data_as(C.POINTER(C.c_float))))
        return cell, coef

    def nnz(self):
        out = np.zeros(4, np.int64)
        _check(lib().ns3d_matrix_nnz(self.h, _p(out, C.c_int64)))
        return [int(x) for x in out]


class Engine:
    """Batched solve engine: contexts attached to it (each driven by its own
    thread) have their BiCGSTAB solves run together, one launch per kernel
    phase over all waiting solves."""

    STAT_NAMES = ["chunks", "launches", "solves", "max_active", "rows_x_iters"]

    def __init__(self, device: int = 0, chunk: int = 8, graphs: bool = True, lanes: int = 3):
        h = C.c_void_p()
        _check(lib().ns3d_engine_create(device, chunk, C.byref(h)))
        self.h = h
        _check(lib().ns3d_engine_set_schedule(self.h, chunk, lanes, int(bool(graphs))))

    @staticmethod
    def set_schedule_of(h, chunk: int, lanes: int, graphs: bool):
        _check(lib().ns3d_engine_set_schedule(h, int(chunk), int(lanes), int(bool(graphs))))

    @staticmethod
    def schedule_of(h) -> dict:
        out = np.zeros(3, np.int32)
        _check(lib().ns3d_engine_get_schedule(h, _p(out, C.c_int32)))
        return dict(chunk=int(out[0]), lanes=int(out[1]), graphs=bool(out[2]))

    def attach(self, ctx):
        _check(lib().ns3d_attach_engine(ctx, self.h))

    @staticmethod
    def detach(ctx):
        _check(lib().ns3d_attach_engine(ctx, None))

    def stats(self):
        out = np.zeros(32)
        _check(lib().ns3d_engine_stats(self.h, _p(out, C.c_double)))
        return self.unpack_stats(out)

    @classmethod
    def unpack_stats(cls, out):
        d = dict(zip(cls.STAT_NAMES, out[:5].tolist()))
        for q, prec in enumerate(("f64", "f32")):
            for k, name in enumerate("ABC"):
                o = 5 + 9 * q + 3 * k
                d[f"{prec}_{name}"] = (float(out[o]), float(out[o + 1]), float(out[o + 2]))
        d["chunk"] = (float(out[23]), float(out[24]), float(out[25]))
        d["host_ms"] = dict(queue=float(out[26]), process=float(out[27]), wait=float(out[28]))
        d["sampled_tiles"] = dict(live=float(out[29]), launched=float(out[30]))
        return d

    def close(self):
        if getattr(self, "h", None):
            lib().ns3d_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_bandwidth(nbytes: int, reps: int = 20, device: int = 0) -> float:
    """GB/s (read + written) of a plain device copy of nbytes (the library's
    own copy kernel)."""
    out = C.c_double(0.0)
    _check(lib().ns3d_copy_bandwidth(int(device), int(nbytes), int(reps), C.byref(out)))
    return out.value


def footprint(shape, nfluid, mixed=False, engine=False):
    """Projected device bytes of a context (upper bound); engine: attached to
    a solve engine (two more solver workspaces)."""
    out = C.c_double()
    _check(lib().ns3d_footprint(*[int(x) for x in shape], int(nfluid), int(bool(mixed)), 3 if engine else 1,
                                C.byref(out)))
    return out.value


def bicgstab_stencil(diag, off, nbr, b, x0, tol, maxit):
    """Stand-alone Eigen-semantics Jacobi-BiCGSTAB on a 7-point matrix."""
    diag = np.ascontiguousarray(diag, np.float64)
    off = np.ascontiguousarray(off, np.float64)
    nbr = np.ascontiguousarray(nbr, np.int32)
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, dtype=np.float64, copy=True)
    info = SolveInfo()
    _check(lib().ns3d_bicgstab_stencil(diag.size, _p(diag, C.c_double), _p(off, C.c_double),
                                       _p(nbr, C.c_int32), _p(b, C.c_double), _p(x, C.c_double), tol, maxit,
                                       C.byref(info)))
    return x, info.as_tuple()


def noalloc_max_rows() -> int:
    rows = C.c_int64()
    _check(lib().ns3d_noalloc_max_rows(C.byref(rows)))
    return rows.value


def bicgstab_stencil_noalloc(diag, off, nbr, b, x, tol, maxit, stream=None):
    """The stand-alone solve with no device allocation: diag, off, nbr, b and
    x are contiguous tensors on the current GPU (float64, float64, int32,
    float64, float64); x holds the guess and is overwritten with the
    solution.  stream: a torch stream to queue on (default: the current
    one)."""
    import torch
    n = int(diag.numel())
    for t, name, dt, size in ((diag, "diag", torch.float64, n), (off, "off", torch.float64, 6 * n),
                              (nbr, "nbr", torch.int32, 6 * n), (b, "b", torch.float64, n),
                              (x, "x", torch.float64, n)):
        if t.dtype != dt or not t.is_cuda or not t.is_contiguous() or t.numel() != size:
            raise ValueError(f"{name}: expected a contiguous {dt} tensor of {size} elements on the GPU")
    s = stream if stream is not None else torch.cuda.current_stream()
    info = SolveInfo()
    _check(lib().ns3d_bicgstab_stencil_noalloc(n, diag.data_ptr(), off.data_ptr(), nbr.data_ptr(), b.data_ptr(),
                                               x.data_ptr(), tol, maxit, C.byref(info), s.cuda_stream))
    return info.as_tuple()
