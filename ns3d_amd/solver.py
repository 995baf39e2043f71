"""ctypes binding of libns3d_host.so (include/ns3d_host.h): one complete
permeability solve -- NavierStokesBase + SIMPLEC/SIMPLER/SIMPLERR/IDEAL driver,
host C++ over the GPU C ABI -- on an in-memory image.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import gpu as _gpu

HERE = os.path.dirname(os.path.abspath(__file__))
# NS3D_HOST_LIBRARY may name another build of the library (one next to a
# libns3d_gpu.so built with other tuning macros, which it then loads).
LIB_PATH = os.environ.get("NS3D_HOST_LIBRARY") or os.path.join(HERE, "libns3d_host.so")
METHODS = {"SIMPLEC": 0, "SIMPLER": 1, "SIMPLERR": 2, "IDEAL": 3}

EXPORTS = ["ns3d_default_run_params", "ns3d_solver_create", "ns3d_solver_info", "ns3d_solver_run",
           "ns3d_solver_fields", "ns3d_solver_destroy", "ns3d_solver_context", "ns3d_host_last_error",
           "ns3d_read_ini", "ns3d_solver_warm_start", "ns3d_server_create", "ns3d_server_submit", "ns3d_server_poll", "ns3d_server_wait",
           "ns3d_server_stats", "ns3d_server_destroy", "ns3d_server_last_error", "ns3d_server_engine",
           "ns3d_server_set_policy", "ns3d_relax_replay", "ns3d_read_image"]


class RunParams(C.Structure):
    _fields_ = [("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("mu", C.c_double),
                ("rho", C.c_double), ("p0", C.c_double), ("gradp", C.c_double), ("alpha", C.c_double),
                ("beta", C.c_double), ("vsolvertol", C.c_double), ("psolvertol", C.c_double),
                ("pcsolvertol", C.c_double), ("simpletol", C.c_double), ("deltaRMR", C.c_double),
                ("maxitsolver", C.c_int32), ("maxitSIMPLE", C.c_int32), ("N1", C.c_int32), ("N2", C.c_int32),
                ("save_step", C.c_int32), ("porevalue", C.c_int32), ("direction", C.c_int32),
                ("border", C.c_int32), ("method", C.c_int32), ("verbose", C.c_int32),
                ("write_files", C.c_int32), ("device", C.c_int32), ("outdir", C.c_char_p),
                ("name", C.c_char_p), ("precision", C.c_int32), ("mixed_min_rows", C.c_int32),
                ("inner_tol", C.c_double), ("mg_min_rows", C.c_int32), ("pad_params", C.c_int32)]


class RunResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("K", C.c_double), ("Q", C.c_double),
                ("v_moy", C.c_double), ("RMR", C.c_double), ("normdiv", C.c_double), ("reldiv", C.c_double),
                ("maxdiv", C.c_double), ("seconds", C.c_double), ("np", C.c_int64), ("nu", C.c_int64),
                ("nv", C.c_int64), ("nw", C.c_int64), ("solver_iterations", C.c_int64 * 4),
                ("phase_seconds", C.c_double * 8), ("solver_work", C.c_int64 * 4), ("fallback", C.c_int32),
                ("pad", C.c_int32)]


PHASES = ["momentum_assembly", "velocity_solves", "pseudo_velocities", "pressure_assembly",
          "pressure_solves", "corrections", "diagnostics", "other"]

_lib = None


def lib():
    global _lib
    if _lib is None:
        _gpu.lib()  # fails loudly when the GPU library is missing
        if not os.path.exists(LIB_PATH):
            raise _gpu.NS3DError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.ns3d_default_run_params.argtypes = [P(RunParams)]
        L.ns3d_default_run_params.restype = None
        L.ns3d_solver_create.argtypes = [P(RunParams), P(C.c_uint8), C.c_int, C.c_int, C.c_int, P(C.c_void_p)]
        L.ns3d_solver_create.restype = C.c_int
        L.ns3d_solver_info.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int64)]
        L.ns3d_solver_info.restype = C.c_int
        L.ns3d_solver_run.argtypes = [C.c_void_p, P(C.c_double), C.c_int, P(C.c_int), P(RunResult)]
        L.ns3d_solver_run.restype = C.c_int
        L.ns3d_solver_fields.argtypes = [C.c_void_p] + [P(C.c_double)] * 4
        L.ns3d_solver_fields.restype = C.c_int
        L.ns3d_solver_destroy.argtypes = [C.c_void_p]
        L.ns3d_solver_destroy.restype = None
        L.ns3d_solver_context.argtypes = [C.c_void_p]
        L.ns3d_solver_context.restype = C.c_void_p
        L.ns3d_host_last_error.restype = C.c_char_p
        L.ns3d_read_ini.argtypes = [C.c_char_p, P(RunParams), C.c_char_p, C.c_int]
        L.ns3d_read_ini.restype = C.c_int
        vp = C.c_void_p
        L.ns3d_solver_warm_start.argtypes = [vp] + [P(C.c_double)] * 4
        L.ns3d_solver_warm_start.restype = C.c_int
        L.ns3d_server_create.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, P(vp)]
        L.ns3d_server_create.restype = C.c_int
        L.ns3d_server_submit.argtypes = [vp, P(RunParams), P(C.c_uint8), C.c_int, C.c_int, C.c_int, P(C.c_int64)]
        L.ns3d_server_submit.restype = C.c_int
        for f in (L.ns3d_server_poll, L.ns3d_server_wait):
            f.argtypes = [vp, C.c_int64, P(RunResult), P(C.c_double)]
            f.restype = C.c_int
        L.ns3d_server_stats.argtypes = [vp, P(C.c_double)]
        L.ns3d_server_stats.restype = C.c_int
        L.ns3d_server_destroy.argtypes = [vp]
        L.ns3d_server_destroy.restype = None
        L.ns3d_server_last_error.restype = C.c_char_p
        L.ns3d_server_engine.argtypes = [vp]
        L.ns3d_server_engine.restype = vp
        L.ns3d_server_set_policy.argtypes = [vp, C.c_int, C.c_double]
        L.ns3d_relax_replay.argtypes = [P(C.c_double), C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        P(C.c_double), P(C.c_double), P(C.c_int)]
        L.ns3d_relax_replay.restype = C.c_int
        L.ns3d_read_image.argtypes = [C.c_char_p, C.c_int, P(C.c_uint8), P(C.c_int)]
        L.ns3d_read_image.restype = C.c_int
        _lib = L
    return _lib


def relax_replay(rmr, beta_cst=True, alpha0=0.8, beta0=1.0, deltaRMR=1e-4):
    """RelaxControl over a scripted sequence of relative mass residuals
    (ns3d_relax_replay): arrays alpha, beta and events after each outer
    iteration (events: 1 jump, 2 flat window, 4 rising window)."""
    r = np.ascontiguousarray(rmr, np.float64)
    a, b = np.zeros(r.size), np.zeros(r.size)
    ev = np.zeros(r.size, np.int32)
    P = C.POINTER
    if lib().ns3d_relax_replay(r.ctypes.data_as(P(C.c_double)), r.size, int(beta_cst), alpha0, beta0, deltaRMR,
                               a.ctypes.data_as(P(C.c_double)), b.ctypes.data_as(P(C.c_double)),
                               ev.ctypes.data_as(P(C.c_int))) != 0:
        raise _gpu.NS3DError(lib().ns3d_host_last_error().decode())
    return a, b, ev


def read_image(path, porevalue=0):
    """A binary image file read and binarised by the host library
    (ns3d_read_image): uint8 (nx, ny, nz), 0 = fluid, 1 = solid."""
    dims = (C.c_int * 3)()
    if lib().ns3d_read_image(str(path).encode(), int(porevalue), None, dims) != 0:
        raise _gpu.NS3DError(lib().ns3d_host_last_error().decode())
    out = np.empty(tuple(dims), np.uint8)
    if lib().ns3d_read_image(str(path).encode(), int(porevalue), out.ctypes.data_as(C.POINTER(C.c_uint8)), dims) != 0:
        raise _gpu.NS3DError(lib().ns3d_host_last_error().decode())
    return out


def default_params(**kw) -> RunParams:
    p = RunParams()
    lib().ns3d_default_run_params(C.byref(p))
    p.verbose = 0
    p.write_files = 0
    for k, v in kw.items():
        if k == "method":
            v = METHODS[v.upper()] if isinstance(v, str) else int(v)
        if k == "precision":
            v = {"fp64": 0, "mixed": 1}[v] if isinstance(v, str) else int(v)
        if k in ("outdir", "name") and isinstance(v, str):
            v = v.encode()
        setattr(p, k, v)
    return p


class Solver:
    """One image, one driver; ``run()`` returns the RES rows and a summary."""

    def __init__(self, img: np.ndarray, **params):
        self.params = default_params(**params)
        img = np.ascontiguousarray(img, dtype=np.uint8)
        h = C.c_void_p()
        rc = lib().ns3d_solver_create(C.byref(self.params), img.ctypes.data_as(C.POINTER(C.c_uint8)),
                                      *img.shape, C.byref(h))
        if rc != 0:
            raise _gpu.NS3DError(lib().ns3d_host_last_error().decode())
        self.h = h
        dims = (C.c_int * 3)()
        cnt = (C.c_int64 * 4)()
        lib().ns3d_solver_info(self.h, dims, cnt)
        self.dims = tuple(dims)
        self.np, self.nu, self.nv, self.nw = (int(x) for x in cnt)

    def run(self, max_rows=100000):
        rows = np.zeros((max_rows, 14))
        n = C.c_int(0)
        res = RunResult()
        rc = lib().ns3d_solver_run(self.h, rows.ctypes.data_as(C.POINTER(C.c_double)), max_rows, C.byref(n),
                                   C.byref(res))
        if rc != 0:
            raise _gpu.NS3DError(lib().ns3d_host_last_error().decode())
        summary = {f: getattr(res, f) for f, _ in RunResult._fields_
                   if f not in ("solver_iterations", "phase_seconds", "solver_work", "pad")}
        summary["solver_iterations"] = list(res.solver_iterations)
        summary["solver_work"] = list(res.solver_work)
        summary["phases"] = dict(zip(PHASES, list(res.phase_seconds)))
        return rows[:min(n.value, max_rows)].copy(), summary

    def context(self):
        """ns3d_ctx* of the GPU library (for timing and structure queries)."""
        return C.c_void_p(lib().ns3d_solver_context(self.h))

    def timing(self, enable=True):
        _gpu._check(_gpu.lib().ns3d_timing(self.context(), int(enable)))

    def warm_start(self, p=None, u=None, v=None, w=None):
        """Initial fields for the next run (None: the default start)."""
        arrs = []
        ptrs = []
        for x, n in ((p, self.np), (u, self.nu), (v, self.nv), (w, self.nw)):
            if x is None:
                ptrs.append(None)
            else:
                a = np.ascontiguousarray(x, dtype=np.float64)
                if a.size != n:
                    raise ValueError("warm_start: wrong field length")
                arrs.append(a)
                ptrs.append(a.ctypes.data_as(C.POINTER(C.c_double)))
        self._warm = arrs  # kept alive until the run
        if lib().ns3d_solver_warm_start(self.h, *ptrs) != 0:
            raise _gpu.NS3DError("ns3d_solver_warm_start failed")

    def attach(self, engine):
        """Hands this solver's BiCGSTAB solves to a batched engine (None
        detaches); run() must then be called from its own thread for solves
        of several solvers to overlap."""
        if engine is None:
            _gpu.Engine.detach(self.context())
        else:
            engine.attach(self.context())

    def timing_read(self):
        out = np.zeros(36)
        _gpu._check(_gpu.lib().ns3d_timing_read(self.context(), out.ctypes.data_as(C.POINTER(C.c_double))))
        res = {}
        for q, prec in enumerate(("f64", "f32")):
            for m, kind in enumerate(("momentum", "pressure")):
                for k, name in enumerate("ABC"):
                    o = 18 * q + 9 * m + 3 * k
                    res[(prec, kind, name)] = (float(out[o]), float(out[o + 1]), float(out[o + 2]))
        return res

    def multigrid_timing(self):
        """{kernel: (ms, launches, bytes)} of the multigrid sweeps' sampled
        finest-level kernels (restrict, prolong, dir, update)."""
        out = np.zeros(12)
        _gpu._check(_gpu.lib().ns3d_multigrid_timing(self.context(), out.ctypes.data_as(C.POINTER(C.c_double))))
        return {name: (float(out[3 * k]), float(out[3 * k + 1]), float(out[3 * k + 2]))
                for k, name in enumerate(("restrict", "prolong", "dir", "update"))}

    def nnz(self):
        out = np.zeros(4, np.int64)
        _gpu._check(_gpu.lib().ns3d_matrix_nnz(self.context(), out.ctypes.data_as(C.POINTER(C.c_int64))))
        return [int(x) for x in out]

    def fields(self):
        p, u, v, w = (np.empty(n) for n in (self.np, self.nu, self.nv, self.nw))
        ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        rc = lib().ns3d_solver_fields(self.h, ptr(p), ptr(u), ptr(v), ptr(w))
        if rc != 0:
            raise _gpu.NS3DError(lib().ns3d_host_last_error().decode())
        return dict(p=p, u=u, v=v, w=w)

    def close(self):
        if getattr(self, "h", None):
            lib().ns3d_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _summary(res):
    summary = {f: getattr(res, f) for f, _ in RunResult._fields_
               if f not in ("solver_iterations", "phase_seconds", "solver_work", "pad")}
    summary["solver_iterations"] = list(res.solver_iterations)
    summary["solver_work"] = list(res.solver_work)
    summary["phases"] = dict(zip(PHASES, list(res.phase_seconds)))
    return summary


FALLBACK_NONE, FALLBACK_FAILED, FALLBACK_UNCONVERGED = 0, 1, 2


class Server:
    """Permeability requests served as they arrive (native server of
    libns3d_host.so): worker slots admit requests by projected device
    footprint within budget_bytes, and all running requests share one solve
    engine.  fallback: FALLBACK_FAILED (or True) runs a mixed-precision
    request again in FP64 when its run fails or gives non-finite values;
    FALLBACK_UNCONVERGED also continues one that ends unconverged at its
    outer-iteration cap in FP64 from its fields."""

    STAT_NAMES = ["submitted", "queued", "running", "done", "failed", "fallbacks", "bytes_in_use",
                  "peak_bytes_in_use", "peak_running", "budget", "seconds"]

    def __init__(self, device=0, budget_bytes=200e9, max_slots=16, chunk=8, fallback=FALLBACK_FAILED):
        h = C.c_void_p()
        if lib().ns3d_server_create(device, float(budget_bytes), int(max_slots), int(chunk), int(fallback),
                                    C.byref(h)) != 0:
            raise _gpu.NS3DError(lib().ns3d_server_last_error().decode())
        self.h = h
        self._keep = {}

    def submit(self, img, **params):
        """Queues one request; returns its ticket."""
        p = default_params(**params)
        img = np.ascontiguousarray(img, dtype=np.uint8)
        t = C.c_int64()
        rc = lib().ns3d_server_submit(self.h, C.byref(p), img.ctypes.data_as(C.POINTER(C.c_uint8)), *img.shape,
                                      C.byref(t))
        if rc != 0:
            raise _gpu.NS3DError(lib().ns3d_server_last_error().decode())
        return int(t.value)

    def _fetch(self, fn, ticket):
        res = RunResult()
        times = (C.c_double * 4)()
        rc = fn(self.h, ticket, C.byref(res), times)
        if rc == 1:
            return None
        if rc < 0:
            raise _gpu.NS3DError(f"request {ticket}: " + lib().ns3d_server_last_error().decode())
        summ = _summary(res)
        summ.update(t_submit=times[0], t_start=times[1], t_end=times[2], footprint=times[3])
        return summ

    def poll(self, ticket):
        """The request's summary (with t_submit, t_start, t_end, footprint)
        once done, else None."""
        return self._fetch(lib().ns3d_server_poll, ticket)

    def wait(self, ticket):
        return self._fetch(lib().ns3d_server_wait, ticket)

    def stats(self):
        out = (C.c_double * 16)()
        lib().ns3d_server_stats(self.h, out)
        return dict(zip(self.STAT_NAMES, list(out)[:11]))

    def engine_stats(self):
        """Statistics of the server's solve engine (see gpu.Engine.stats)."""
        out = np.zeros(32)
        _gpu._check(_gpu.lib().ns3d_engine_stats(C.c_void_p(lib().ns3d_server_engine(self.h)),
                                                 out.ctypes.data_as(C.POINTER(C.c_double))))
        return _gpu.Engine.unpack_stats(out)

    def set_graphs(self, on: bool):
        """Chunks of the server's engine as HIP graphs (default) or launched
        kernel by kernel (see gpu.Engine)."""
        _gpu._check(_gpu.lib().ns3d_engine_set_graphs(C.c_void_p(lib().ns3d_server_engine(self.h)), int(bool(on))))

    def set_schedule(self, chunk: int, lanes: int, graphs: bool = True):
        """The engine's schedule (only while nothing is queued or running)."""
        _gpu.Engine.set_schedule_of(C.c_void_p(lib().ns3d_server_engine(self.h)), chunk, lanes, graphs)

    def set_policy(self, largest_first: bool = False, max_rows: float = 0.0):
        """Admission order (arrival, or largest first) and the cap on fluid
        cells in flight (0: none); see ns3d_server_set_policy."""
        if lib().ns3d_server_set_policy(self.h, int(bool(largest_first)), float(max_rows)) != 0:
            raise _gpu.NS3DError(lib().ns3d_server_last_error().decode())

    def schedule(self) -> dict:
        return _gpu.Engine.schedule_of(C.c_void_p(lib().ns3d_server_engine(self.h)))

    def close(self):
        if getattr(self, "h", None):
            lib().ns3d_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_batch(images, params, engine=None, chunk=8, graphs=True):
    """Solves several images concurrently: one thread per image runs its
    driver, their BiCGSTAB solves batched on one engine (created here unless
    given; graphs: its chunks submitted as HIP graphs).  params: one dict per
    image (or one dict for all).  Returns [(rows, summary)] in order;
    summaries carry the final fields."""
    import threading
    if isinstance(params, dict):
        params = [params] * len(images)
    own = engine is None
    eng = _gpu.Engine(device=params[0].get("device", 0) if params else 0, chunk=chunk, graphs=graphs) \
        if own else engine
    solvers = [Solver(img, **p) for img, p in zip(images, params)]
    out = [None] * len(solvers)
    errs = []

    def work(i):
        try:
            rows, summ = solvers[i].run()
            summ["fields"] = solvers[i].fields()
            out[i] = (rows, summ)
        except Exception as e:  # reported after the join
            errs.append((i, e))

    try:
        for s in solvers:
            s.attach(eng)
        threads = [threading.Thread(target=work, args=(i,)) for i in range(len(solvers))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    finally:
        for s in solvers:
            s.close()
        if own:
            stats = eng.stats()
            eng.close()
    if errs:
        raise errs[0][1]
    if own:
        for r in out:
            r[1]["engine"] = stats
    return out


def solve(img, **params):
    s = Solver(img, **params)
    try:
        rows, summ = s.run()
        summ["fields"] = s.fields()
        return rows, summ
    finally:
        s.close()
