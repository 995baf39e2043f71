/*
 * ns3d_host.h -- C ABI of libns3d_host.so: a whole permeability solve (one of
 * the four outer-loop drivers over one image) for callers that do not want
 * to drive the per-method ABI of ns3d.h themselves.
 *
 * It replaces the reference's entry point: NavierStokesBase construction
 * from an image + parameters (src/NSBase.cxx:50-194) followed by
 * XSOLVER(...).Solve(base) (src/navierstokes3D.cxx:516-542).  The drivers are
 * host C++ (ns3d_amd/host/) calling libns3d_gpu.so.
 */
#ifndef NS3D_HOST_H
#define NS3D_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { NS3D_METHOD_SIMPLEC = 0, NS3D_METHOD_SIMPLER = 1, NS3D_METHOD_SIMPLERR = 2, NS3D_METHOD_IDEAL = 3 };

/* The keys of the reference's .ini file (navierstokes3D.cxx:162-201), with
 * the same meaning and defaults (see ns3d_default_run_params). */
typedef struct {
  double dx, dy, dz;            /* dy, dz <= 0: same as dx */
  double mu, rho, p0, gradp;
  double alpha, beta;
  double vsolvertol, psolvertol, pcsolvertol, simpletol, deltaRMR;
  int32_t maxitsolver, maxitSIMPLE, N1, N2, save_step;
  int32_t porevalue, direction, border;
  int32_t method;               /* NS3D_METHOD_* */
  int32_t verbose;              /* 1: print like the reference (stdout); 2: one
                                   line per outer iteration on stderr */
  int32_t write_files;          /* 1: write INFO/RES logs and .vti fields */
  int32_t device;               /* GPU index, -1 = current */
  const char* outdir;           /* directory for the output files */
  const char* name;             /* image file name (output names derive from it) */
  int32_t precision;            /* NS3D_PRECISION_FP64 (default) or _MIXED (ns3d.h) */
  int32_t mixed_min_rows;       /* mixed form: systems with fewer rows are solved in
                                   FP64 (ns3d_set_mixed_min_rows; 0: none) */
  double inner_tol;             /* mixed form: relative tolerance of the FP32 sweeps */
  int32_t mg_min_rows;          /* mixed form: pressure systems from this many rows
                                   on are swept by multigrid CG (ns3d_set_multigrid;
                                   0: never; default 200000) */
  int32_t pad_params;
} ns3d_run_params;

typedef struct {
  int32_t iterations;           /* outer iterations done (the drivers' it) */
  int32_t converged;            /* CheckConvergence returned true */
  double K, Q, v_moy, RMR;      /* last permeability / residual */
  double normdiv, reldiv, maxdiv;
  double seconds;               /* wall time of Solve */
  int64_t np, nu, nv, nw;
  int64_t solver_iterations[4]; /* summed BiCGSTAB iterations: u, v, w, p */
  double phase_seconds[8];      /* momentum asm, velocity solves, hats, p asm,
                                   p solves, corrections, diagnostics, other */
  int64_t solver_work[4];       /* summed executed BiCGSTAB iterations (a zero
                                   right-hand side counts 0, not maxit) */
  int32_t fallback;             /* served requests: 1 if an FP64 run replaced or
                                   continued the mixed-precision one */
  int32_t pad;
} ns3d_run_result;

typedef struct ns3d_solver ns3d_solver;

void ns3d_default_run_params(ns3d_run_params* p);

/* img is the raw image (voxel == porevalue is fluid), C order (nx, ny, nz)
 * with x slowest; rotation and border are applied from the parameters. */
int ns3d_solver_create(const ns3d_run_params* p, const uint8_t* img, int nx, int ny, int nz,
                       ns3d_solver** out);
/* Grid after rotation/border: dims[3]; counts[4] = np, nu, nv, nw. */
int ns3d_solver_info(ns3d_solver* s, int dims[3], int64_t counts[4]);
/* Runs the driver.  rows (may be NULL) receives up to max_rows RES rows of 14
 * doubles: it Q v_moy RMR K gradP normdiv reldiv velocityNorm maxdiv it_u
 * it_v it_w it_pc (it_pc = 0 before the first pressure solve). */
int ns3d_solver_run(ns3d_solver* s, double* rows, int max_rows, int* nrows, ns3d_run_result* res);
/* Final fields on the node sets (lengths from ns3d_solver_info). */
int ns3d_solver_fields(ns3d_solver* s, double* p, double* u, double* v, double* w);
/* Warm start: initial p, u, v, w on the node sets (lengths from
 * ns3d_solver_info) for the next run, in place of the linear pressure
 * profile and zero velocities (NULL keeps that default for the field) --
 * the in-memory form of the pressurefield / vxfield / vyfield / vzfield
 * inputs (NSBase.cxx:2575-2690). */
int ns3d_solver_warm_start(ns3d_solver* s, const double* p, const double* u, const double* v, const double* w);
void ns3d_solver_destroy(ns3d_solver* s);
/* The GPU context behind a solver (for ns3d_timing / ns3d_matrix_nnz). */
struct ns3d_ctx* ns3d_solver_context(ns3d_solver* s);
const char* ns3d_host_last_error(void);

/* ---- serving (no counterpart in the reference) ---------------------------
 * A server takes permeability requests as they arrive.  max_slots worker
 * threads run requests concurrently; a queued request starts (in arrival
 * order) when a slot is free and its projected device footprint
 * (ns3d_footprint) fits in what budget_bytes has left.  All running requests
 * share one solve engine (ns3d_engine_create with `chunk`), so their
 * BiCGSTAB solves run batched and a new request joins at the next chunk
 * boundary.  fallback (NS3D_FALLBACK_*): what happens to a mixed-precision
 * request whose run fails or ends unconverged (result field `fallback` is 1
 * when an FP64 run replaced or continued it).  Its linear solves meet the
 * same FP64 tolerances as the plain path, so a mixed run that stops at its
 * outer-iteration cap is where the FP64 run would be too; continuing it in
 * FP64 (mode 2) is for callers who want that second opinion.  The image is copied
 * at submission; outdir/name strings too.  times[4] of poll/wait: submit,
 * start and end time (seconds since the server was created) and the
 * projected footprint in bytes. */
enum {
  NS3D_FALLBACK_NONE = 0,        /* the mixed result stands as it is */
  NS3D_FALLBACK_FAILED = 1,      /* a failed run or non-finite K / residual: rerun in FP64 */
  NS3D_FALLBACK_UNCONVERGED = 2  /* also: unconverged at the cap -> continue in FP64 from its fields
                                    with max(10, maxitSIMPLE / 4) more outer iterations */
};
typedef struct ns3d_server ns3d_server;
int ns3d_server_create(int device, double budget_bytes, int max_slots, int chunk, int fallback,
                       ns3d_server** out);
/* 0 queued; 1 bad arguments; 2 error; 3 the request alone exceeds the budget. */
int ns3d_server_submit(ns3d_server* s, const ns3d_run_params* p, const uint8_t* img, int nx, int ny, int nz,
                       int64_t* ticket);
/* 0 done (res, times filled); 1 queued or running; -1 unknown ticket; -2
 * the request failed (message in ns3d_server_last_error, times filled). */
int ns3d_server_poll(ns3d_server* s, int64_t ticket, ns3d_run_result* res, double times[4]);
int ns3d_server_wait(ns3d_server* s, int64_t ticket, ns3d_run_result* res, double times[4]);
/* out[16]: submitted, queued, running, done, failed, fallbacks, bytes in use,
 * peak bytes in use, peak running, budget, seconds since creation. */
int ns3d_server_stats(ns3d_server* s, double out[16]);
/* The server's solve engine (for ns3d_engine_stats). */
struct ns3d_engine* ns3d_server_engine(ns3d_server* s);
/* Admission policy.  order 0 (default): requests start in arrival order;
 * 1: the largest queued request (fluid cells) first, which shortens a burst
 * whose long requests would otherwise start late.  Either way a request that
 * does not fit is skipped for the next one that does.  max_rows > 0 caps
 * the fluid cells of the running requests (a request alone always runs):
 * once the GPU is full, more requests in flight only stretch every
 * request's iteration cadence.  0 (default): no cap.  Returns 0, or 1 on
 * bad arguments. */
int ns3d_server_set_policy(ns3d_server* s, int order, double max_rows);
/* Finishes every queued and running request, then frees the server. */
void ns3d_server_destroy(ns3d_server* s);
const char* ns3d_server_last_error(void);

/* Replays the outer loop's relaxation control (RelaxControl,
 * NSBase.cxx:196-301, after the drivers' tracking of the best residual,
 * SIMPLEC.cxx:155-166) over n relative mass residuals taken as outer
 * iterations 0..n-1; after iteration i: alpha_out[i], beta_out[i] and
 * events[i] (bits: 1 the residual jumped above 2.5x its best, 2 the 30-entry
 * window was flat, 4 it was rising).  beta_cst as the drivers pass it
 * (SIMPLERR 0, the others 1).  No GPU is used: for studying how a residual
 * history drives the relaxation factors.  Returns 0, or 1 on bad arguments. */
int ns3d_relax_replay(const double* rmr, int n, int beta_cst, double alpha0, double beta0, double deltaRMR,
                      double* alpha_out, double* beta_out, int* events);

/* Reads a binary image file -- XML ImageData (.vti: appended raw or base64
 * data, zlib or none, UInt32/UInt64 headers) or legacy STRUCTURED_POINTS
 * (.vtk) -- and binarises it as readvtkfile_uint8 does (vtkio.cxx:96-150):
 * voxel == porevalue -> 0 (fluid), any other value -> 1.  out (may be NULL:
 * only dims is filled) gets dims[0]*dims[1]*dims[2] bytes, C order (x
 * slowest) as the reference's Blitz++ array.  Returns 0, or 1 with
 * ns3d_host_last_error(). */
int ns3d_read_image(const char* path, int porevalue, uint8_t* out, int dims[3]);

/* Reads a .ini file with the reference's grammar (ReadInput,
 * navierstokes3D.cxx:266-463); the image path is copied to image (size cap).
 * Returns 0 on success, 1 on a fatal input error (message on stderr). */
int ns3d_read_ini(const char* path, ns3d_run_params* p, char* image, int cap);

#ifdef __cplusplus
}
#endif

#endif /* NS3D_HOST_H */
