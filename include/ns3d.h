/*
 * ns3d.h -- C ABI of the MI355X implementation of NS3D's SIMPLE-family outer
 * iteration (libns3d_gpu.so).
 *
 * One context holds one image on one GPU: the fluid-index arrays, the four
 * staggered linear systems (u, v, w momentum and the pressure / pressure-
 * correction system) in device memory, and every field the outer-loop drivers
 * use.  Each entry point below replaces one method of the reference's
 * NavierStokesBase class (src/NSBase.h:63-594), with the same argument
 * meaning; the outer-loop drivers (SIMPLEC/SIMPLER/SIMPLERR/IDEAL) are host
 * code calling these entry points in the reference's order.
 *
 * Conventions
 *  - Images are uint8, 0 = fluid, 1 = solid, C order with shape (nx, ny, nz):
 *    byte (i*ny + j)*nz + k is voxel (x=i, y=j, z=k).  Rotation for flow along
 *    y/z and the x border are applied by the caller (ns3d_prepare_image).
 *  - Index grids are int32, -1 = no node; u's grid has shape (nx+1, ny, nz).
 *  - All floating-point state is FP64.  Host buffers are caller-owned and are
 *    only read or written during the call.
 *  - Errors are returned as ns3d_status; ns3d_last_error() gives the message
 *    of the last failure on the calling thread.  Nothing throws across the ABI.
 *  - One host thread per context.  Different contexts may be driven from
 *    different threads; each context owns its own HIP stream.
 */
#ifndef NS3D_H
#define NS3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ns3d_ctx ns3d_ctx;

typedef enum {
  NS3D_OK = 0,
  NS3D_ERR_INVALID = 1, /* bad argument or state */
  NS3D_ERR_OOM = 2,     /* device allocation failed */
  NS3D_ERR_HIP = 3,     /* HIP runtime error */
  NS3D_ERR_NODEV = 4    /* no usable GPU */
} ns3d_status;

/* Linear systems. */
enum { NS3D_SYS_U = 0, NS3D_SYS_V = 1, NS3D_SYS_W = 2, NS3D_SYS_P = 3 };

/* Velocity triples held by a context: (u, v, w), (u_temp, v_temp, w_temp),
 * (u_hat, v_hat, w_hat) -- the drivers' u/u_temp/u_hat vectors
 * (SIMPLER.h:62-83, IDEAL.h:67-89). */
enum { NS3D_SET_VEL = 0, NS3D_SET_TEMP = 1, NS3D_SET_HAT = 2 };

/* Vectors for ns3d_get_vector / ns3d_set_vector. */
enum {
  NS3D_VEC_U = 0, NS3D_VEC_V, NS3D_VEC_W,
  NS3D_VEC_UT, NS3D_VEC_VT, NS3D_VEC_WT,
  NS3D_VEC_UH, NS3D_VEC_VH, NS3D_VEC_WH,
  NS3D_VEC_P, NS3D_VEC_PC, NS3D_VEC_B
};

/* Pressure-type systems (ns3d_assemble_pressure). */
enum { NS3D_PRESSURE_FULL = 0, NS3D_PC_SIMPLE = 1, NS3D_PC_SIMPLEC = 2 };

/* Arithmetic of the linear solves (ns3d_set_precision). */
enum { NS3D_PRECISION_FP64 = 0, NS3D_PRECISION_MIXED = 1 };

/* Velocity corrections (ns3d_correct_velocity). */
enum { NS3D_CORR_SIMPLE = 0, NS3D_CORR_SIMPLEC = 1, NS3D_CORR_SIMPLER = 2 };

/* Result of one linear solve, with Eigen's BiCGSTAB meaning: iterations(),
 * info() (0 Success, 1 NumericalIssue, 2 NoConvergence, 3 InvalidInput) and
 * error() = |r|/|b|.  work counts the iterations actually executed: 0 for a
 * zero right-hand side (which Eigen reports as maxit iterations), summed over
 * the tightened retries of ns3d_compute_pressure_field. */
typedef struct {
  int64_t iterations;
  int32_t info;
  int32_t restarts;
  double error;
  int64_t work;
} ns3d_solve_info;

/* Physical and numerical constants of one solve.  The per-iteration
 * relaxation factors alpha and beta are passed to the calls that use them,
 * since the drivers change them (NSBase.cxx:196-301). */
typedef struct {
  double dx, dy, dz;   /* grid spacing (dy, dz already defaulted to dx) */
  double mu, rho;      /* viscosity, density */
  double p0, p1;       /* inlet / outlet pressure (after NSBase.cxx:155-161) */
  int32_t border;      /* x border added by the caller */
  int32_t reserved;
} ns3d_params;

/* Last error message on this thread ("" if none). */
const char* ns3d_last_error(void);

/* Number of GPUs visible; <0 on a runtime error. */
int ns3d_device_count(void);

/* Rotation (direction 2: y->x, 3: z->x) and fluid border along x, on the host.
 * Replaces NavierStokesBase::Rotation (NSBase.cxx:2451-2492) and AddBorder
 * (NSBase.cxx:2425-2448; pads with 0 = fluid).  With out == NULL only dims is
 * filled; out must hold dims[0]*dims[1]*dims[2] bytes. */
ns3d_status ns3d_prepare_image(const uint8_t* img, int nx, int ny, int nz, int direction,
                               int border, uint8_t* out, int dims[3]);

/* Uploads the image and builds the fluid-index arrays with a device scan,
 * bit-exact with NavierStokesBase::InitIndexArrays (NSBase.cxx:2495-2573),
 * then the structure of the four systems.  device < 0 keeps the current
 * device. */
ns3d_status ns3d_create(const uint8_t* img, int nx, int ny, int nz, int device, ns3d_ctx** out);
void ns3d_destroy(ns3d_ctx* ctx);

/* np, nu, nv, nw (NSBase.h:76). */
ns3d_status ns3d_counts(ns3d_ctx* ctx, int64_t out[4]);

/* Copies pindex/vindex/windex (nx*ny*nz each) and uindex ((nx+1)*ny*nz) to
 * host buffers; any pointer may be NULL (NSBase.h:90-93). */
ns3d_status ns3d_get_index(ns3d_ctx* ctx, int32_t* p, int32_t* u, int32_t* v, int32_t* w);

/* Sets the constants of the solve (NavierStokesBase constructor,
 * NSBase.cxx:50-194). */
ns3d_status ns3d_set_params(ns3d_ctx* ctx, const ns3d_params* params);

/* Zeroes every velocity triple, p, pc and b (driver initialisation, e.g.
 * SIMPLEC.cxx:95-106). */
ns3d_status ns3d_zero_fields(ns3d_ctx* ctx);

/* Linear pressure profile p = (p1-p0)*(i+1)/nx + p0 (Initpvector,
 * NSBase.cxx:2671-2681). */
ns3d_status ns3d_init_pressure_linear(ns3d_ctx* ctx);

/* Host <-> device copies of one vector (length from ns3d_counts). */
ns3d_status ns3d_set_vector(ns3d_ctx* ctx, int which, const double* host, int64_t n);
ns3d_status ns3d_get_vector(ns3d_ctx* ctx, int which, double* host, int64_t n);

/* dst triple := src triple (the drivers' u_temp = u, SIMPLER.cxx:161-163). */
ns3d_status ns3d_copy_velocity(ns3d_ctx* ctx, int dst_set, int src_set);

/* AssembleAllMomentum (NSBase.cxx:853-884): u, v, w rows from the VEL triple
 * (AssembleMomentum_x/y/z, NSBase.cxx:886-1442) plus the Jacobi setup of the
 * three solvers (compute()). */
ns3d_status ns3d_assemble_momentum(ns3d_ctx* ctx, double alpha);

/* ComputeVelocityField (NSBase.cxx:1773-1884): RHS (Assemble_RHS_u/v/w,
 * NSBase.cxx:1445-1546), BiCGSTAB from the current VEL triple, revert on
 * 0 iterations or failure, optional clamp to +-delta_v_max.  info[3] gets u,
 * v, w (w untouched when nw == 0). */
ns3d_status ns3d_compute_velocity_field(ns3d_ctx* ctx, double alpha, int limit, double delta_v_max,
                                        double tol, int64_t maxit, ns3d_solve_info info[3]);

/* Compute_u_hat / v_hat / w_hat (NSBase.cxx:1548-1755) into the HAT triple,
 * with the off-diagonal products taken on uc_set and relaxation on uold_set. */
ns3d_status ns3d_compute_hat(ns3d_ctx* ctx, double alpha, int uc_set, int uold_set);

/* AssemblePressure (kind FULL, NSBase.cxx:478-610), AssemblePressureCorrection
 * SIMPLE (:734-851) or SIMPLEC (:612-732) on the velocities of vel_set, plus
 * the Jacobi setup of the pressure solver.  relax/beta apply to FULL only. */
ns3d_status ns3d_assemble_pressure(ns3d_ctx* ctx, int kind, int relax, double beta, int vel_set);

/* ComputePressureCorrection (NSBase.cxx:2337-2365): pc from the guess pc. */
ns3d_status ns3d_compute_pressure_correction(ns3d_ctx* ctx, double tol, int64_t maxit,
                                             ns3d_solve_info* info);

/* ComputePressureField (NSBase.cxx:2383-2423): solve from p, tighten the
 * tolerance by 10x while a solve takes 0 iterations, keep the result only if
 * the last solve iterated. */
ns3d_status ns3d_compute_pressure_field(ns3d_ctx* ctx, double tol, int64_t maxit,
                                        ns3d_solve_info* info);

/* CorrectVelocityFieldSIMPLE / SIMPLEC (NSBase.cxx:2082-2263): target_set +=
 * clamped -grad(pc)/a.  CorrectVelocityFieldSIMPLER (:2265-2335): target_set
 * = hat - grad(p)/a with p_0/p_1 at the inlet/outlet. */
ns3d_status ns3d_correct_velocity(ns3d_ctx* ctx, int kind, double delta_v_max, int target_set,
                                  double p_0, double p_1);

/* CorrectPressureField (NSBase.cxx:2367-2380): p += beta*pc; out = beta*max,
 * beta*min, beta*norm of pc. */
ns3d_status ns3d_correct_pressure(ns3d_ctx* ctx, double beta, double out[3]);

/* ComputeDiv (NSBase.cxx:1958-2080) on vel_set: out = MaxDeltaV, normdiv,
 * velocityNorm, maxdiv. */
ns3d_status ns3d_compute_div(ns3d_ctx* ctx, int vel_set, double out[4]);

/* ComputePermeability (NSBase.cxx:1886-1956): out = Q, v_moy, K, p_inlet,
 * p_outlet. */
ns3d_status ns3d_compute_permeability(ns3d_ctx* ctx, double out[5]);

/* Matrix export for tests: diag[n], off[6*n] (E,W,N,S,T,B blocks of n),
 * nbr[6*n] (neighbour rows, the row itself where absent), b[n] when sys is
 * NS3D_SYS_P (else may be NULL), rowsum[n] for momentum systems (may be NULL). */
ns3d_status ns3d_get_matrix(ns3d_ctx* ctx, int sys, double* diag, double* off, int32_t* nbr,
                            double* rowsum);

/* Stand-alone Jacobi-BiCGSTAB (Eigen semantics) on a 7-point matrix given as
 * diag[n], off[6*n], nbr[6*n] (same layout as ns3d_get_matrix); x holds the
 * guess on entry and the solution on exit.  Used to test the solver alone. */
ns3d_status ns3d_bicgstab_stencil(int n, const double* diag, const double* off, const int32_t* nbr,
                                  const double* b, double* x, double tol, int64_t maxit,
                                  ns3d_solve_info* info);

/* ns3d_bicgstab_stencil without any device allocation: the same solve, the
 * same kernels, bit-identical results, with diag, off, nbr, b and x in device
 * memory of the current device and the work queued on `stream` (a
 * hipStream_t; NULL = the null stream), which the call synchronizes.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (serving loops, harnesses that forbid hipMalloc).  The
 * workspace (coefficients, M^-1, the BiCGSTAB vectors, partials, state) is
 * one fixed array per GPU, held until the process ends, sized for
 * n <= ns3d_noalloc_max_rows() (1048576 rows, 120.1 MiB); a larger n is
 * NS3D_ERR_INVALID, as are pointers that are not device memory of the
 * current device and neighbour rows outside [0, n) (checked on the device
 * before the solve).  Calls that overlap on one GPU share that array and
 * race: serialize them. */
ns3d_status ns3d_bicgstab_stencil_noalloc(int n, const double* diag, const double* off, const int32_t* nbr,
                                          const double* b, double* x, double tol, int64_t maxit,
                                          ns3d_solve_info* info, void* stream);

/* The row capacity of ns3d_bicgstab_stencil_noalloc's workspace. */
ns3d_status ns3d_noalloc_max_rows(int64_t* rows);

/* Linear-solve arithmetic.  NS3D_PRECISION_FP64 (default) is the reference's
 * Eigen BiCGSTAB in FP64.  NS3D_PRECISION_MIXED keeps assembly, residuals,
 * corrections and reductions in FP64 and solves each system by FP64 defect
 * correction around FP32-storage BiCGSTAB sweeps (FP32 coefficients and
 * vectors, FP64 arithmetic in registers): r = b - A x in FP64, inner solve of
 * A d = r to max(inner_tol, tol |b| / 2|r|), x += d, until |b - A x| <= tol |b|.
 * Reported iterations are summed inner iterations; error is the true FP64
 * relative residual.  When a correction step reduces |b - A x| by less than
 * half (FP32 has reached its accuracy on that system), the solve finishes
 * with the plain FP64 BiCGSTAB from the current x.  This form is not part of
 * the reference. */
ns3d_status ns3d_set_precision(ns3d_ctx* ctx, int mode, double inner_tol);

/* Mixed form only: systems with fewer than `rows` rows are solved with the
 * plain FP64 BiCGSTAB.  FP32 sweeps save HBM bandwidth, which small systems
 * hardly spend (their iterations are bound by launch latency), while every
 * defect-correction step costs a synchronisation.  0 (default): every system
 * in the mixed form.  Not part of the reference. */
ns3d_status ns3d_set_mixed_min_rows(ns3d_ctx* ctx, int64_t rows);

/* Mixed form only: pressure-type systems (p' and p) with at least `rows`
 * rows get FP32 sweeps of conjugate gradients preconditioned by one
 * aggregation-multigrid V-cycle (2x2x2 blocks per level, Galerkin coarse
 * operators rebuilt after every pressure assembly) instead of Jacobi
 * BiCGSTAB; the FP64 defect correction around them, its tolerance and its
 * true-residual contract are unchanged.  The pressure matrices are
 * symmetric and definite (negated M-matrices), so CG applies.  0: never;
 * default 200000 (smaller systems are cheaper batched in the engine).  Not
 * part of the reference (its solver is Eigen's Jacobi BiCGSTAB, which the
 * plain FP64 form keeps). */
ns3d_status ns3d_set_multigrid(ns3d_ctx* ctx, int64_t rows);
/* Levels of the pressure system's multigrid hierarchy (0 until the first
 * multigrid sweep has built it) and their row counts (up to 12 entries
 * written to rows, which may be NULL). */
ns3d_status ns3d_multigrid_levels(ns3d_ctx* ctx, int* levels, int64_t* rows);
/* Level `level` (1 .. levels-1) of that hierarchy as it is on the device
 * (for checks of the aggregation and the Galerkin operators): the linear
 * cell of each row on the level's grid (each dimension of level l+1 is
 * (d+1)/2 of level l's) into cell[n], and the FP32 diagonal, E, N and T
 * couplings into coef[4][n] (either may be NULL).  Not part of the
 * reference. */
ns3d_status ns3d_multigrid_level(ns3d_ctx* ctx, int level, int32_t* cell, float* coef);
/* Device time of the multigrid sweeps' finest-level kernels (with
 * ns3d_timing(ctx, 1): the first iteration of each sweep is launched with
 * dispatch events).  out[3k..3k+2] = milliseconds, sampled launches and
 * algorithmic bytes of kernel k: 0 the pre-smoothed residual of the first
 * restriction (nnz*8 + 12n), 1 the last prolongation with smoothing (nnz*8
 * + 20n + 4 per coarse row), 2 the direction update with A p (nnz*8 + 16n),
 * 3 the x / r update (24n); n pressure rows, nnz their stored entries.  Not
 * part of the reference. */
ns3d_status ns3d_multigrid_timing(ns3d_ctx* ctx, double out[12]);

/* Measurement: a plain device copy of `bytes` (read once, written once) run
 * `reps` times on `device` (-1: current), timed with events; *gbs = bytes
 * read + written per second / 1e9.  The achievable streaming bandwidth that
 * the kernels' HBM fractions are read against, and (under rocprofv3 --pmc)
 * the known-size transfer that calibrates FETCH_SIZE / WRITE_SIZE.
 * Allocates and frees its own buffers.  Not part of the reference. */
ns3d_status ns3d_copy_bandwidth(int device, int64_t bytes, int reps, double* gbs);

/* Device time of the fused BiCGSTAB kernels (enable with ns3d_timing(ctx,
 * 1)): one iteration per queued chunk of iterations is sampled, its three
 * kernels launched with start/stop events taken from their own dispatches
 * on the context's stream; samples of launches after a solve has stopped
 * are dropped.  out[36]: for storage precision q (0 FP64, 1 FP32), system
 * kind m (0 momentum, 1 pressure) and kernel k (0: A = p update + SpMV +
 * r0.v, 1: B = s + SpMV + t.s/t.t, 2: C = x/r update + r0.r/r.r),
 * out[18q + 9m + 3k] = milliseconds, sampled launches, algorithmic bytes
 * (nnz (s + 4) per SpMV plus s bytes per row per vector stream: A 7, B 4,
 * C 9; s = 8 or 4). */
ns3d_status ns3d_timing(ns3d_ctx* ctx, int enable);
ns3d_status ns3d_timing_read(ns3d_ctx* ctx, double out[36]);

/* ---- batched solves (no counterpart in the reference) -------------------
 * An engine owns one stream and a dispatcher thread.  Contexts attached to
 * it hand every BiCGSTAB solve (FP64, and the FP32 inner solves of the mixed
 * form) to the engine, which runs the solves of all attached contexts that
 * are waiting together: one launch per kernel phase over a table of tiles of
 * all of them, solves joining and leaving at chunk boundaries (every `chunk`
 * iterations).  The calling thread blocks until its solve has stopped; each
 * attached context must be driven by its own host thread for solves to
 * overlap.  Results are those of the context's own path (same kernels and
 * tiles, same summation order), alone or in a batch.  The engine must
 * outlive the contexts attached to it, or be detached from them first. */
typedef struct ns3d_engine ns3d_engine;
ns3d_status ns3d_engine_create(int device, int chunk, ns3d_engine** out);
void ns3d_engine_destroy(ns3d_engine* eng);
/* eng NULL detaches.  Attaching also gives the context two more solver
 * workspaces, so that the three momentum solves of
 * ns3d_compute_velocity_field go to the engine together.  While a context
 * is the only one attached to its engine, its solves run on its own stream
 * as if detached (nothing to batch with; same arithmetic), and go to the
 * engine again once another context attaches.  While attached, the
 * context's own kernels and copies run on the engine's context stream (one
 * for all attached contexts, on a hardware queue apart from the engine's
 * lanes); attach and detach first wait for the context's queued work. */
ns3d_status ns3d_attach_engine(ns3d_ctx* ctx, ns3d_engine* eng);
/* on (default): each chunk of iterations is submitted as one HIP graph per
 * engine stream (reused while the grids stay the same; the kernels read the
 * table sizes from device memory), except every 8th round, which is launched
 * kernel by kernel with dispatch events for ns3d_engine_stats; off: every
 * round kernel by kernel.  Results are the same either way. */
ns3d_status ns3d_engine_set_graphs(ns3d_engine* eng, int on);
/* The engine's schedule: chunk (iterations queued per round, 1..64), lanes
 * (streams the solves are spread over, 1..4; default 3) and graphs (as
 * ns3d_engine_set_graphs).  Only while no solve is pending or running
 * (NS3D_ERR_INVALID otherwise).  No choice changes any result; an autotuner
 * (ns3d_amd/autotune.py) picks them by timing a request sample.  get:
 * out = {chunk, lanes, graphs}. */
ns3d_status ns3d_engine_set_schedule(ns3d_engine* eng, int chunk, int lanes, int graphs);
ns3d_status ns3d_engine_get_schedule(ns3d_engine* eng, int out[3]);
/* out[32]: chunks, launches, solves done, max active solves, sum of rows x
 * iterations; then per precision q (0 FP64, 1 FP32) and kernel k (A, B, C),
 * out[5 + 9q + 3k] = sampled milliseconds, sampled launches, algorithmic
 * bytes of the solves that iterated in them (one iteration per chunk is
 * sampled, each launch timed by its own dispatch); out[23..25] = whole
 * chunks (the engine runs its solves on two streams; from the first A start
 * to the last C stop over all of them): milliseconds, count, algorithmic
 * bytes of the iterations they executed; out[26..28] = the dispatcher
 * thread's milliseconds admitting and queueing rounds, processing their
 * states, and waiting for the round in flight; out[29..30] = over the
 * sampled iterations, the tiles of solves that iterated and all tiles
 * launched. */
ns3d_status ns3d_engine_stats(ns3d_engine* eng, double out[32]);

/* Upper bound of the device bytes a context of an nx x ny x nz image with
 * nfluid fluid voxels allocates (mixed != 0: with the mixed form's buffers,
 * the multigrid hierarchy included when the image could reach the default
 * ns3d_set_multigrid threshold; lanes 1: alone, 3: attached to an engine),
 * for admission control; ns3d_ctx_bytes gives what a context did request. */
ns3d_status ns3d_footprint(int nx, int ny, int nz, int64_t nfluid, int mixed, int lanes, double* bytes);
ns3d_status ns3d_ctx_bytes(ns3d_ctx* ctx, double* bytes);

/* Stored entries (diagonal + existing neighbours) of the u, v, w, p systems. */
ns3d_status ns3d_matrix_nnz(ns3d_ctx* ctx, int64_t out[4]);

/* Synchronises the context's stream. */
ns3d_status ns3d_sync(ns3d_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* NS3D_H */
