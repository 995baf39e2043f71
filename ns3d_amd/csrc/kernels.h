// kernels.h -- host-callable launchers of the HIP kernels (one per .hip file
// group) and the device-side BiCGSTAB state.  Included by the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "device.h"

namespace ns3d {

// ---- fluid-index compaction (kernels_index.hip) ----------------------------
struct ScanBuffers {
  int64_t ncell;          // nx*ny*nz
  int nblocks;            // scan blocks of kScanCells cells
  int32_t* blockcounts;   // [4][nblocks]
  int32_t* blockoffsets;  // [4][nblocks]
  int32_t* totals;        // [4]: np, nu, nv, nw
};
constexpr int kScanThreads = 256;
constexpr int kScanCpt = 16;
constexpr int kScanCells = kScanThreads * kScanCpt;

hipError_t launch_scan_count(const uint8_t* img, int nx, int ny, int nz, ScanBuffers& sb, hipStream_t s);
hipError_t launch_scan_write(const uint8_t* img, int nx, int ny, int nz, ScanBuffers& sb,
                             int32_t* pg, int32_t* ug, int32_t* vg, int32_t* wg, int32_t* pcell,
                             int32_t* ucell, int32_t* vcell, int32_t* wcell, hipStream_t s);
hipError_t launch_build_nbr(const Geo& g, int sys, const int32_t* cell, int n, int32_t* nbr,
                            unsigned long long* nnz, hipStream_t s);

// ---- assembly (kernels_assemble.hip) ----------------------------------------
struct Phys {
  double dx, dy, dz, mu, rho, p0, p1, alpha, beta, sf;
};
struct Vel3 {
  const double* u;
  const double* v;
  const double* w;
};
hipError_t launch_assemble_momentum(int sys, const Geo& g, SysView S, Vel3 vc, Phys ph, hipStream_t s);
hipError_t launch_assemble_pressure(const Geo& g, SysView P, SysView U, SysView V, SysView W, Vel3 vc,
                                    const double* pold, double* b, int kind, int relax, Phys ph,
                                    hipStream_t s);
hipError_t launch_rhs(int sys, const Geo& g, SysView S, const double* p, const double* xold, double* b,
                      Phys ph, hipStream_t s);
hipError_t launch_hat(int sys, const Geo& g, SysView S, const double* xc, const double* xold, double* hat,
                      Phys ph, hipStream_t s);

// ---- BiCGSTAB (kernels_solver.hip) ------------------------------------------
enum BStatus { BS_RUNNING = 0, BS_DONE = 1, BS_RESTART = 2, BS_ZERO_RHS = 3 };
// Solver scalars in device memory.  The row kernels of one iteration never
// read a field that their own blocks write: per-iteration scalars live in
// two slots indexed by the parity of the executed-iteration count, it_a is
// written by kernel A and read by B and C, it_c is written by C and read by
// A.  Eigen's iteration counter is it_c - ibase (ibase moves on the first
// restart, which resets Eigen's i to 0).
struct BState {
  double rho[2], alpha[2], w[2];
  double rho_fresh;     // rho of iteration fresh_at, set by init / restart
  double r0_sqnorm, rhs_sqnorm, tol2, tol;
  double rr, r0v, ts, tt;
  double eps2;          // restart threshold factor (SolveDesc::eps2, or Eigen's epsilon^2)
  long long it_a, it_c, ibase, fresh_at, maxit, restarts;
  int status;
  int pad;
};
// Matrix storage as the solver reads it.
//   ML_INDEXED: nbr[6][n] int32 neighbour rows, coef[7][n];
//   ML_PACKED: one 64-bit word of neighbour distances per row (see
//     pack_neighbours), coef[7][n];
//   ML_PACKED_SYM: packed distances, and only the diagonal and the E, N, T
//     coefficients are read: the W/S/B coefficient of a row is the E/N/T
//     coefficient of that neighbour (bit-symmetric pressure matrices).
enum MatLayout { ML_INDEXED = 0, ML_PACKED = 1, ML_PACKED_SYM = 2 };
// Packed neighbour distances: T, B in 2 bits each (1, or 2 next to the
// outlet u plane), N, S in 11 bits, E, W in 19 bits; 0 = no neighbour.
constexpr int kPkTB = 2, kPkNS = 11, kPkEW = 19;

// Matrix view used by the solver, with coefficient storage type T (double
// for the plain path, float for the inner solves of the mixed form).
template <typename T>
struct SMat {
  int n;
  int layout;           // MatLayout
  const int32_t* nbr;   // [6][n] (ML_INDEXED)
  const uint64_t* pk;   // [n] (ML_PACKED*)
  const T* coef;        // [7][n]
  const T* invdiag;     // [n]
  const int* chk;       // plane-band chunk list (build_chunks), or null: contiguous tiles
};

// Plane-band order for large systems (n >= kChunkMinRows).  A plane of such a
// system holds ~1e5 rows, so the E/W gathers of the row kernels reach one
// plane away: with each block sweeping one contiguous tile, a row is read
// again (as a neighbour) ~1e7 rows of traffic later and comes from HBM three
// times.  Instead the rows are cut into kBands bands of j (one per XCD) and
// each band into chunks of at most kBlock rows of one plane, in plane order;
// the blocks of a band sweep its chunks together (block k of bx takes chunks
// k, k + bx, ...), so the planes i-1, i, i+1 of a band are touched within a
// few hundred kilobytes of that XCD's L2 traffic.  Layout of the list:
// chk[0..kBands] = first chunk of each band (chk[kBands] = total), chk[9]
// unused, then int2 {first row, rows} per chunk from chk + 10.
constexpr int kBands = 8;
#ifndef NS3D_CHUNK_MIN_ROWS
#define NS3D_CHUNK_MIN_ROWS 4000000
#endif
constexpr int64_t kChunkMinRows = NS3D_CHUNK_MIN_ROWS;
constexpr int kChunkHeader = 10;
// The chunk list of a system from its row -> cell map (cell = (i*ny + j)*nz + k,
// rows in ascending cell order); empty if the system stays contiguous (fewer
// than min_rows rows, or fewer than kBands lines of j).
std::vector<int> build_chunks(const int32_t* cell, int64_t n, int ny, int nz, int64_t min_rows);
// BiCGSTAB vectors of storage type T; scalars and reductions are FP64.
template <typename T>
struct SWork {
  T* r;
  T* r0;
  T* p[2];
  T* v[2];
  T* t;
  BState* st;
};
using BWork = SWork<double>;

// One linear solve as the batched kernels see it: the matrix, the vectors,
// the unknown and right-hand side, the tolerance and iteration cap, the
// range [t0, t1) of its tiles in the launch's tile table, and its own block
// partials: tile t writes partials[slot * pstride + (t - t0)] (kPartialSlots
// slots: A: r0.v in slot 0; B: t.s, t.t in 1-2; C: r0.r, r.r in 3-4; init /
// restart in 5-6).  Partials belong to the solve, not to the table, so they
// stay valid when the table is rebuilt between two kernels of an iteration.
template <typename T>
struct SolveDesc {
  SMat<T> S;
  SWork<T> w;
  T* x;
  const T* b;
  double tol;
  long long maxit;
  double* partials;
  int pstride;
  int t0, t1;
  int pad;
  // restart threshold: |rho| < eps2 |r0|^2 restarts (0: Eigen's epsilon^2)
  double eps2;
};
// A tile of one solve's system, processed by one block: rows [begin, end)
// (step 0), or chunks begin, begin + step, ... below end of its band's
// plane-band chunk list (step = blocks of the band).
struct Tile {
  int solve;
  int step;
  long long begin, end;
};
// One entry of a batched launch's tile table: a tile with its solve's
// descriptor and iteration base (the solve's executed iterations when the
// chunk was queued) copied in, so a block reaches all its operands with one
// table load and no dependent one.
template <typename T>
struct TileDesc {
  SolveDesc<T> d;
  long long begin, end;
  long long it0;
  int solve;
  int step;
};
// The operands of one launch over a set of solves of storage type T.
// Either tables (d: one descriptor per solve, for the per-solve kernels;
// tiles) or one solve passed by value (single = 1: `one`, tiles of `per`
// rows, one_it0).  q is the launch's index in its chunk, so a
// row kernel knows its iteration (it0 + q) -- and with it the p/v buffer and
// the scalar slot -- without a dependent load; it exits if the solve has
// stopped or its own count disagrees.
constexpr int kPartialSlots = 8;
template <typename T>
struct Batch {
  const SolveDesc<T>* d;
  const TileDesc<T>* tiles;
  // when set, the row kernels read the tile count from here (a table whose
  // size changes between launches of a reused graph) and the grid may have
  // more blocks than tiles
  const int* ntiles_dev;
  int nsolves, ntiles;
  int q;
  int single;
  long long one_it0;
  long long per;
  SolveDesc<T> one;
};
// Batch of the one solve `d` (tiles as make_tiles builds them).
template <typename T>
Batch<T> single_batch(const SolveDesc<T>& d, long long it0, int q);
// Tiles of one solve of n rows: between 1 and kMaxTilesPerSolve tiles of
// a multiple of kBlock rows, at least kTileRows rows each where n allows.
#ifndef NS3D_MAX_TILES
#define NS3D_MAX_TILES 1024
#endif
constexpr int kMaxTilesPerSolve = NS3D_MAX_TILES;
#ifndef NS3D_TILE_ROWS
#define NS3D_TILE_ROWS 1024
#endif
constexpr int kTileRows = NS3D_TILE_ROWS;
// (chunked: a system with a plane-band chunk list, whose tile count is a
// multiple of kBands, at least kBands)
int tiles_for(int64_t n, bool chunked = false);
// Appends the tiles of solve `solve` (n rows) to out; returns its [t0, t1).
// band_off: the system's chunk list header (its first kBands + 1 entries,
// host copy) for plane-band tiles, or null for contiguous ones.
void make_tiles(int solve, int64_t n, const int* band_off, std::vector<Tile>& out, int& t0, int& t1);
// Blocks of a launch over ntiles tiles (a multiple of the XCD count).
int tile_grid(int ntiles);

// init: r = b - A x, r0 = r, p = v = 0, |r|^2 and |b|^2 -> BState (status
// RUNNING, DONE or ZERO_RHS), for every solve of the batch.
template <typename T>
hipError_t launch_bicg_init(Batch<T> B, hipStream_t s);
// restart (solves whose status is RESTART; others exit at once).
template <typename T>
hipError_t launch_bicg_restart(Batch<T> B, hipStream_t s);
// The three row kernels of one iteration.  ev (may be null): {start, stop}
// taken from the kernel's own dispatch (hipExtLaunchKernelGGL).
template <typename T>
hipError_t launch_bicg_A(Batch<T> B, hipStream_t s, hipEvent_t* ev = nullptr);
template <typename T>
hipError_t launch_bicg_B(Batch<T> B, hipStream_t s, hipEvent_t* ev = nullptr);
template <typename T>
hipError_t launch_bicg_C(Batch<T> B, hipStream_t s, hipEvent_t* ev = nullptr);
// Adds the row kernel k (0 A, 1 B, 2 C) over B with `grid` blocks to graph
// g, after `dep` (may be null).
template <typename T>
hipError_t graph_add_bicg(hipGraph_t g, int k, const Batch<T>& B, int grid, hipGraphNode_t dep,
                          hipGraphNode_t* out);
hipError_t launch_invdiag(SysView S, hipStream_t s);
// Packs nbr[6][n] into pk[n]; *bad (zeroed by the caller) becomes nonzero if
// a distance does not fit or a neighbour is not on its expected side.
hipError_t launch_pack_neighbours(const int32_t* nbr, int n, uint64_t* pk, int* bad, hipStream_t s);

// Mixed form: FP32 copy of a system's coefficients and preconditioner.
hipError_t launch_coef_to_f32(const double* coef, const double* invdiag, int64_t n, float* coef32,
                              float* invdiag32, hipStream_t s);
// r = b - A x in FP64, r32 = (float) r; result[0] = |r|^2, result[1] = |b|^2.
// d (may be null): the residual of the trial iterate x + (double) d, x itself
// unchanged (the mixed form applies a correction only once it is accepted).
hipError_t launch_residual64(SMat<double> A, const double* x, const double* b, double* r, float* r32,
                             double* partials, double* result, hipStream_t s, const float* d = nullptr);
// x += (double) d
// x += a d (a = 1: the correction; a = -1: taking it back)
hipError_t launch_add_f32(double* x, const float* d, int64_t n, hipStream_t s, double a = 1.0);
hipError_t launch_fill_f32(float* x, int64_t n, hipStream_t s);

// ---- aggregation multigrid for pressure systems (kernels_mg.hip) ------------
// Rows the coarsest level may hold (its solve runs in one work-group of
// kMgCoarseThreads threads, two rows each), and the most levels a hierarchy
// has.
constexpr int kMgCoarseRows = 512;
constexpr int kMgCoarseThreads = 256;
constexpr int kMgMaxLevels = 12;
// One level: rows numbered in C order of the level's grid (nx, ny, nz);
// level 0 is the pressure system itself (its index grid, row cells, packed
// distances and FP32 coefficients, slots 0 / 1+DE / 1+DN / 1+DT of [7][n]),
// coarser levels store [4][n] (diagonal, E, N, T: slots 0..3).  agg maps a
// row to its row on the next level (null on the coarsest); b and x are the
// level's right-hand side and result.
struct MgLevel {
  int n, nx, ny, nz;
  int sE, sN, sT;
  const int32_t* grid;
  const int32_t* cell;
  const uint64_t* pk;
  const float* coef;
  const float* invd;
  const int32_t* agg;
  float* b;
  float* x;
};
// State of one flexible-CG solve in device memory: scalars of iteration i
// in slot i & 1; it_c (iterations completed) is written by the update
// kernel and read by the first restriction, which admits the iteration and
// writes it_a, read by every later kernel of the iteration.
struct MgState {
  double rho[2], alpha[2];
  double rr, rr0, tol2;
  long long it_a, it_c, maxit;
  int status;  // BStatus
  int pad;
};
// The flexible-CG operands: level 0 (its b is the residual r, its x the
// preconditioned residual z), the solution x, the directions p (two
// buffers), q = A p, per-block partials ([4][kMaxRedBlocks]), the state,
// the row kernels' block count, the smoothing weight and the coarse
// correction's scale.
struct MgFcg {
  MgLevel f;
  float* x;
  float* p[2];
  float* q;
  double* part;
  MgState* st;
  int nb;
  float omega, scale;
};
hipError_t launch_mg_fill_i32(int32_t* x, int64_t n, int32_t v, hipStream_t s);
// flags[the coarse cell of each row of F] = 1 (flags over C's grid, zeroed).
hipError_t launch_mg_mark(const MgLevel& F, const MgLevel& C, int32_t* flags, hipStream_t s);
// Blocks of the flag scan over ncell cells (counts needs that many + 1).
int mg_scan_blocks(int64_t ncell);
// cell_or_null == null: counts = exclusive block offsets, counts[nb] = total;
// else: grid (the flags) becomes the row of each flagged cell (-1 elsewhere)
// and cell[row] its cell.
hipError_t launch_mg_scan(int32_t* grid, int64_t ncell, int32_t* counts, int32_t* cell_or_null, hipStream_t s);
hipError_t launch_mg_agg(const MgLevel& F, const MgLevel& C, int32_t* agg, hipStream_t s);
// Packed distances of C's rows; *bad = 1 if one does not fit.
hipError_t launch_mg_pk(const MgLevel& C, uint64_t* pk, int* bad, hipStream_t s);
// C's coefficients ([4][C.n]) and D^-1 as the Galerkin product of F.
hipError_t launch_mg_galerkin(const MgLevel& F, const MgLevel& C, float* coef, float* invd, hipStream_t s);
// r = rhs, x = 0, state for tolerance tol (relative to |rhs|) and the cap.
hipError_t launch_mg_init(const MgFcg& G, const float* rhs, double tol, long long maxit, hipStream_t s);
// One flexible-CG iteration over levels lv[0..nlev) (nlev >= 2).  ev (may
// be null): {start, stop} events for each of the finest level's four
// kernels (first restriction, last prolongation, direction, update), taken
// from their own dispatches.
hipError_t launch_mg_iteration(const MgFcg& G, const MgLevel* lv, int nlev, int coarse_iters, hipStream_t s,
                               const hipEvent_t* ev = nullptr);

// ---- fields, corrections, diagnostics (kernels_field.hip) -------------------
hipError_t launch_fill(double* x, int64_t n, double v, hipStream_t s);
// dst = src, `words` 16-byte words.
hipError_t launch_copy16(const void* src, void* dst, int64_t words, hipStream_t s);
// *bad = 1.0 if an entry of idx[count] lies outside [0, n) (bad untouched
// otherwise: the caller clears it).
hipError_t launch_check_index(const int32_t* idx, int64_t count, int n, double* bad, hipStream_t s);
hipError_t launch_init_pressure_linear(const Geo& g, const int32_t* pcell, int np, double* p, double p0,
                                       double p1, hipStream_t s);
hipError_t launch_clamp(double* x, const double* temp, int64_t n, double dvmax, hipStream_t s);
hipError_t launch_correct_velocity(const Geo& g, SysView P, SysView U, SysView V, SysView W, int kind,
                                   const double* q, double* tu, double* tv, double* tw, const double* uh,
                                   const double* vh, const double* wh, double dvmax, double p_0, double p_1,
                                   double dx, double dy, double dz, hipStream_t s);
hipError_t launch_correct_pressure(double* p, const double* pc, int64_t n, double beta, double* partials,
                                   double* result, hipStream_t s);
hipError_t launch_compute_div(const Geo& g, SysView P, const double* u, const double* v, const double* w,
                              double dx, double dy, double dz, double* partials, double* result, hipStream_t s);
hipError_t launch_permeability(const Geo& g, const double* u, const double* p, int border, double* partials,
                               double* result, hipStream_t s);

}  // namespace ns3d
