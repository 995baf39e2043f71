// capi.hip -- the C ABI of libns3d_gpu.so (include/ns3d.h): context
// management, device memory, and the host side of each NavierStokesBase
// method (the loop control of the linear solves, scalar post-processing).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ns3d.h"
#include "engine.h"
#include "kernels.h"

using namespace ns3d;

namespace {

thread_local std::string g_err;

struct HipFail {
  hipError_t e;
  const char* what;
};

#define HCHK(expr)                                       \
  do {                                                   \
    hipError_t e_ = (expr);                              \
    if (e_ != hipSuccess) throw HipFail{e_, #expr};      \
  } while (0)

struct Invalid {
  std::string msg;
};

struct SysMem {
  int n = 0;
  int d0 = 0;
  int32_t* grid = nullptr;   // owned by the context (index grids)
  int32_t* cell = nullptr;
  int32_t* nbr = nullptr;
  double* coef = nullptr;
  double* invdiag = nullptr;
  double* rowsum = nullptr;
  int64_t last_iters = 16;   // chunk-size hint for the next solve (plain)
  int64_t last_iters32 = 16; // same for the FP32 inner solves (mixed)
  // mixed form: FP32 sweeps on this system once failed to reduce the
  // residual by even the inner tolerance before an FP64 finish was needed;
  // its later solves are plain FP64
  bool fp32_stalls = false;
  int64_t nnz = 0;           // stored entries (diagonal + existing neighbours)
  float* coef32 = nullptr;   // mixed form: FP32 copies
  float* invdiag32 = nullptr;
  uint64_t* pk = nullptr;    // packed neighbour distances (layout != ML_INDEXED)
  int layout = ML_INDEXED;   // how the solver reads the matrix (MatLayout)
  // plane-band chunk list of a large system (build_chunks; null: contiguous
  // tiles) and the host copy of its band offsets, for the engine's tiles
  int* chk = nullptr;
  std::vector<int> band_off;
  const int* band_off_host() const { return band_off.empty() ? nullptr : band_off.data(); }
  SysView view() const { return SysView{n, d0, grid, cell, nbr, coef, invdiag, rowsum}; }
  SMat<double> mat64() const { return SMat<double>{n, layout, nbr, pk, coef, invdiag, chk}; }
  SMat<float> mat32() const { return SMat<float>{n, layout, nbr, pk, coef32, invdiag32, chk}; }
};

// Device time of the fused BiCGSTAB kernels, per storage precision (0 FP64,
// 1 FP32), system kind (0 momentum, 1 pressure) and kernel (A, B, C), over
// the sampled iterations (one per queued chunk; start and stop are the
// dispatch's own timestamps): milliseconds, sampled launches, algorithmic
// bytes.
struct Timing {
  bool on = false;
  double ms[2][2][3] = {}, launches[2][2][3] = {}, bytes[2][2][3] = {};
};

// Workspace of one linear solve in flight.
struct Lane {
  double* vb = nullptr;     // momentum right-hand side
  double* temp = nullptr;   // saved field for revert/clamp
  BWork bw{};
  SWork<float> bw32{};      // mixed form: FP32 inner-solve vectors (state shared with bw)
  float* d32 = nullptr;     // inner-solve unknown (the correction)
  float* r32 = nullptr;     // inner-solve right-hand side
  double* rres = nullptr;   // FP64 residual of the defect correction
  double* partials = nullptr;  // [kPartialSlots][kMaxTilesPerSolve]
  double* result = nullptr;    // reduction results (device)
  BState* hstate = nullptr;    // pinned copy of the solver state
  double* hresult = nullptr;   // pinned copy of results
};

// The multigrid hierarchy of the pressure system (mixed form), built on the
// first inner solve that uses it.  Levels hold device pointers into the
// context's allocations; coef / invd are the writable views of levels >= 1.
// Default of ns3d_set_multigrid: pressure systems from this many rows on.
constexpr int64_t kMgDefaultMinRows = 200000;

struct MgHier {
  bool tried = false;  // build attempted (ok: it succeeded)
  bool ok = false;
  bool dirty = true;   // coarse operators stale (the pressure system was reassembled)
  int nlev = 0;
  MgLevel lv[kMgMaxLevels]{};
  float* coef[kMgMaxLevels] = {};
  float* invd[kMgMaxLevels] = {};
  MgFcg G{};
  MgState* hst = nullptr;  // pinned copy of the state
  int64_t hint = 8;        // iterations of the last solve (sizes the first chunk)
  int coarse_iters = 16;   // CG steps of the coarsest solve
  // chunks of iterations are replayed as graphs (one per chunk length and
  // solution vector), captured on a stream of their own; NS3D_MG_GRAPHS=0
  // launches them kernel by kernel
  bool graphs = true;
  // sampled kernel times (with the context's timing on): events of the
  // finest level's four kernels, and per kernel milliseconds, launches and
  // algorithmic bytes
  hipEvent_t tev[8] = {};
  double ms[4] = {}, launches[4] = {}, bytes[4] = {};
  hipStream_t cap = nullptr;
  std::vector<std::pair<std::pair<const void*, int>, hipGraphExec_t>> execs;
};

}  // namespace

struct ns3d_ctx {
  int device = 0;
  hipStream_t stream = nullptr;      // where the context's work goes: its own stream, or its engine's
  hipStream_t own_stream = nullptr;  // the stream it created
  hipEvent_t ev_sync = nullptr;      // marks the end of its own work so far (ctx_sync)
  int nx = 0, ny = 0, nz = 0;
  int64_t ncell = 0;
  uint8_t* img = nullptr;
  int32_t *pg = nullptr, *ug = nullptr, *vg = nullptr, *wg = nullptr;
  SysMem sys[4];  // u, v, w, p
  // fields
  double* vel[3][3] = {{nullptr}};  // [set][component]
  double* p = nullptr;
  double* pc = nullptr;
  double* b = nullptr;
  // solver workspaces: lane 0 serves every solve on the context's stream and
  // the reductions; with an engine attached, lanes 1 and 2 let the three
  // momentum solves of one ComputeVelocityField run at once
  Lane ln[3];
  int nlanes = 0;
  int64_t lane_rows = 0;
  int precision = NS3D_PRECISION_FP64;
  double inner_tol = 1e-4;
  int64_t mixed_min_rows = 0;  // mixed form: smaller systems are solved in FP64
  bool engine_used = false;     // the last solves went through the attached engine
  ns3d_params prm{};
  bool have_params = false;
  std::vector<void*> allocs;
  double alloc_bytes = 0;   // device bytes requested by this context
  int64_t fp64_finishes = 0;  // mixed form: solves finished in FP64 after FP32 stagnated
  Timing tm;
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::vector<hipEvent_t> tev;  // start / stop of A, B, C of the sampled iteration
  ns3d_engine* engine = nullptr;  // batched solve engine (null: solve on the context's stream)
  // mixed form: pressure systems with at least this many rows are solved by
  // multigrid-preconditioned flexible CG (0: never)
  int64_t mg_min_rows = kMgDefaultMinRows;
  MgHier mg;
  // BiCGSTAB restart threshold factor (0: Eigen's epsilon^2).  Set from
  // NS3D_RESTART_EPS2 at creation: tests raise it to drive solves through
  // the restart path, which Eigen's threshold reaches only rarely
  double restart_eps2 = 0.0;

  Geo geo() const { return Geo{nx, ny, nz, pg, ug, vg, wg}; }

  template <typename T>
  T* alloc(int64_t count) {
    void* ptr = nullptr;
    const size_t bytes = (size_t)(count > 0 ? count : 1) * sizeof(T);
    hipError_t e = hipMalloc(&ptr, bytes);
    if (e != hipSuccess) throw HipFail{e, "hipMalloc"};
    allocs.push_back(ptr);
    alloc_bytes += (double)bytes;
    return (T*)ptr;
  }
};

namespace {

// Waits for the work this context has queued so far.  An attached context's
// stream is shared with the engine's other contexts, so it waits on an event
// recorded after its own work, not for whatever the others queued since.
hipError_t ctx_sync(ns3d_ctx* c) {
  hipError_t e = hipEventRecord(c->ev_sync, c->stream);
  if (e != hipSuccess) return e;
  return hipEventSynchronize(c->ev_sync);
}

ns3d_status fail_hip(const HipFail& f) {
  char buf[512];
  snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s", (int)f.e, hipGetErrorString(f.e), f.what);
  g_err = buf;
  return f.e == hipErrorOutOfMemory ? NS3D_ERR_OOM : NS3D_ERR_HIP;
}

ns3d_status fail_invalid(const std::string& m) {
  g_err = m;
  return NS3D_ERR_INVALID;
}

template <typename F>
ns3d_status guarded(F&& f) {
  try {
    f();
    return NS3D_OK;
  } catch (const HipFail& e) {
    return fail_hip(e);
  } catch (const Invalid& e) {
    return fail_invalid(e.msg);
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return NS3D_ERR_OOM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return NS3D_ERR_HIP;
  }
}

void need(bool c, const char* msg) {
  if (!c) throw Invalid{msg};
}

Phys phys_of(const ns3d_ctx* c, double alpha, double beta) {
  Phys ph;
  ph.dx = c->prm.dx; ph.dy = c->prm.dy; ph.dz = c->prm.dz;
  ph.mu = c->prm.mu; ph.rho = c->prm.rho;
  ph.p0 = c->prm.p0; ph.p1 = c->prm.p1;
  ph.alpha = alpha; ph.beta = beta;
  ph.sf = 0.0;  // schemefactor (NSBase.cxx:171)
  return ph;
}

double* vec_ptr(ns3d_ctx* c, int which, int64_t& n) {
  switch (which) {
    case NS3D_VEC_U: case NS3D_VEC_UT: case NS3D_VEC_UH: n = c->sys[0].n; break;
    case NS3D_VEC_V: case NS3D_VEC_VT: case NS3D_VEC_VH: n = c->sys[1].n; break;
    case NS3D_VEC_W: case NS3D_VEC_WT: case NS3D_VEC_WH: n = c->sys[2].n; break;
    case NS3D_VEC_P: case NS3D_VEC_PC: case NS3D_VEC_B: n = c->sys[3].n; break;
    default: throw Invalid{"unknown vector selector"};
  }
  if (which <= NS3D_VEC_WH) return c->vel[which / 3][which % 3];
  if (which == NS3D_VEC_P) return c->p;
  if (which == NS3D_VEC_PC) return c->pc;
  return c->b;
}

Vel3 vel3(ns3d_ctx* c, int set) {
  if (set < 0 || set > 2) throw Invalid{"unknown velocity set"};
  return Vel3{c->vel[set][0], c->vel[set][1], c->vel[set][2]};
}

// Algorithmic bytes of the fused BiCGSTAB kernels per launch, for storage of
// vb bytes per value: the matrix is nnz*(vb+4) (coefficient + 32-bit column
// per stored entry), plus vb bytes per row per vector stream -- A: r, p, v,
// M^-1, r0 read, p, v written (7); B: r, v, M^-1 read, t written (4); C: x, p,
// r, v, t, M^-1, r0 read, x, r written (9).
double kernel_bytes(int64_t n, int64_t nnz, int k, int vb) {
  static const double streams[3] = {7.0, 4.0, 9.0};
  const double mat = (k < 2) ? (double)nnz * (vb + 4.0) : 0.0;
  return mat + streams[k] * vb * (double)n;
}

// One BiCGSTAB call: system, matrix, unknown (with its guess), right-hand
// side, vectors, tolerance, cap, the lane whose partials and pinned state it
// uses, and the chunk-size hint of its system.
template <typename T>
struct SolveJob {
  int sid = 0;
  SMat<T> M{};
  T* x = nullptr;
  const T* b = nullptr;
  SWork<T> w{};
  double tol = 0;
  int64_t maxit = 0;
  Lane* L = nullptr;
  int64_t* hint = nullptr;
};

// Host loop of one queued BiCGSTAB solve on stream s: init, then chunks of
// iterations with status polling between them.  The solver state is copied
// into *hs after each chunk and sync() waits for it.  hint: iterations of the
// previous solve of this system, sizes the first chunk; updated on return.
// tev (may be null): six events that time kernels A, B, C of one iteration
// in the middle of each chunk, after which sampled() reads them if that
// iteration did work.
template <typename T, typename Sync, typename Sampled>
void queued_solve(hipStream_t s, const SolveDesc<T>& d, BState* hs, int64_t& hint, hipEvent_t* tev, Sync&& sync,
                  Sampled&& sampled) {
  HCHK(launch_bicg_init<T>(single_batch(d, 0, 0), s));
  HCHK(hipMemcpyAsync(hs, d.w.st, sizeof(BState), hipMemcpyDeviceToHost, s));
  sync();
  int64_t chunk = hint > 4 ? (hint * 3) / 4 : 4;
  int64_t launched = 0;
  const int64_t maxit = d.maxit;
  const int64_t cap = 4 * (maxit > 0 ? maxit : 1) + 64;
  while (hs->status == BS_RUNNING || hs->status == BS_RESTART) {
    const long long done_before = hs->it_c;
    if (hs->status == BS_RESTART) HCHK(launch_bicg_restart<T>(single_batch(d, done_before, 0), s));
    if (chunk > maxit + 1) chunk = maxit + 1;
    const int64_t qt = tev ? chunk / 2 : -1;
    for (int64_t q = 0; q < chunk; ++q) {
      hipEvent_t* e = (q == qt) ? tev : nullptr;
      const Batch<T> B = single_batch(d, done_before, (int)q);
      HCHK(launch_bicg_A<T>(B, s, e));
      HCHK(launch_bicg_B<T>(B, s, e ? e + 2 : nullptr));
      HCHK(launch_bicg_C<T>(B, s, e ? e + 4 : nullptr));
    }
    HCHK(hipMemcpyAsync(hs, d.w.st, sizeof(BState), hipMemcpyDeviceToHost, s));
    sync();
    if (qt >= 0 && hs->it_c - done_before > qt) sampled();
    launched += chunk;
    if (launched > cap) throw Invalid{"BiCGSTAB did not stop within its iteration cap"};
    chunk = 16;
  }
  if (hs->status != BS_ZERO_RHS) hint = hs->it_c > 0 ? hs->it_c : 4;
}

// One solve queued on the context's stream.  Leaves the final solver state
// in the lane's hstate; with timing on, one iteration per chunk is timed.
template <typename T>
void stream_solve(ns3d_ctx* c, const SolveJob<T>& j) {
  const SysMem& S = c->sys[j.sid];
  const int kind = j.sid == 3 ? 1 : 0;
  const int prec = sizeof(T) == 4 ? 1 : 0;
  // one solve: its descriptor travels by value in the kernel arguments
  const SolveDesc<T> d{j.M, j.w, j.x, j.b, j.tol, (long long)j.maxit, j.L->partials, kMaxTilesPerSolve, 0, 0, 0,
                       c->restart_eps2};
  queued_solve<T>(c->stream, d, j.L->hstate, *j.hint, c->tm.on ? c->tev.data() : nullptr,
                  [&] { HCHK(ctx_sync(c)); },
                  [&] {
                    for (int k = 0; k < 3; ++k) {
                      float ms = 0.f;
                      HCHK(hipEventElapsedTime(&ms, c->tev[2 * k], c->tev[2 * k + 1]));
                      c->tm.ms[prec][kind][k] += ms;
                      c->tm.launches[prec][kind][k] += 1;
                      c->tm.bytes[prec][kind][k] += kernel_bytes(S.n, S.nnz, k, (int)sizeof(T));
                    }
                  });
}

// Runs the solves, each leaving its final state in its lane's hstate: with an
// engine that other contexts share, all at once (batched with the other
// contexts' solves), else one after the other on the context's stream.  Each
// solve's arithmetic is the same either way, so a context may switch between
// the two from one call to the next.
template <typename T>
void run_solves(ns3d_ctx* c, SolveJob<T>* jobs, int n) {
  if (n <= 0) return;
  if (!c->engine || c->engine->attached.load() <= 1) {
    // rounds the engine queued before this context's last batched solve
    // returned may still name its lanes' state: let them drain first
    if (c->engine && c->engine_used) c->engine->fence();
    c->engine_used = false;
    for (int i = 0; i < n; ++i) stream_solve<T>(c, jobs[i]);
    return;
  }
  c->engine_used = true;
  if (n > 3) throw Invalid{"run_solves: at most three solves at once"};
  EngineRequest req[3];
  EngineRequest* rp[3];
  HCHK(hipEventRecord(c->ev[0], c->stream));
  for (int i = 0; i < n; ++i) {
    const SolveJob<T>& j = jobs[i];
    const SolveDesc<T> d{j.M, j.w, j.x, j.b, j.tol, (long long)j.maxit, j.L->partials, kMaxTilesPerSolve, 0, 0, 0,
                         c->restart_eps2};
    req[i].prec = sizeof(T) == 4 ? 1 : 0;
    if constexpr (sizeof(T) == 4) req[i].d32 = d;
    else req[i].d64 = d;
    req[i].n = c->sys[j.sid].n;
    req[i].nnz = c->sys[j.sid].nnz;
    req[i].kind = j.sid == 3 ? 1 : 0;
    req[i].band_off = c->sys[j.sid].band_off_host();
    req[i].ready = c->ev[0];
    rp[i] = &req[i];
  }
  c->engine->solve_many(rp, n);
  for (int i = 0; i < n; ++i) {
    *jobs[i].L->hstate = req[i].result;
    const BState& st = req[i].result;
    if (st.status != BS_ZERO_RHS) *jobs[i].hint = st.it_c > 0 ? st.it_c : 4;
  }
}

// The result of a plain FP64 solve from its final state, as Eigen reports
// it.  Returns true for a zero right-hand side, whose solution is x = 0 (the
// caller writes it) with iterations = maxit and error = tol.
bool solve_info(const BState& st, double tol, int64_t maxit, ns3d_solve_info& info) {
  info.restarts = 0;
  info.work = 0;
  if (st.status == BS_ZERO_RHS) {
    info.iterations = maxit;
    info.error = tol;
    info.info = 0;
    return true;
  }
  info.iterations = st.it_c - st.ibase;
  info.work = st.it_c;
  info.error = std::sqrt(st.rr / st.rhs_sqnorm);
  info.info = (info.error <= tol) ? 0 : 2;
  info.restarts = (int32_t)st.restarts;
  return false;
}

// n BiCGSTAB solves (system sid[i], unknown x[i] with its guess, right-hand
// side b[i], lane L[i]; at most three, on distinct systems and lanes), Eigen
// semantics, FP64.
void bicgstab64_many(ns3d_ctx* c, int n, const int* sid, double* const* x, const double* const* b, double tol,
                     int64_t maxit, ns3d_solve_info* info, Lane* const* L) {
  SolveJob<double> jobs[3];
  int idx[3];
  int m = 0;
  for (int i = 0; i < n; ++i) {
    SysMem& S = c->sys[sid[i]];
    if (S.n == 0) {
      info[i].iterations = 0; info[i].info = 0; info[i].restarts = 0; info[i].error = 0; info[i].work = 0;
      continue;
    }
    jobs[m] = SolveJob<double>{sid[i], S.mat64(), x[i], b[i], L[i]->bw, tol, maxit, L[i], &S.last_iters};
    idx[m++] = i;
  }
  run_solves<double>(c, jobs, m);
  for (int q = 0; q < m; ++q) {
    const int i = idx[q];
    if (solve_info(*L[i]->hstate, tol, maxit, info[i])) HCHK(launch_fill(x[i], c->sys[sid[i]].n, 0.0, c->stream));
  }
}

// ---- multigrid inner solve of the pressure system (mixed form) -------------

float env_float(const char* name, float dflt) {
  const char* e = std::getenv(name);
  return e ? (float)std::atof(e) : dflt;
}

// Builds the aggregation hierarchy of the pressure system (kernels_mg.hip):
// level l+1 holds the 2x2x2 blocks of level l's grid with at least one row,
// down to at most kMgCoarseRows rows.  Leaves c->mg.ok false (the inner
// solves stay Jacobi BiCGSTAB) when the system is too small, is not stored
// packed-symmetric, or a level's neighbour distances do not fit the packed
// format.
void mg_build(ns3d_ctx* c) {
  MgHier& H = c->mg;
  H.tried = true;
  const SysMem& S = c->sys[3];
  if (S.layout != ML_PACKED_SYM || !S.coef32 || S.n <= kMgCoarseRows) return;
  hipStream_t s = c->stream;
  MgLevel lv[kMgMaxLevels]{};
  float* coef[kMgMaxLevels] = {};
  float* invd[kMgMaxLevels] = {};
  MgLevel& L0 = lv[0];
  L0.n = S.n; L0.nx = c->nx; L0.ny = c->ny; L0.nz = c->nz;
  L0.sE = 1 + DE; L0.sN = 1 + DN; L0.sT = 1 + DT;
  L0.grid = c->pg; L0.cell = S.cell; L0.pk = S.pk; L0.coef = S.coef32; L0.invd = S.invdiag32;
  L0.b = c->alloc<float>(S.n);
  L0.x = c->alloc<float>(S.n);
  int nlev = 1;
  int* bad = c->alloc<int>(1);
  HCHK(hipMemsetAsync(bad, 0, sizeof(int), s));
  while (lv[nlev - 1].n > kMgCoarseRows) {
    if (nlev == kMgMaxLevels) return;
    MgLevel& F = lv[nlev - 1];
    MgLevel C{};
    C.nx = (F.nx + 1) / 2; C.ny = (F.ny + 1) / 2; C.nz = (F.nz + 1) / 2;
    if (C.nx == F.nx && C.ny == F.ny && C.nz == F.nz) return;
    const int64_t ncell = (int64_t)C.nx * C.ny * C.nz;
    int32_t* grid = c->alloc<int32_t>(ncell);
    HCHK(launch_mg_fill_i32(grid, ncell, 0, s));
    HCHK(launch_mg_mark(F, C, grid, s));
    const int nb = mg_scan_blocks(ncell);
    int32_t* counts = c->alloc<int32_t>(nb + 1);
    HCHK(launch_mg_scan(grid, ncell, counts, nullptr, s));
    int32_t total = 0;
    HCHK(hipMemcpyAsync(&total, counts + nb, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HCHK(ctx_sync(c));
    C.n = total;
    C.grid = grid;
    int32_t* cell = c->alloc<int32_t>(C.n);
    C.cell = cell;
    HCHK(launch_mg_scan(grid, ncell, counts, cell, s));
    int32_t* agg = c->alloc<int32_t>(F.n);
    HCHK(launch_mg_agg(F, C, agg, s));
    F.agg = agg;
    uint64_t* pk = c->alloc<uint64_t>(C.n);
    C.pk = pk;
    HCHK(launch_mg_pk(C, pk, bad, s));
    C.sE = 1; C.sN = 2; C.sT = 3;
    coef[nlev] = c->alloc<float>((int64_t)4 * C.n);
    invd[nlev] = c->alloc<float>(C.n);
    C.coef = coef[nlev];
    C.invd = invd[nlev];
    C.b = c->alloc<float>(C.n);
    C.x = c->alloc<float>(C.n);
    lv[nlev++] = C;
  }
  int hbad = 0;
  HCHK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  HCHK(ctx_sync(c));
  if (hbad || nlev < 2) return;
  MgFcg& G = H.G;
  G.f = lv[0];
  G.p[0] = c->alloc<float>(S.n);
  G.p[1] = c->alloc<float>(S.n);
  G.q = c->alloc<float>(S.n);
  G.part = c->alloc<double>((int64_t)4 * kMaxRedBlocks);
  G.st = c->alloc<MgState>(1);
  G.nb = reduce_blocks(S.n);
  // Jacobi damping 6/7: omega * lambda_max(D^-1 A) <= 12/7 < 2 for every
  // diagonally dominant 7-point row, so the smoother stays convergent (1.0
  // is not: it loses the high-frequency damping and the CG stalls).  The
  // coarse correction is over-scaled by 1.4, the usual remedy for the
  // piecewise-constant prolongation's too-small coarse-grid error.  Swept on
  // the 128^3 RSA and 384^3 Boolean pressure systems: 27% and 38% fewer CG
  // iterations than (2/3, 1.0).
  G.omega = env_float("NS3D_MG_OMEGA", 6.0f / 7.0f);
  G.scale = env_float("NS3D_MG_SCALE", 1.4f);
  H.coarse_iters = (int)env_float("NS3D_MG_COARSE_ITERS", (float)H.coarse_iters);
  H.graphs = env_float("NS3D_MG_GRAPHS", 1.0f) != 0.0f;
  HCHK(hipHostMalloc((void**)&H.hst, sizeof(MgState), hipHostMallocDefault));
  for (int l = 0; l < nlev; ++l) {
    H.lv[l] = lv[l];
    H.coef[l] = coef[l];
    H.invd[l] = invd[l];
  }
  H.nlev = nlev;
  H.ok = true;
  H.dirty = true;
}

// Whether system sid's FP32 sweeps are multigrid-preconditioned CG (the
// pressure system of a mixed-form context, from mg_min_rows rows on).
bool mg_on(ns3d_ctx* c, int sid) {
  if (sid != 3 || c->precision != NS3D_PRECISION_MIXED || c->mg_min_rows <= 0) return false;
  const SysMem& S = c->sys[3];
  if (S.n < c->mg_min_rows || !S.coef32) return false;
  if (!c->mg.tried) mg_build(c);
  return c->mg.ok;
}

// The graph of `chunk` iterations of the current operands (built on first
// use).  Its kernels take everything from device memory, so one graph
// serves every solve with the same solution vector.
hipGraphExec_t mg_graph(ns3d_ctx* c, int chunk) {
  MgHier& H = c->mg;
  const std::pair<const void*, int> key{H.G.x, chunk};
  for (auto& e : H.execs)
    if (e.first == key) return e.second;
  if (!H.cap) HCHK(hipStreamCreateWithFlags(&H.cap, hipStreamNonBlocking));
  HCHK(hipStreamBeginCapture(H.cap, hipStreamCaptureModeThreadLocal));
  hipError_t err = hipSuccess;
  for (int q = 0; q < chunk && err == hipSuccess; ++q) err = launch_mg_iteration(H.G, H.lv, H.nlev, H.coarse_iters, H.cap);
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(H.cap, &g);
  HCHK(err);
  HCHK(e2);
  hipGraphExec_t ex = nullptr;
  const hipError_t e3 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HCHK(e3);
  H.execs.push_back({key, ex});
  return ex;
}

// One FP32 sweep A32 x = rhs (x from 0) to a relative tolerance tol, at
// most maxit iterations, by flexible CG with one V-cycle per iteration;
// the final state is returned in BiCGSTAB's terms (status DONE or
// ZERO_RHS, it_c iterations, rr = |r|^2, rhs_sqnorm = |rhs|^2).
void mg_solve(ns3d_ctx* c, const float* rhs, float* x, double tol, int64_t maxit, BState& out) {
  MgHier& H = c->mg;
  // on the context's stream (an attached context's is the engine's context
  // stream, which has a hardware queue of its own: a stream sharing a
  // lane's queue would wait behind that lane's chunks)
  hipStream_t s = c->stream;
  auto sync = [&] { HCHK(ctx_sync(c)); };
  if (H.dirty) {
    for (int l = 0; l + 1 < H.nlev; ++l) HCHK(launch_mg_galerkin(H.lv[l], H.lv[l + 1], H.coef[l + 1], H.invd[l + 1], s));
    H.dirty = false;
  }
  H.G.x = x;
  HCHK(launch_mg_init(H.G, rhs, tol, (long long)maxit, s));
  HCHK(hipMemcpyAsync(H.hst, H.G.st, sizeof(MgState), hipMemcpyDeviceToHost, s));
  sync();
  // chunks of 2, 4, 8 or 16 iterations: the first sized by the last solve,
  // then 4 between polls
  int chunk = 2;
  while (chunk < 16 && 2 * chunk <= H.hint * 3 / 4) chunk *= 2;
  int64_t launched = 0;
  // with timing on, the first chunk goes kernel by kernel and its first
  // iteration's finest-level kernels carry dispatch events
  bool sample = c->tm.on;
  if (sample && !H.tev[0])
    for (auto& e : H.tev) HCHK(hipEventCreate(&e));
  while (H.hst->status == BS_RUNNING) {
    const bool direct = !H.graphs || sample;
    if (!direct) {
      HCHK(hipGraphLaunch(mg_graph(c, chunk), s));
    } else {
      for (int q = 0; q < chunk; ++q)
        HCHK(launch_mg_iteration(H.G, H.lv, H.nlev, H.coarse_iters, s, (sample && q == 0) ? H.tev : nullptr));
    }
    HCHK(hipMemcpyAsync(H.hst, H.G.st, sizeof(MgState), hipMemcpyDeviceToHost, s));
    sync();
    if (sample && H.hst->it_c > 0) {
      // the sampled iteration ran: its kernels' times and algorithmic bytes
      const MgLevel& F = H.lv[0];
      const double n = F.n, nnz = (double)c->sys[3].nnz, nc = H.lv[1].n;
      const double b[4] = {nnz * 8 + 12 * n, nnz * 8 + 20 * n + 4 * nc, nnz * 8 + 16 * n, 24 * n};
      for (int k = 0; k < 4; ++k) {
        float ms = 0.f;
        HCHK(hipEventElapsedTime(&ms, H.tev[2 * k], H.tev[2 * k + 1]));
        H.ms[k] += ms;
        H.launches[k] += 1;
        H.bytes[k] += b[k];
      }
    }
    sample = false;
    launched += chunk;
    if (launched > 2 * maxit + 64) throw Invalid{"multigrid CG did not stop within its iteration cap"};
    chunk = 4;
  }
  const MgState& st = *H.hst;
  if (std::getenv("NS3D_MG_TRACE"))
    std::fprintf(stderr, "mg_solve: n %d levels %d tol %.3e status %d its %lld |r0| %.6e |r| %.6e rho %.6e %.6e alpha %.6e %.6e\n",
                 H.lv[0].n, H.nlev, tol, st.status, st.it_c, std::sqrt(st.rr0), std::sqrt(st.rr), st.rho[0], st.rho[1],
                 st.alpha[0], st.alpha[1]);
  if (st.status != BS_ZERO_RHS) H.hint = std::max<int64_t>(2, st.it_c);
  out = BState{};
  out.status = st.status == BS_ZERO_RHS ? BS_ZERO_RHS : BS_DONE;
  out.it_c = st.it_c;
  out.rr = st.rr;
  out.rhs_sqnorm = st.rr0;
}

// Mixed form: FP64 defect correction around FP32 inner BiCGSTAB solves.
//   r = b - A x (FP64); while |r| > tol |b|: solve A32 d = r32 to a relative
//   tolerance max(inner_tol, tol |b| / (2 |r|)); x += d; r = b - A x.
// A sweep that does nothing, or after which |r| has not halved, means FP32
// has reached its accuracy on this system: the solve is finished in FP64
// from the current iterate.  A sweep that breaks down or makes |r| larger is
// not applied (it is taken back), and the FP32 sweeps together use at most
// half of maxit, so the FP64 finish always has the rest: on a large,
// ill-conditioned system FP32 BiCGSTAB's true residual stalls far above its
// recursive one.  When the FP32 sweeps of a solve used up their half of the
// budget or did not reduce |r| by even the inner tolerance before the FP64
// finish, or the finish took more than half as many iterations as those
// sweeps, FP32 does not carry that system: its later solves are plain FP64
// (mixed_ready).  iterations = summed inner iterations; error =
// |b - A x| / |b| (true FP64 residual); a zero right-hand side and an
// already-converged guess behave as in the plain solve (x = 0 with
// iterations = maxit, and 0 iterations).  The n solves advance in lockstep
// (their sweeps and finishes run together), each with exactly the steps it
// would take alone.
void bicgstab_mixed_many(ns3d_ctx* c, int n, const int* sid, double* const* x, const double* const* b, double tol,
                         int64_t maxit, ns3d_solve_info* info, Lane* const* L) {
  hipStream_t s = c->stream;
  struct Track {
    double rr = 0, bb = 0, tol2 = 0, rr_old = 0, err = 0, rr0 = 0;
    int64_t total = 0, total32 = 0;
    int outer = 0;
    bool live = false, stagnated = false;
  };
  Track t[3];
  bool use[3] = {false, false, false};
  // the FP32 sweeps get at most half the iteration budget, so that an FP64
  // finish always has the other half
  const int64_t budget32 = std::max<int64_t>(1, maxit / 2);
  // |b - A x|^2 (trial: of the trial iterate x + d, x unchanged) and |b|^2
  auto residuals = [&](const bool* which, bool trial) {
    bool any = false;
    for (int i = 0; i < n; ++i) {
      if (!which[i]) continue;
      HCHK(launch_residual64(c->sys[sid[i]].mat64(), x[i], b[i], L[i]->rres, L[i]->r32, L[i]->partials,
                             L[i]->result, s, trial ? L[i]->d32 : nullptr));
      HCHK(hipMemcpyAsync(L[i]->hresult, L[i]->result, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
      any = true;
    }
    if (!any) return;
    HCHK(ctx_sync(c));
    for (int i = 0; i < n; ++i)
      if (which[i]) {
        t[i].rr = L[i]->hresult[0];
        t[i].bb = L[i]->hresult[1];
      }
  };
  for (int i = 0; i < n; ++i) {
    info[i].restarts = 0;
    info[i].work = 0;
    use[i] = c->sys[sid[i]].n > 0;
    if (!use[i]) { info[i].iterations = 0; info[i].info = 0; info[i].error = 0; }
  }
  residuals(use, false);
  for (int i = 0; i < n; ++i) {
    if (!use[i]) continue;
    if (t[i].bb == 0) {
      HCHK(launch_fill(x[i], c->sys[sid[i]].n, 0.0, s));
      info[i].iterations = maxit; info[i].error = tol; info[i].info = 0;
      use[i] = false;
      continue;
    }
    t[i].tol2 = tol * tol * t[i].bb;
    t[i].rr0 = t[i].rr;
    t[i].live = true;
  }
  for (;;) {
    SolveJob<float> jobs[3];
    int idx[3], all[3];
    int m = 0, na = 0;
    for (int i = 0; i < n; ++i) {
      Track& k = t[i];
      if (!use[i] || !k.live) continue;
      if (!(k.rr > k.tol2 && k.total < budget32 && k.outer < 100)) {
        // out of FP32 budget or sweeps short of the tolerance: FP64 finishes
        if (k.rr > k.tol2) k.stagnated = true;
        k.live = false;
        continue;
      }
      SysMem& S = c->sys[sid[i]];
      const double rel = std::sqrt(k.rr / k.bb);
      const double tin = std::max(c->inner_tol, 0.5 * tol / rel);
      all[na++] = i;
      if (mg_on(c, sid[i])) {
        // the pressure system's sweep: multigrid-preconditioned CG, on the
        // context's stream, its final state in the lane's pinned copy
        mg_solve(c, L[i]->r32, L[i]->d32, tin, budget32 - k.total, *L[i]->hstate);
        continue;
      }
      HCHK(launch_fill_f32(L[i]->d32, S.n, s));
      jobs[m] = SolveJob<float>{sid[i], S.mat32(), L[i]->d32, L[i]->r32, L[i]->bw32, tin, budget32 - k.total, L[i],
                                &S.last_iters32};
      idx[m++] = i;
    }
    if (na == 0) break;
    run_solves<float>(c, jobs, m);
    bool again[3] = {false, false, false};
    for (int q = 0; q < na; ++q) {
      const int i = all[q];
      Track& k = t[i];
      const BState& st = *L[i]->hstate;
      const int64_t its = (st.status == BS_ZERO_RHS) ? 0 : st.it_c;
      info[i].restarts += (int32_t)st.restarts;
      k.total += its;
      ++k.outer;
      // no iteration, or a sweep that broke down (non-finite residual):
      // its correction is not applied
      if (its == 0 || !std::isfinite(st.rr)) {
        k.stagnated = true;
        k.live = false;
        continue;
      }
      k.rr_old = k.rr;
      again[i] = true;
    }
    // the FP64 residual of x + d decides whether the correction is applied
    residuals(again, true);
    for (int i = 0; i < n; ++i) {
      if (!again[i]) continue;
      Track& k = t[i];
      if (!(k.rr <= k.rr_old)) {
        // the correction would make |r| larger, or is not finite (FP32 sweeps
        // past their accuracy, or overflowed, on an ill-conditioned system):
        // x stays as it was
        k.rr = k.rr_old;
        k.stagnated = true;
        k.live = false;
        continue;
      }
      HCHK(launch_add_f32(x[i], L[i]->d32, c->sys[sid[i]].n, s));
      if (k.rr > 0.25 * k.rr_old) {  // |r| fell by less than 2x
        k.stagnated = true;
        k.live = false;
      }
    }
  }
  // finish in FP64 from the current iterate (same tolerance, the remaining
  // iteration budget) where FP32 stagnated short of the tolerance
  SolveJob<double> fin[3];
  int idx[3];
  int m = 0;
  for (int i = 0; i < n; ++i) {
    if (!use[i]) continue;
    Track& k = t[i];
    k.err = std::sqrt(k.rr / k.bb);
    if (k.stagnated && k.rr > k.tol2 && k.total < maxit) {
      SysMem& S = c->sys[sid[i]];
      if (!(k.rr <= c->inner_tol * c->inner_tol * k.rr0) || k.total >= budget32) S.fp32_stalls = true;
      k.total32 = k.total;
      fin[m] = SolveJob<double>{sid[i], S.mat64(), x[i], b[i], L[i]->bw, tol, maxit - k.total, L[i], &S.last_iters};
      idx[m++] = i;
    }
  }
  run_solves<double>(c, fin, m);
  for (int q = 0; q < m; ++q) {
    const int i = idx[q];
    const BState& st = *L[i]->hstate;
    t[i].total += st.it_c;
    info[i].restarts += (int32_t)st.restarts;
    t[i].err = std::sqrt(st.rr / st.rhs_sqnorm);
    // an FP64 finish longer than half the FP32 sweeps before it: FP32 did
    // not carry this system
    if (2 * st.it_c > t[i].total32) c->sys[sid[i]].fp32_stalls = true;
    c->fp64_finishes += 1;
  }
  for (int i = 0; i < n; ++i) {
    if (!use[i]) continue;
    info[i].iterations = t[i].total;
    info[i].work = t[i].total;
    info[i].error = t[i].err;
    info[i].info = (info[i].error <= tol) ? 0 : 2;
  }
}

bool mixed_ready(const ns3d_ctx* c, const int* sid, int n) {
  if (c->precision != NS3D_PRECISION_MIXED) return false;
  for (int i = 0; i < n; ++i) {
    const SysMem& S = c->sys[sid[i]];
    if (S.n > 0 && (!S.coef32 || S.n < c->mixed_min_rows || S.fp32_stalls)) return false;
  }
  return true;
}

void bicgstab_many(ns3d_ctx* c, int n, const int* sid, double* const* x, const double* const* b, double tol,
                   int64_t maxit, ns3d_solve_info* info, Lane* const* L) {
  if (mixed_ready(c, sid, n)) bicgstab_mixed_many(c, n, sid, x, b, tol, maxit, info, L);
  else bicgstab64_many(c, n, sid, x, b, tol, maxit, info, L);
}

// One BiCGSTAB solve on system sid with x as guess (lane 0).
void bicgstab(ns3d_ctx* c, int sid, double* x, const double* b, double tol, int64_t maxit, ns3d_solve_info* info) {
  Lane* L = &c->ln[0];
  bicgstab_many(c, 1, &sid, &x, &b, tol, maxit, info, &L);
}

// A lane's FP64 workspace (vectors of lane_rows rows, partials, state,
// pinned copies), plus the FP32 one when the context is in the mixed form.
void alloc_lane32(ns3d_ctx* c, Lane& L) {
  if (L.d32) return;
  const int64_t nmax = std::max<int64_t>(c->lane_rows, 1);
  SWork<float>& w = L.bw32;
  w.r = c->alloc<float>(nmax);
  w.r0 = c->alloc<float>(nmax);
  w.p[0] = c->alloc<float>(nmax);
  w.p[1] = c->alloc<float>(nmax);
  w.v[0] = c->alloc<float>(nmax);
  w.v[1] = c->alloc<float>(nmax);
  w.t = c->alloc<float>(nmax);
  w.st = L.bw.st;
  L.d32 = c->alloc<float>(nmax);
  L.r32 = c->alloc<float>(nmax);
  L.rres = c->alloc<double>(nmax);
}

void alloc_lane(ns3d_ctx* c, Lane& L) {
  const int64_t nmax = std::max<int64_t>(c->lane_rows, 1);
  L.vb = c->alloc<double>(nmax);
  L.temp = c->alloc<double>(nmax);
  BWork& w = L.bw;
  w.r = c->alloc<double>(nmax);
  w.r0 = c->alloc<double>(nmax);
  w.p[0] = c->alloc<double>(nmax);
  w.p[1] = c->alloc<double>(nmax);
  w.v[0] = c->alloc<double>(nmax);
  w.v[1] = c->alloc<double>(nmax);
  w.t = c->alloc<double>(nmax);
  L.partials = c->alloc<double>((int64_t)kPartialSlots * std::max(kMaxTilesPerSolve, kMaxRedBlocks));
  w.st = c->alloc<BState>(1);
  HCHK(hipMemsetAsync(w.st, 0, sizeof(BState), c->stream));
  L.result = c->alloc<double>(16);
  HCHK(hipHostMalloc((void**)&L.hstate, sizeof(BState), hipHostMallocDefault));
  HCHK(hipHostMalloc((void**)&L.hresult, 16 * sizeof(double), hipHostMallocDefault));
  if (c->precision == NS3D_PRECISION_MIXED) alloc_lane32(c, L);
}

// FP32 copies of a system after its assembly (mixed form only).
void refresh_f32(ns3d_ctx* c, int sid) {
  if (c->precision != NS3D_PRECISION_MIXED) return;
  SysMem& S = c->sys[sid];
  if (S.n == 0) return;
  if (!S.coef32) {
    S.coef32 = c->alloc<float>((int64_t)7 * S.n);
    S.invdiag32 = c->alloc<float>(S.n);
  }
  HCHK(launch_coef_to_f32(S.coef, S.invdiag, S.n, S.coef32, S.invdiag32, c->stream));
  if (sid == 3) c->mg.dirty = true;
}

void alloc_system(ns3d_ctx* c, int sid, int n, int d0, int32_t* grid, bool momentum) {
  SysMem& S = c->sys[sid];
  S.n = n;
  S.d0 = d0;
  S.grid = grid;
  S.cell = c->alloc<int32_t>(n);
  S.nbr = c->alloc<int32_t>((int64_t)6 * n);
  S.coef = c->alloc<double>((int64_t)7 * n);
  S.invdiag = c->alloc<double>(n);
  S.rowsum = momentum ? c->alloc<double>(n) : nullptr;
}

}  // namespace

extern "C" {

const char* ns3d_last_error(void) { return g_err.c_str(); }

int ns3d_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return -1;
  return n;
}

ns3d_status ns3d_prepare_image(const uint8_t* img, int nx, int ny, int nz, int direction, int border,
                               uint8_t* out, int dims[3]) {
  if (!img || !dims || nx <= 0 || ny <= 0 || nz <= 0 || border < 0 || direction < 1 || direction > 3)
    return fail_invalid("ns3d_prepare_image: bad arguments");
  int ox = nx, oy = ny, oz = nz;
  if (direction == 2) { ox = ny; oy = nx; oz = nz; }
  if (direction == 3) { ox = nz; oy = ny; oz = nx; }
  dims[0] = ox + 2 * border;
  dims[1] = oy;
  dims[2] = oz;
  if (!out) return NS3D_OK;
  const size_t plane = (size_t)oy * oz;
  std::memset(out, 0, (size_t)dims[0] * plane);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const uint8_t v = img[((size_t)i * ny + j) * nz + k];
        int a, bq, cq;
        if (direction == 2) { a = j; bq = nx - 1 - i; cq = k; }       // NSBase.cxx:2468
        else if (direction == 3) { a = k; bq = j; cq = i; }           // NSBase.cxx:2482
        else { a = i; bq = j; cq = k; }
        out[(size_t)(a + border) * plane + (size_t)bq * oz + cq] = v;
      }
  return NS3D_OK;
}

ns3d_status ns3d_create(const uint8_t* img, int nx, int ny, int nz, int device, ns3d_ctx** out) {
  if (!img || !out || nx <= 0 || ny <= 0 || nz <= 0) return fail_invalid("ns3d_create: bad arguments");
  const int64_t ncell = (int64_t)nx * ny * nz;
  if (ncell + (int64_t)ny * nz >= ((int64_t)1 << 31) - 1)
    return fail_invalid("ns3d_create: image too large for 32-bit node indices");
  *out = nullptr;
  ns3d_ctx* c = new ns3d_ctx();
  ns3d_status st = guarded([&] {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) throw Invalid{"no GPU visible"};
    if (device >= 0) {
      if (device >= ndev) throw Invalid{"ns3d_create: device index out of range"};
      HCHK(hipSetDevice(device));
    }
    HCHK(hipGetDevice(&c->device));
    if (const char* e = std::getenv("NS3D_RESTART_EPS2")) c->restart_eps2 = std::max(0.0, std::atof(e));
    HCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = c->stream;
    HCHK(hipEventCreateWithFlags(&c->ev_sync, hipEventDisableTiming));
    HCHK(hipEventCreate(&c->ev[0]));
    HCHK(hipEventCreate(&c->ev[1]));
    c->nx = nx; c->ny = ny; c->nz = nz; c->ncell = ncell;
    hipStream_t s = c->stream;
    c->img = c->alloc<uint8_t>(ncell);
    HCHK(hipMemcpyAsync(c->img, img, (size_t)ncell, hipMemcpyHostToDevice, s));
    const int64_t ugrid = (int64_t)(nx + 1) * ny * nz;
    c->pg = c->alloc<int32_t>(ncell);
    c->vg = c->alloc<int32_t>(ncell);
    c->wg = c->alloc<int32_t>(ncell);
    c->ug = c->alloc<int32_t>(ugrid);
    ScanBuffers sb;
    sb.ncell = ncell;
    sb.nblocks = (int)((ncell + kScanCells - 1) / kScanCells);
    sb.blockcounts = c->alloc<int32_t>((int64_t)4 * sb.nblocks);
    sb.blockoffsets = c->alloc<int32_t>((int64_t)4 * sb.nblocks);
    sb.totals = c->alloc<int32_t>(4);
    HCHK(launch_scan_count(c->img, nx, ny, nz, sb, s));
    int32_t tot[4];
    HCHK(hipMemcpyAsync(tot, sb.totals, sizeof(tot), hipMemcpyDeviceToHost, s));
    HCHK(ctx_sync(c));
    const int np = tot[0], nu = tot[1], nv = tot[2], nw = tot[3];
    alloc_system(c, 0, nu, nx + 1, c->ug, true);
    alloc_system(c, 1, nv, nx, c->vg, true);
    alloc_system(c, 2, nw, nx, c->wg, true);
    alloc_system(c, 3, np, nx, c->pg, false);
    HCHK(launch_scan_write(c->img, nx, ny, nz, sb, c->pg, c->ug, c->vg, c->wg, c->sys[3].cell,
                           c->sys[0].cell, c->sys[1].cell, c->sys[2].cell, s));
    const Geo g = c->geo();
    unsigned long long* dnnz = c->alloc<unsigned long long>(4);
    HCHK(hipMemsetAsync(dnnz, 0, 4 * sizeof(unsigned long long), s));
    for (int q = 0; q < 4; ++q)
      HCHK(launch_build_nbr(g, q, c->sys[q].cell, c->sys[q].n, c->sys[q].nbr, dnnz + q, s));
    unsigned long long hnnz[4];
    HCHK(hipMemcpyAsync(hnnz, dnnz, sizeof(hnnz), hipMemcpyDeviceToHost, s));
    HCHK(ctx_sync(c));
    for (int q = 0; q < 4; ++q) c->sys[q].nnz = (int64_t)hnnz[q];
    // packed neighbour distances where every distance fits (always, for the
    // monotone numbering of the index scan, unless a plane is very large)
    int* dbad = c->alloc<int>(4);
    HCHK(hipMemsetAsync(dbad, 0, 4 * sizeof(int), s));
    for (int q = 0; q < 4; ++q) {
      c->sys[q].pk = c->alloc<uint64_t>(c->sys[q].n);
      HCHK(launch_pack_neighbours(c->sys[q].nbr, c->sys[q].n, c->sys[q].pk, dbad + q, s));
    }
    int hbad[4];
    HCHK(hipMemcpyAsync(hbad, dbad, sizeof(hbad), hipMemcpyDeviceToHost, s));
    HCHK(ctx_sync(c));
    for (int q = 0; q < 4; ++q)
      c->sys[q].layout = hbad[q] ? ML_INDEXED : (q == 3 ? ML_PACKED_SYM : ML_PACKED);
    // plane-band order for the solver kernels of large systems (the row
    // threshold can be lowered with NS3D_CHUNK_MIN_ROWS, e.g. to test the
    // order on small images)
    int64_t min_rows = kChunkMinRows;
    if (const char* e = std::getenv("NS3D_CHUNK_MIN_ROWS")) min_rows = std::max<int64_t>(1, std::atoll(e));
    std::vector<int32_t> hcell;
    for (int q = 0; q < 4; ++q) {
      SysMem& S = c->sys[q];
      if (S.n < min_rows || ny < kBands) continue;
      hcell.resize((size_t)S.n);
      HCHK(hipMemcpyAsync(hcell.data(), S.cell, (size_t)S.n * 4, hipMemcpyDeviceToHost, s));
      HCHK(ctx_sync(c));
      const std::vector<int> chk = build_chunks(hcell.data(), S.n, ny, nz, min_rows);
      if (chk.empty()) continue;
      S.chk = c->alloc<int>((int64_t)chk.size());
      HCHK(hipMemcpyAsync(S.chk, chk.data(), chk.size() * 4, hipMemcpyHostToDevice, s));
      S.band_off.assign(chk.begin(), chk.begin() + kBands + 1);
      HCHK(ctx_sync(c));
    }
    for (int set = 0; set < 3; ++set)
      for (int comp = 0; comp < 3; ++comp) c->vel[set][comp] = c->alloc<double>(c->sys[comp].n);
    c->p = c->alloc<double>(np);
    c->pc = c->alloc<double>(np);
    c->b = c->alloc<double>(np);
    int nmax = np;
    for (int q = 0; q < 3; ++q) nmax = std::max(nmax, c->sys[q].n);
    c->lane_rows = nmax;
    alloc_lane(c, c->ln[0]);
    c->nlanes = 1;
    HCHK(ctx_sync(c));
  });
  if (st != NS3D_OK) {
    ns3d_destroy(c);
    return st;
  }
  const ns3d_status z = ns3d_zero_fields(c);
  if (z != NS3D_OK) {
    ns3d_destroy(c);
    return z;
  }
  *out = c;
  return NS3D_OK;
}

void ns3d_destroy(ns3d_ctx* c) {
  if (!c) return;
  // rounds an engine queued before this context's last solve returned may
  // still name its state buffers
  if (c->engine) {
    c->engine->fence();
    c->engine->attached.fetch_sub(1);
  }
  // its own work done (an attached context's stream is shared)
  if (c->stream) {
    if (c->ev_sync) (void)ctx_sync(c);
    else (void)hipStreamSynchronize(c->stream);
  }
  c->stream = c->own_stream;
  for (void* p : c->allocs) (void)hipFree(p);
  for (Lane& L : c->ln) {
    if (L.hstate) (void)hipHostFree(L.hstate);
    if (L.hresult) (void)hipHostFree(L.hresult);
  }
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->ev_sync) (void)hipEventDestroy(c->ev_sync);
  if (c->mg.hst) (void)hipHostFree(c->mg.hst);
  for (auto& e : c->mg.execs) (void)hipGraphExecDestroy(e.second);
  for (auto& e : c->mg.tev)
    if (e) (void)hipEventDestroy(e);
  if (c->mg.cap) (void)hipStreamDestroy(c->mg.cap);
  for (auto& e : c->tev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

ns3d_status ns3d_counts(ns3d_ctx* c, int64_t out[4]) {
  if (!c || !out) return fail_invalid("ns3d_counts: null argument");
  out[0] = c->sys[3].n; out[1] = c->sys[0].n; out[2] = c->sys[1].n; out[3] = c->sys[2].n;
  return NS3D_OK;
}

ns3d_status ns3d_get_index(ns3d_ctx* c, int32_t* p, int32_t* u, int32_t* v, int32_t* w) {
  if (!c) return fail_invalid("ns3d_get_index: null context");
  return guarded([&] {
    const size_t nc = (size_t)c->ncell * 4, nuc = (size_t)(c->nx + 1) * c->ny * c->nz * 4;
    if (p) HCHK(hipMemcpyAsync(p, c->pg, nc, hipMemcpyDeviceToHost, c->stream));
    if (u) HCHK(hipMemcpyAsync(u, c->ug, nuc, hipMemcpyDeviceToHost, c->stream));
    if (v) HCHK(hipMemcpyAsync(v, c->vg, nc, hipMemcpyDeviceToHost, c->stream));
    if (w) HCHK(hipMemcpyAsync(w, c->wg, nc, hipMemcpyDeviceToHost, c->stream));
    HCHK(ctx_sync(c));
  });
}

ns3d_status ns3d_set_params(ns3d_ctx* c, const ns3d_params* prm) {
  if (!c || !prm) return fail_invalid("ns3d_set_params: null argument");
  if (!(prm->dx > 0) || !(prm->dy > 0) || !(prm->dz > 0)) return fail_invalid("ns3d_set_params: spacing must be > 0");
  if (prm->border < 0 || 2 * prm->border >= c->nx) return fail_invalid("ns3d_set_params: border does not fit the image");
  c->prm = *prm;
  c->have_params = true;
  return NS3D_OK;
}

ns3d_status ns3d_zero_fields(ns3d_ctx* c) {
  if (!c) return fail_invalid("ns3d_zero_fields: null context");
  return guarded([&] {
    for (int set = 0; set < 3; ++set)
      for (int comp = 0; comp < 3; ++comp) HCHK(launch_fill(c->vel[set][comp], c->sys[comp].n, 0.0, c->stream));
    HCHK(launch_fill(c->p, c->sys[3].n, 0.0, c->stream));
    HCHK(launch_fill(c->pc, c->sys[3].n, 0.0, c->stream));
    HCHK(launch_fill(c->b, c->sys[3].n, 0.0, c->stream));
    HCHK(ctx_sync(c));
  });
}

ns3d_status ns3d_init_pressure_linear(ns3d_ctx* c) {
  if (!c || !c->have_params) return fail_invalid("ns3d_init_pressure_linear: parameters not set");
  return guarded([&] {
    HCHK(launch_init_pressure_linear(c->geo(), c->sys[3].cell, c->sys[3].n, c->p, c->prm.p0, c->prm.p1,
                                     c->stream));
    HCHK(ctx_sync(c));
  });
}

ns3d_status ns3d_set_vector(ns3d_ctx* c, int which, const double* host, int64_t n) {
  if (!c || !host) return fail_invalid("ns3d_set_vector: null argument");
  return guarded([&] {
    int64_t len;
    double* d = vec_ptr(c, which, len);
    need(n == len, "ns3d_set_vector: length mismatch");
    if (len) HCHK(hipMemcpyAsync(d, host, (size_t)len * 8, hipMemcpyHostToDevice, c->stream));
    HCHK(ctx_sync(c));
  });
}

ns3d_status ns3d_get_vector(ns3d_ctx* c, int which, double* host, int64_t n) {
  if (!c || !host) return fail_invalid("ns3d_get_vector: null argument");
  return guarded([&] {
    int64_t len;
    double* d = vec_ptr(c, which, len);
    need(n == len, "ns3d_get_vector: length mismatch");
    if (len) HCHK(hipMemcpyAsync(host, d, (size_t)len * 8, hipMemcpyDeviceToHost, c->stream));
    HCHK(ctx_sync(c));
  });
}

ns3d_status ns3d_copy_velocity(ns3d_ctx* c, int dst, int src) {
  if (!c || dst < 0 || dst > 2 || src < 0 || src > 2) return fail_invalid("ns3d_copy_velocity: bad set");
  return guarded([&] {
    if (dst == src) return;
    for (int q = 0; q < 3; ++q)
      if (c->sys[q].n)
        HCHK(hipMemcpyAsync(c->vel[dst][q], c->vel[src][q], (size_t)c->sys[q].n * 8, hipMemcpyDeviceToDevice,
                            c->stream));
  });
}

ns3d_status ns3d_assemble_momentum(ns3d_ctx* c, double alpha) {
  if (!c || !c->have_params) return fail_invalid("ns3d_assemble_momentum: parameters not set");
  return guarded([&] {
    const Phys ph = phys_of(c, alpha, 1.0);
    const Geo g = c->geo();
    const Vel3 vc = vel3(c, NS3D_SET_VEL);
    HCHK(launch_assemble_momentum(0, g, c->sys[0].view(), vc, ph, c->stream));
    HCHK(launch_assemble_momentum(1, g, c->sys[1].view(), vc, ph, c->stream));
    if (c->sys[2].n > 0) HCHK(launch_assemble_momentum(2, g, c->sys[2].view(), vc, ph, c->stream));
    for (int q = 0; q < 3; ++q) refresh_f32(c, q);
  });
}

ns3d_status ns3d_compute_velocity_field(ns3d_ctx* c, double alpha, int limit, double dvmax, double tol,
                                        int64_t maxit, ns3d_solve_info info[3]) {
  if (!c || !c->have_params || !info) return fail_invalid("ns3d_compute_velocity_field: bad arguments");
  return guarded([&] {
    const Phys ph = phys_of(c, alpha, 1.0);
    const Geo g = c->geo();
    hipStream_t s = c->stream;
    for (int q = 0; q < 3; ++q) info[q] = ns3d_solve_info{};
    // the three systems are independent (each right-hand side reads the
    // pressure and its own component only): with an engine and three lanes
    // they are solved at once, else one after the other on lane 0
    const bool together = c->engine && c->nlanes >= 3;
    int sid[3];
    double* xs[3];
    const double* bs[3];
    Lane* Ls[3];
    int m = 0;
    auto finish = [&](int q, Lane& L) {
      const int64_t n = c->sys[q].n;
      double* x = c->vel[NS3D_SET_VEL][q];
      if (info[q].iterations == 0 || info[q].info != 0)
        HCHK(hipMemcpyAsync(x, L.temp, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
      if (limit) HCHK(launch_clamp(x, L.temp, n, dvmax, s));
    };
    for (int q = 0; q < 3; ++q) {
      SysMem& S = c->sys[q];
      if (q == 2 && S.n <= 0) break;
      Lane& L = c->ln[together ? q : 0];
      double* x = c->vel[NS3D_SET_VEL][q];
      HCHK(launch_rhs(q, g, S.view(), c->p, x, L.vb, ph, s));
      HCHK(hipMemcpyAsync(L.temp, x, (size_t)S.n * 8, hipMemcpyDeviceToDevice, s));
      if (!together) {
        bicgstab(c, q, x, L.vb, tol, maxit, &info[q]);
        finish(q, L);
        continue;
      }
      sid[m] = q; xs[m] = x; bs[m] = L.vb; Ls[m] = &L;
      ++m;
    }
    if (together) {
      bicgstab_many(c, m, sid, xs, bs, tol, maxit, info, Ls);
      for (int q = 0; q < m; ++q) finish(q, c->ln[q]);
    }
    HCHK(ctx_sync(c));
  });
}

ns3d_status ns3d_compute_hat(ns3d_ctx* c, double alpha, int uc_set, int uold_set) {
  if (!c || !c->have_params) return fail_invalid("ns3d_compute_hat: parameters not set");
  return guarded([&] {
    need(uc_set != NS3D_SET_HAT && uold_set != NS3D_SET_HAT, "ns3d_compute_hat: HAT is the output set");
    const Phys ph = phys_of(c, alpha, 1.0);
    const Geo g = c->geo();
    for (int q = 0; q < 3; ++q)
      HCHK(launch_hat(q, g, c->sys[q].view(), c->vel[uc_set][q], c->vel[uold_set][q], c->vel[NS3D_SET_HAT][q],
                      ph, c->stream));
  });
}

ns3d_status ns3d_assemble_pressure(ns3d_ctx* c, int kind, int relax, double beta, int vel_set) {
  if (!c || !c->have_params) return fail_invalid("ns3d_assemble_pressure: parameters not set");
  if (kind < 0 || kind > 2) return fail_invalid("ns3d_assemble_pressure: unknown kind");
  return guarded([&] {
    const Phys ph = phys_of(c, 1.0, beta);
    HCHK(launch_assemble_pressure(c->geo(), c->sys[3].view(), c->sys[0].view(), c->sys[1].view(),
                                  c->sys[2].view(), vel3(c, vel_set), c->p, c->b, kind, relax, ph, c->stream));
    refresh_f32(c, 3);
  });
}

ns3d_status ns3d_compute_pressure_correction(ns3d_ctx* c, double tol, int64_t maxit, ns3d_solve_info* info) {
  if (!c || !info) return fail_invalid("ns3d_compute_pressure_correction: bad arguments");
  return guarded([&] {
    bicgstab(c, 3, c->pc, c->b, tol, maxit, info);
    HCHK(ctx_sync(c));
  });
}

ns3d_status ns3d_compute_pressure_field(ns3d_ctx* c, double tol, int64_t maxit, ns3d_solve_info* info) {
  if (!c || !info) return fail_invalid("ns3d_compute_pressure_field: bad arguments");
  return guarded([&] {
    hipStream_t s = c->stream;
    const int64_t n = c->sys[3].n;
    double* ptemp = c->ln[0].temp;  // temp holds max(n) doubles
    HCHK(hipMemcpyAsync(ptemp, c->p, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
    bicgstab(c, 3, ptemp, c->b, tol, maxit, info);
    int64_t work = info->work;
    while (info->iterations == 0 && info->info == 0 && tol > 1e-15) {
      tol *= 0.1;
      bicgstab(c, 3, ptemp, c->b, tol, maxit, info);
      work += info->work;
    }
    info->work = work;
    if (info->iterations > 0) HCHK(hipMemcpyAsync(c->p, ptemp, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
    HCHK(ctx_sync(c));
  });
}

ns3d_status ns3d_correct_velocity(ns3d_ctx* c, int kind, double dvmax, int target_set, double p_0, double p_1) {
  if (!c || !c->have_params) return fail_invalid("ns3d_correct_velocity: parameters not set");
  if (kind < 0 || kind > 2 || target_set < 0 || target_set > 2) return fail_invalid("ns3d_correct_velocity: bad kind/set");
  return guarded([&] {
    const double* q = (kind == NS3D_CORR_SIMPLER) ? c->p : c->pc;
    HCHK(launch_correct_velocity(c->geo(), c->sys[3].view(), c->sys[0].view(), c->sys[1].view(), c->sys[2].view(),
                                 kind, q, c->vel[target_set][0], c->vel[target_set][1], c->vel[target_set][2],
                                 c->vel[NS3D_SET_HAT][0], c->vel[NS3D_SET_HAT][1], c->vel[NS3D_SET_HAT][2], dvmax,
                                 p_0, p_1, c->prm.dx, c->prm.dy, c->prm.dz, c->stream));
  });
}

ns3d_status ns3d_correct_pressure(ns3d_ctx* c, double beta, double out[3]) {
  if (!c || !out) return fail_invalid("ns3d_correct_pressure: bad arguments");
  return guarded([&] {
    const int64_t n = c->sys[3].n;
    need(n > 0, "ns3d_correct_pressure: no pressure nodes");
    HCHK(launch_correct_pressure(c->p, c->pc, n, beta, c->ln[0].partials, c->ln[0].result, c->stream));
    HCHK(hipMemcpyAsync(c->ln[0].hresult, c->ln[0].result, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HCHK(ctx_sync(c));
    out[0] = beta * c->ln[0].hresult[0];
    out[1] = beta * c->ln[0].hresult[1];
    out[2] = beta * std::sqrt(c->ln[0].hresult[2]);
  });
}

ns3d_status ns3d_compute_div(ns3d_ctx* c, int vel_set, double out[4]) {
  if (!c || !c->have_params || !out) return fail_invalid("ns3d_compute_div: bad arguments");
  return guarded([&] {
    const Vel3 v = vel3(c, vel_set);
    if (c->sys[3].n == 0) { out[0] = out[1] = out[2] = out[3] = 0; return; }
    HCHK(launch_compute_div(c->geo(), c->sys[3].view(), v.u, v.v, v.w, c->prm.dx, c->prm.dy, c->prm.dz,
                            c->ln[0].partials, c->ln[0].result, c->stream));
    HCHK(hipMemcpyAsync(c->ln[0].hresult, c->ln[0].result, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HCHK(ctx_sync(c));
    out[0] = c->ln[0].hresult[0];
    out[1] = std::sqrt(c->ln[0].hresult[1]);
    out[2] = c->ln[0].hresult[2] / (double)c->sys[3].n;
    out[3] = c->ln[0].hresult[3];
  });
}

ns3d_status ns3d_compute_permeability(ns3d_ctx* c, double out[5]) {
  if (!c || !c->have_params || !out) return fail_invalid("ns3d_compute_permeability: bad arguments");
  return guarded([&] {
    const int border = c->prm.border;
    HCHK(launch_permeability(c->geo(), c->vel[NS3D_SET_VEL][0], c->p, border, c->ln[0].partials, c->ln[0].result, c->stream));
    HCHK(hipMemcpyAsync(c->ln[0].hresult, c->ln[0].result, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HCHK(ctx_sync(c));
    const double* r = c->ln[0].hresult;
    double vm = r[0];
    double p_in, p_out;
    if (border <= 0) {
      p_in = c->prm.p0;
      p_out = c->prm.p1;
    } else {
      p_in = r[3];
      p_out = r[1];
      if (r[4] > 0) p_in /= r[4];
      if (r[2] > 0) p_out /= r[2];
    }
    const double Q = vm * (c->prm.dy * c->prm.dz);
    vm /= (double)(c->ny * c->nz);
    double K = 0.0;
    if (std::fabs(p_in - p_out) > 0.0)
      K = vm * c->prm.mu * c->prm.dx * (double)(c->nx - 2 * border) / (p_in - p_out);
    out[0] = Q; out[1] = vm; out[2] = K; out[3] = p_in; out[4] = p_out;
  });
}

ns3d_status ns3d_get_matrix(ns3d_ctx* c, int sys, double* diag, double* off, int32_t* nbr, double* rowsum) {
  if (!c || sys < 0 || sys > 3) return fail_invalid("ns3d_get_matrix: bad arguments");
  return guarded([&] {
    const SysMem& S = c->sys[sys];
    const size_t n = (size_t)S.n;
    hipStream_t s = c->stream;
    if (diag) HCHK(hipMemcpyAsync(diag, S.coef, n * 8, hipMemcpyDeviceToHost, s));
    if (off) HCHK(hipMemcpyAsync(off, S.coef + n, 6 * n * 8, hipMemcpyDeviceToHost, s));
    if (nbr) HCHK(hipMemcpyAsync(nbr, S.nbr, 6 * n * 4, hipMemcpyDeviceToHost, s));
    if (rowsum && S.rowsum) HCHK(hipMemcpyAsync(rowsum, S.rowsum, n * 8, hipMemcpyDeviceToHost, s));
    HCHK(ctx_sync(c));
  });
}

ns3d_status ns3d_bicgstab_stencil(int n, const double* diag, const double* off, const int32_t* nbr,
                                  const double* b, double* x, double tol, int64_t maxit, ns3d_solve_info* info) {
  if (n <= 0 || !diag || !off || !nbr || !b || !x || !info) return fail_invalid("ns3d_bicgstab_stencil: bad arguments");
  for (int64_t q = 0; q < (int64_t)6 * n; ++q)
    if (nbr[q] < 0 || nbr[q] >= n) return fail_invalid("ns3d_bicgstab_stencil: neighbour index out of range");
  // A throw-away context holding one "p" system of n rows.
  ns3d_ctx* c = new ns3d_ctx();
  ns3d_status st = guarded([&] {
    HCHK(hipGetDevice(&c->device));
    if (const char* e = std::getenv("NS3D_RESTART_EPS2")) c->restart_eps2 = std::max(0.0, std::atof(e));
    HCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = c->stream;
    HCHK(hipEventCreateWithFlags(&c->ev_sync, hipEventDisableTiming));
    HCHK(hipEventCreate(&c->ev[0]));
    HCHK(hipEventCreate(&c->ev[1]));
    hipStream_t s = c->stream;
    SysMem& S = c->sys[3];
    S.n = n;
    S.nbr = c->alloc<int32_t>((int64_t)6 * n);
    S.coef = c->alloc<double>((int64_t)7 * n);
    S.invdiag = c->alloc<double>(n);
    HCHK(hipMemcpyAsync(S.nbr, nbr, (size_t)6 * n * 4, hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(S.coef, diag, (size_t)n * 8, hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(S.coef + n, off, (size_t)6 * n * 8, hipMemcpyHostToDevice, s));
    HCHK(launch_invdiag(S.view(), s));
    double* db = c->alloc<double>(n);
    double* dx = c->alloc<double>(n);
    HCHK(hipMemcpyAsync(db, b, (size_t)n * 8, hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(dx, x, (size_t)n * 8, hipMemcpyHostToDevice, s));
    c->lane_rows = n;
    alloc_lane(c, c->ln[0]);
    c->nlanes = 1;
    bicgstab(c, 3, dx, db, tol, maxit, info);
    HCHK(hipMemcpyAsync(x, dx, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    HCHK(ctx_sync(c));
  });
  ns3d_destroy(c);
  return st;
}

}  // extern "C"

// ---- the stand-alone solve without allocation --------------------------------
// Workspace of ns3d_bicgstab_stencil_noalloc, one copy per GPU for the life of
// the process: the matrix coefficients [7][n] and M^-1, the seven BiCGSTAB
// vectors, the block partials, the solver state and a flag (in doubles).
// Vector regions start on 256-byte boundaries.  At the default capacity of
// 2^20 rows the array takes 120.1 MiB of device memory.
#ifndef NS3D_NOALLOC_MAX_ROWS
#define NS3D_NOALLOC_MAX_ROWS 1048576
#endif
namespace {
constexpr int64_t kNoAllocVec = NS3D_NOALLOC_MAX_ROWS;  // a multiple of 32
constexpr int64_t kNoAllocPartials = (int64_t)kPartialSlots * std::max(kMaxTilesPerSolve, kMaxRedBlocks);
constexpr int64_t kNoAllocDoubles = 15 * kNoAllocVec + kNoAllocPartials + 64;
static_assert(kNoAllocVec % 32 == 0, "row capacity must keep the regions 256-byte aligned");
static_assert(sizeof(BState) <= 56 * sizeof(double), "solver state does not fit its slot");
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

constexpr int kNoAllocDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kNoAllocDevices] = {};

// The device address of this GPU's scratch copy, looked up once per device
// (with that device current); the array's name is not a device address in
// host code.
double* scratch_of(int dev) {
  if (dev < 0 || dev >= kNoAllocDevices) throw Invalid{"ns3d_bicgstab_stencil_noalloc: device ordinal out of range"};
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double*)ptr;
  }
  return g_scratch_ptr[dev];
}

// Throws unless p is device memory of device dev.
void need_device_ptr(const void* p, int dev, const char* what) {
  hipPointerAttribute_t a{};
  const hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Invalid{std::string("ns3d_bicgstab_stencil_noalloc: ") + what + " is not device memory"};
  }
  if ((a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged) || a.device != dev)
    throw Invalid{std::string("ns3d_bicgstab_stencil_noalloc: ") + what + " is not device memory of the current device"};
}

}  // namespace

extern "C" {

ns3d_status ns3d_bicgstab_stencil_noalloc(int n, const double* diag, const double* off, const int32_t* nbr,
                                          const double* b, double* x, double tol, int64_t maxit,
                                          ns3d_solve_info* info, void* stream) {
  if (n <= 0 || !diag || !off || !nbr || !b || !x || !info)
    return fail_invalid("ns3d_bicgstab_stencil_noalloc: bad arguments");
  if (n > kNoAllocVec)
    return fail_invalid("ns3d_bicgstab_stencil_noalloc: n = " + std::to_string(n) +
                        " exceeds the workspace maximum of " + std::to_string(kNoAllocVec) + " rows");
  return guarded([&] {
    int dev = 0;
    HCHK(hipGetDevice(&dev));
    need_device_ptr(diag, dev, "diag");
    need_device_ptr(off, dev, "off");
    need_device_ptr(nbr, dev, "nbr");
    need_device_ptr(b, dev, "b");
    need_device_ptr(x, dev, "x");
    hipStream_t s = (hipStream_t)stream;
    double* base = scratch_of(dev);
    const int64_t m = ((int64_t)n + 31) / 32 * 32;
    double* coef = base;                      // [7][n]: diag, then E, W, N, S, T, B
    double* invdiag = base + 7 * m;
    double* vec = base + 8 * m;               // r, r0, p0, p1, v0, v1, t
    double* partials = base + 15 * m;
    BState* st = (BState*)(partials + kNoAllocPartials);
    double* bad = partials + kNoAllocPartials + 56;
    // neighbour rows in range, checked on the device before any solver
    // kernel gathers through them
    HCHK(launch_fill(bad, 1, 0.0, s));
    HCHK(launch_check_index(nbr, (int64_t)6 * n, n, bad, s));
    double hbad = 0.0;
    HCHK(hipMemcpyAsync(&hbad, bad, sizeof(double), hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    if (hbad != 0.0) throw Invalid{"ns3d_bicgstab_stencil_noalloc: neighbour index out of range"};
    HCHK(hipMemcpyAsync(coef, diag, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
    HCHK(hipMemcpyAsync(coef + n, off, (size_t)6 * n * 8, hipMemcpyDeviceToDevice, s));
    HCHK(launch_invdiag(SysView{n, 0, nullptr, nullptr, nbr, coef, invdiag, nullptr}, s));
    const SMat<double> M{n, ML_INDEXED, nbr, nullptr, coef, invdiag, nullptr};
    const BWork w{vec, vec + m, {vec + 2 * m, vec + 3 * m}, {vec + 4 * m, vec + 5 * m}, vec + 6 * m, st};
    const SolveDesc<double> d{M, w, x, b, tol, (long long)maxit, partials, kMaxTilesPerSolve, 0, 0, 0};
    BState hs{};
    int64_t hint = 16;
    queued_solve<double>(s, d, &hs, hint, nullptr, [&] { HCHK(hipStreamSynchronize(s)); }, [] {});
    if (solve_info(hs, tol, maxit, *info)) HCHK(launch_fill(x, n, 0.0, s));
    HCHK(hipStreamSynchronize(s));
  });
}

ns3d_status ns3d_noalloc_max_rows(int64_t* rows) {
  if (!rows) return fail_invalid("ns3d_noalloc_max_rows: null argument");
  *rows = kNoAllocVec;
  return NS3D_OK;
}

ns3d_status ns3d_timing(ns3d_ctx* c, int enable) {
  if (!c) return fail_invalid("ns3d_timing: null context");
  return guarded([&] {
    c->tm = Timing();
    c->tm.on = enable != 0;
    if (c->tm.on && c->tev.empty()) {
      c->tev.resize(6, nullptr);
      for (auto& e : c->tev) HCHK(hipEventCreate(&e));
    }
  });
}

ns3d_status ns3d_timing_read(ns3d_ctx* c, double out[36]) {
  if (!c || !out) return fail_invalid("ns3d_timing_read: bad arguments");
  for (int prec = 0; prec < 2; ++prec)
    for (int kind = 0; kind < 2; ++kind) {
      for (int k = 0; k < 3; ++k) {
        double* o = out + prec * 18 + kind * 9 + k * 3;
        o[0] = c->tm.ms[prec][kind][k];
        o[1] = c->tm.launches[prec][kind][k];
        o[2] = c->tm.bytes[prec][kind][k];
      }
    }
  return NS3D_OK;
}

ns3d_status ns3d_set_precision(ns3d_ctx* c, int mode, double inner_tol) {
  if (!c || (mode != NS3D_PRECISION_FP64 && mode != NS3D_PRECISION_MIXED))
    return fail_invalid("ns3d_set_precision: bad arguments");
  if (mode == NS3D_PRECISION_MIXED && !(inner_tol > 0 && inner_tol < 1))
    return fail_invalid("ns3d_set_precision: inner tolerance must be in (0, 1)");
  return guarded([&] {
    if (mode == NS3D_PRECISION_MIXED)
      for (int q = 0; q < c->nlanes; ++q) alloc_lane32(c, c->ln[q]);
    c->precision = mode;
    if (mode == NS3D_PRECISION_MIXED) c->inner_tol = inner_tol;
    for (int q = 0; q < 4; ++q) refresh_f32(c, q);
    HCHK(ctx_sync(c));
  });
}

ns3d_status ns3d_set_multigrid(ns3d_ctx* c, int64_t rows) {
  if (!c || rows < 0) return fail_invalid("ns3d_set_multigrid: bad arguments");
  c->mg_min_rows = rows;
  return NS3D_OK;
}

ns3d_status ns3d_multigrid_levels(ns3d_ctx* c, int* levels, int64_t* rows) {
  if (!c || !levels) return fail_invalid("ns3d_multigrid_levels: bad arguments");
  *levels = c->mg.ok ? c->mg.nlev : 0;
  if (rows)
    for (int l = 0; l < *levels; ++l) rows[l] = c->mg.lv[l].n;
  return NS3D_OK;
}

ns3d_status ns3d_multigrid_level(ns3d_ctx* c, int level, int32_t* cell, float* coef) {
  if (!c || !c->mg.ok || level < 1 || level >= c->mg.nlev) return fail_invalid("ns3d_multigrid_level: no such level");
  return guarded([&] {
    const MgLevel& L = c->mg.lv[level];
    if (cell) HCHK(hipMemcpyAsync(cell, L.cell, (size_t)L.n * 4, hipMemcpyDeviceToHost, c->stream));
    if (coef) HCHK(hipMemcpyAsync(coef, L.coef, (size_t)L.n * 16, hipMemcpyDeviceToHost, c->stream));
    HCHK(ctx_sync(c));
  });
}

ns3d_status ns3d_copy_bandwidth(int device, int64_t bytes, int reps, double* gbs) {
  if (bytes < 16 || reps < 1 || !gbs) return fail_invalid("ns3d_copy_bandwidth: bad arguments");
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const ns3d_status st = guarded([&] {
    if (device >= 0) HCHK(hipSetDevice(device));
    const int64_t words = bytes / 16;
    HCHK(hipMalloc(&a, (size_t)words * 16));
    HCHK(hipMalloc(&b, (size_t)words * 16));
    HCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HCHK(hipEventCreate(&e0));
    HCHK(hipEventCreate(&e1));
    HCHK(hipMemsetAsync(a, 1, (size_t)words * 16, s));
    HCHK(launch_copy16(a, b, words, s));  // warm-up
    HCHK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) HCHK(launch_copy16(a, b, words, s));
    HCHK(hipEventRecord(e1, s));
    HCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HCHK(hipEventElapsedTime(&ms, e0, e1));
    *gbs = 2.0 * 16.0 * (double)words * reps / (ms * 1e-3) / 1e9;
  });
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return st;
}

ns3d_status ns3d_multigrid_timing(ns3d_ctx* c, double out[12]) {
  if (!c || !out) return fail_invalid("ns3d_multigrid_timing: bad arguments");
  for (int k = 0; k < 4; ++k) {
    out[3 * k] = c->mg.ms[k];
    out[3 * k + 1] = c->mg.launches[k];
    out[3 * k + 2] = c->mg.bytes[k];
  }
  return NS3D_OK;
}

ns3d_status ns3d_set_mixed_min_rows(ns3d_ctx* c, int64_t rows) {
  if (!c || rows < 0) return fail_invalid("ns3d_set_mixed_min_rows: bad arguments");
  c->mixed_min_rows = rows;
  return NS3D_OK;
}

ns3d_status ns3d_matrix_nnz(ns3d_ctx* c, int64_t out[4]) {
  if (!c || !out) return fail_invalid("ns3d_matrix_nnz: bad arguments");
  for (int q = 0; q < 4; ++q) out[q] = c->sys[q].nnz;
  return NS3D_OK;
}

ns3d_status ns3d_engine_create(int device, int chunk, ns3d_engine** out) {
  if (!out) return fail_invalid("ns3d_engine_create: null output");
  *out = nullptr;
  return guarded([&] {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) throw Invalid{"no GPU visible"};
    if (device < 0) HCHK(hipGetDevice(&device));
    if (device >= ndev) throw Invalid{"ns3d_engine_create: device index out of range"};
    if (chunk < 1 || chunk > 256) throw Invalid{"ns3d_engine_create: chunk must be in [1, 256]"};
    *out = new ns3d_engine(device, chunk);
  });
}

void ns3d_engine_destroy(ns3d_engine* e) { delete e; }

ns3d_status ns3d_attach_engine(ns3d_ctx* c, ns3d_engine* e) {
  if (!c) return fail_invalid("ns3d_attach_engine: null context");
  if (e && e->device != c->device) return fail_invalid("ns3d_attach_engine: engine and context on different devices");
  if (e && c->nlanes < 3 && c->sys[0].n > 0) {
    // lanes for the momentum solves to run at once
    const ns3d_status st = guarded([&] {
      for (int q = c->nlanes; q < 3; ++q) alloc_lane(c, c->ln[q]);
      c->nlanes = 3;
      HCHK(ctx_sync(c));
    });
    if (st != NS3D_OK) return st;
  }
  if (c->engine != e) {
    if (c->engine) {
      c->engine->fence();
      c->engine->attached.fetch_sub(1);
    }
    // the context's work moves to the engine's context stream (or back to
    // its own) once everything queued so far is done
    const ns3d_status st = guarded([&] { HCHK(ctx_sync(c)); });
    if (st != NS3D_OK) return st;
    c->stream = (e && e->ctx_stream) ? e->ctx_stream : c->own_stream;
    if (e) e->attached.fetch_add(1);
    c->engine = e;
    c->engine_used = false;
  }
  return NS3D_OK;
}

ns3d_status ns3d_engine_set_schedule(ns3d_engine* e, int chunk, int lanes, int graphs) {
  if (!e) return fail_invalid("ns3d_engine_set_schedule: null engine");
  if (chunk < 1 || chunk > 64 || lanes < 1 || lanes > ns3d_engine::kMaxLanes)
    return fail_invalid("ns3d_engine_set_schedule: chunk must be in [1, 64] and lanes in [1, 4]");
  return guarded([&] {
    if (!e->set_schedule(chunk, lanes, graphs != 0))
      throw Invalid{"ns3d_engine_set_schedule: the engine has solves pending or running"};
  });
}

ns3d_status ns3d_engine_get_schedule(ns3d_engine* e, int out[3]) {
  if (!e || !out) return fail_invalid("ns3d_engine_get_schedule: bad arguments");
  out[0] = e->chunk;
  out[1] = e->nlanes;
  out[2] = e->graphs ? 1 : 0;
  return NS3D_OK;
}

ns3d_status ns3d_engine_set_graphs(ns3d_engine* e, int on) {
  if (!e) return fail_invalid("ns3d_engine_set_graphs: null engine");
  e->set_graphs(on != 0);
  return NS3D_OK;
}

ns3d_status ns3d_engine_stats(ns3d_engine* e, double out[32]) {
  if (!e || !out) return fail_invalid("ns3d_engine_stats: bad arguments");
  e->stats(out, 32);
  return NS3D_OK;
}

ns3d_status ns3d_footprint(int nx, int ny, int nz, int64_t nfluid, int mixed, int lanes, double* bytes) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || nfluid < 0 || !bytes || (lanes != 1 && lanes != 3))
    return fail_invalid("ns3d_footprint: bad arguments");
  const double N = (double)nx * ny * nz, plane = (double)ny * nz, nf = (double)nfluid;
  // rows: p, v, w <= fluid cells; u <= fluid cells + one outlet plane
  const double nu = nf + plane, nmax = nu, rows_all = 3 * nf + nu;
  double b = N + 12 * N + 4 * (N + plane);            // image, index grids
  b += 2 * 16 * std::ceil(N / kScanCells) + 64;        // scan block counts / offsets
  b += rows_all * (4 + 24 + 56 + 8 + 8) + 3 * nmax * 8;  // cell, nbr, coef, invdiag, pk; rowsum (u, v, w)
  b += 4 * kMaxTilesPerSolve * sizeof(Tile);           // tile tables
  b += 8 * (rows_all / kBlock + 4.0 * kBands * (nx + 1)) + 4 * (4 * kChunkHeader + 64.0 * 1024);  // chunk lists
  b += 3 * 8 * (3 * nf + plane) + 24 * nf;             // velocity sets, p, pc, b
  // per lane: momentum rhs, saved field, BiCGSTAB vectors, partials, state
  const double lane = (2 + 7) * 8 * nmax + 8.0 * kPartialSlots * std::max(kMaxTilesPerSolve, kMaxRedBlocks) + 1024;
  b += lanes * lane + 4096;
  if (mixed) b += rows_all * (28 + 4) + lanes * nmax * (7 * 4 + 2 * 4 + 8);  // FP32 matrices; inner vectors, residual
  // multigrid of the pressure system (from the default threshold on, where
  // the pressure rows may reach it): level-0 vectors and aggregate map (r,
  // z, p[2], q, agg); coarser levels hold at most N/7 cells and rows (grid
  // 4 B per cell; cell, pk, coef, D^-1, agg, b, x 44 B per row); partials;
  // allocation rounding of ten allocations per level (each level has at
  // most 1/4 of the rows of the one above, down to kMgCoarseRows)
  if (mixed && nf >= (double)kMgDefaultMinRows) {
    int levels = 2;
    for (double r = nf; r > kMgCoarseRows && levels < kMgMaxLevels; r /= 4) ++levels;
    b += 24 * nf + 48 * N / 7 + 32.0 * kMaxRedBlocks + 64.0 * 1024 * (8 + 10 * levels);
  }
  b += 64.0 * 1024 * (48 + 26 * (lanes - 1));          // allocation rounding (64 KiB per allocation)
  // whole bytes: sums and differences of footprints stay exact in FP64
  *bytes = std::ceil(b);
  return NS3D_OK;
}

ns3d_status ns3d_ctx_bytes(ns3d_ctx* c, double* bytes) {
  if (!c || !bytes) return fail_invalid("ns3d_ctx_bytes: bad arguments");
  *bytes = c->alloc_bytes;
  return NS3D_OK;
}

ns3d_status ns3d_sync(ns3d_ctx* c) {
  if (!c) return fail_invalid("ns3d_sync: null context");
  return guarded([&] { HCHK(ctx_sync(c)); });
}

}  // extern "C"
