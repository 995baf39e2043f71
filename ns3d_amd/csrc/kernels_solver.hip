// kernels_solver.hip -- fused right-Jacobi BiCGSTAB with Eigen's iteration
// contract (Eigen/src/IterativeLinearSolvers/BiCGSTAB.h, 3.3/3.4):
//
//   r = b - A x; r0 = r; if |b|^2 == 0: x = 0, done
//   rho = alpha = w = 1; v = p = 0; tol2 = tol^2 |b|^2
//   while |r|^2 > tol2 and i < maxit:
//     rho_old = rho; rho = r0.r
//     if |rho| < eps^2 |r0|^2: r = b - A x; r0 = r; rho = |r0|^2 (restart;
//                               i = 0 on the first restart)
//     beta = (rho/rho_old)(alpha/w); p = r + beta (p - w v); y = M^-1 p
//     v = A y; alpha = rho / r0.v; s = r - alpha v; z = M^-1 s; t = A z
//     w = t.s / t.t (0 if t.t == 0); x += alpha y + w z; r = s - w t; ++i
//
// One iteration is three row kernels; each writes block partials of its dot
// products and the next one starts by summing them (every block sums the
// same partials in the same order, so all blocks hold identical scalars):
//   A: sum C's partials (r0.r, r.r): loop test, restart test, rho, beta;
//      p' = r + beta (p - w v) and y = M^-1 p' recomputed at the row and its
//      six neighbours, v' = A y, partial r0.v'
//   B: sum A's partials: alpha; s = r - alpha v' and z = M^-1 s on the fly,
//      t = A z, partials t.s, t.t
//   C: sum B's partials: w; x += alpha M^-1 p' + w M^-1 s, r = s - w t,
//      partials r0.r, r.r
// p and v are double-buffered (buffer parity = executed iterations), so the
// recomputation at neighbours never reads a value written in the same kernel.
// Block (0, thread 0) of each kernel records the scalars in BState for the
// host, which polls the status between chunks of queued iterations; kernels
// launched after the solve has stopped exit at once.  init and restart keep
// a one-block finalize kernel (once per solve / rare).
// Vectors and coefficients are stored as T (double for the plain path, float
// for the inner solves of the mixed form); every operation is evaluated in
// FP64 registers and reductions accumulate in FP64, so the FP64
// instantiation computes exactly the expressions above.
#include <hip/hip_ext.h>

#include <algorithm>
#include <type_traits>

#include "kernels.h"

namespace ns3d {

namespace {

constexpr double kEps = 2.220446049250313e-16;  // NumTraits<double>::epsilon()
constexpr int kSlotA = 0, kSlotB = 1, kSlotC = 3, kSlotI = 5;



// The seven points of a row in SpMV order: W, S, B, the row itself, T, N,
// E -- ascending column order for monotone numbering (i-1 < j-1 < k-1 <
// self < k+1 < j+1 < i+1).  An absent neighbour is the row itself with a
// zero coefficient.  Coefficients are kept in their storage type.
template <typename T>
struct RowPts {
  int q[7];
  T c[7];
};

template <typename T, int L, typename M>
__device__ __forceinline__ void row_pts(const M& S, int row, int64_t n, RowPts<T>& P) {
  int nW, nS, nB, nT, nN, nE;
  bool hW = true, hS = true, hB = true;
  if constexpr (L == ML_INDEXED) {
    nW = S.nbr[DW * n + row]; nS = S.nbr[DS * n + row]; nB = S.nbr[DB * n + row];
    nT = S.nbr[DT * n + row]; nN = S.nbr[DN * n + row]; nE = S.nbr[DE * n + row];
  } else {
    const uint64_t pw = S.pk[row];
    const int dT = (int)(pw & 3u), dB = (int)((pw >> 2) & 3u);
    const int dN = (int)((pw >> 4) & 0x7ffu), dS = (int)((pw >> 15) & 0x7ffu);
    const int dE = (int)((pw >> 26) & 0x7ffffu), dW = (int)((pw >> 45) & 0x7ffffu);
    nT = row + dT; nB = row - dB; nN = row + dN; nS = row - dS; nE = row + dE; nW = row - dW;
    hW = dW != 0; hS = dS != 0; hB = dB != 0;
  }
  P.q[0] = nW; P.q[1] = nS; P.q[2] = nB; P.q[3] = row; P.q[4] = nT; P.q[5] = nN; P.q[6] = nE;
  P.c[3] = S.coef[row];
  P.c[6] = S.coef[(1 + DE) * n + row];
  P.c[5] = S.coef[(1 + DN) * n + row];
  P.c[4] = S.coef[(1 + DT) * n + row];
  if constexpr (L == ML_PACKED_SYM) {
    const T gW = S.coef[(1 + DE) * n + nW];
    const T gS = S.coef[(1 + DN) * n + nS];
    const T gB = S.coef[(1 + DT) * n + nB];
    P.c[0] = hW ? gW : (T)0;
    P.c[1] = hS ? gS : (T)0;
    P.c[2] = hB ? gB : (T)0;
  } else {
    P.c[0] = S.coef[(1 + DW) * n + row];
    P.c[1] = S.coef[(1 + DS) * n + row];
    P.c[2] = S.coef[(1 + DB) * n + row];
  }
}

// sum over the points in order of c_k * y(k), in FP64
template <typename T, typename F>
__device__ __forceinline__ double spmv_pts(const RowPts<T>& P, F&& y) {
  double t = (double)P.c[0] * y(0);
#pragma unroll
  for (int k = 1; k < 7; ++k) t += (double)P.c[k] * y(k);
  return t;
}

template <typename T, int L, typename M, typename F>
__device__ __forceinline__ double stencil_row(const M& S, int row, int64_t n, F&& x) {
  RowPts<T> P;
  row_pts<T, L>(S, row, n, P);
  return spmv_pts(P, [&](int k) { return x(P.q[k]); });
}

// Loads of one row of kernel A (everything that does not depend on the
// iteration's scalars): the points, r, p, v and M^-1 at each point, r0.
template <typename T>
struct AIn {
  RowPts<T> P;
  T r[7], p[7], v[7], m[7];
  T r0;
};
template <typename T, int L, typename M, typename W, typename P>
__device__ __forceinline__ void a_fetch(const M& S, const W& w, P pold, P vold, int row, int64_t n, AIn<T>& f) {
  row_pts<T, L>(S, row, n, f.P);
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const int q = f.P.q[k];
    f.r[k] = w.r[q];
    f.p[k] = pold[q];
    f.v[k] = vold[q];
    f.m[k] = S.invdiag[q];
  }
  f.r0 = w.r0[row];
}

// Loads of one row of kernel B: the points, r, v' and M^-1 at each point.
template <typename T>
struct BIn {
  RowPts<T> P;
  T r[7], v[7], m[7];
};
template <typename T, int L, typename M, typename W, typename P>
__device__ __forceinline__ void b_fetch(const M& S, const W& w, P vn, int row, int64_t n, BIn<T>& f) {
  row_pts<T, L>(S, row, n, f.P);
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const int q = f.P.q[k];
    f.r[k] = w.r[q];
    f.v[k] = vn[q];
    f.m[k] = S.invdiag[q];
  }
}

// Loads of one row of kernel C.
template <typename T>
struct CIn {
  T m, p, r, v, t, x, r0;
};
template <typename T, typename P, typename X, typename W, typename V>
__device__ __forceinline__ void c_fetch(P invdiag, X x, const W& w, V pn, V vn, int64_t row, CIn<T>& f) {
  f.m = invdiag[row];
  f.p = pn[row];
  f.r = w.r[row];
  f.v = vn[row];
  f.t = w.t[row];
  f.x = x[row];
  f.r0 = w.r0[row];
}

// Tile of this block (logical tiles of one XCD are contiguous), -1 past the
// end of the table (the grid is padded to a multiple of the XCD count).
__device__ __forceinline__ int block_tile(int ntiles) {
  const int t = logical_block();
  return t < ntiles ? t : -1;
}
// The same with the tile count read from device memory when the batch says
// so: each XCD takes one contiguous eighth of the tiles, whatever the grid.
template <typename T>
__device__ __forceinline__ int block_tile(const Batch<T>& B) {
  if (!B.ntiles_dev) return block_tile(B.ntiles);
  const int nt = *B.ntiles_dev;
  const int per = (nt + kXcd - 1) / kXcd;
  const int i = (int)blockIdx.x / kXcd;
  if (i >= per) return -1;
  const int t = ((int)blockIdx.x % kXcd) * per + i;
  return t < nt ? t : -1;
}

// Pointers read from a table are generic pointers to the compiler, which
// then emits flat loads and stores (slower, and more registers).  The
// kernels work on views whose pointers are typed global (address space 1).
#define NS3D_GLOBAL __attribute__((address_space(1)))
template <typename T>
struct GMat {
  int n;
  int layout;
  const NS3D_GLOBAL int32_t* nbr;
  const NS3D_GLOBAL uint64_t* pk;
  const NS3D_GLOBAL T* coef;
  const NS3D_GLOBAL T* invdiag;
  const NS3D_GLOBAL int* chk;
};
template <typename T>
struct GWork {
  NS3D_GLOBAL T* r;
  NS3D_GLOBAL T* r0;
  NS3D_GLOBAL T* p[2];
  NS3D_GLOBAL T* v[2];
  NS3D_GLOBAL T* t;
  NS3D_GLOBAL BState* st;
};
template <typename T>
struct GDesc {
  GMat<T> S;
  GWork<T> w;
  NS3D_GLOBAL T* x;
  const NS3D_GLOBAL T* b;
  double tol;
  long long maxit;
  NS3D_GLOBAL double* partials;
  int pstride;
  int t0, t1;
  double eps2;
  // slot q of this solve's partials (tile-local index)
  __device__ __forceinline__ NS3D_GLOBAL double* slot(int q) const { return partials + (int64_t)q * pstride; }
};
template <typename P>
__device__ __forceinline__ NS3D_GLOBAL P* gl(P* p) {
  return (NS3D_GLOBAL P*)p;
}
template <typename T>
__device__ __forceinline__ GDesc<T> to_gdesc(const SolveDesc<T>& D) {
  GDesc<T> G;
  G.S = GMat<T>{D.S.n, D.S.layout, gl(D.S.nbr), gl(D.S.pk), gl(D.S.coef), gl(D.S.invdiag), gl(D.S.chk)};
  G.w.r = gl(D.w.r);
  G.w.r0 = gl(D.w.r0);
  G.w.p[0] = gl(D.w.p[0]);
  G.w.p[1] = gl(D.w.p[1]);
  G.w.v[0] = gl(D.w.v[0]);
  G.w.v[1] = gl(D.w.v[1]);
  G.w.t = gl(D.w.t);
  G.w.st = gl(D.w.st);
  G.x = gl(D.x);
  G.b = gl(D.b);
  G.tol = D.tol;
  G.maxit = D.maxit;
  G.partials = gl(D.partials);
  G.pstride = D.pstride;
  G.t0 = D.t0;
  G.t1 = D.t1;
  G.eps2 = D.eps2;
  return G;
}
template <typename T>
__device__ __forceinline__ GDesc<T> load_desc(const SolveDesc<T>* p) {
  return to_gdesc(*p);
}

// The solve and rows of tile t of a launch, and the iteration the launch
// stands for.  One solve passed by value needs no table loads.
template <typename T>
__device__ __forceinline__ void tile_of(const Batch<T>& B, int t, Tile& tl, GDesc<T>& D, long long& it) {
  if (B.single) {
    D = to_gdesc(B.one);
    tl.solve = 0;
    if (D.S.chk) {
      // plane-band tiles: band t / bx, block t % bx of it
      const int bx = B.ntiles / kBands, band = t / bx;
      tl.step = bx;
      tl.begin = D.S.chk[band] + t % bx;
      tl.end = D.S.chk[band + 1];
    } else {
      const long long n = D.S.n, b = (long long)t * B.per;
      tl.step = 0;
      tl.begin = b < n ? b : n;
      tl.end = b + B.per < n ? b + B.per : n;
    }
    it = B.one_it0 + B.q;
  } else {
    const TileDesc<T>* e = B.tiles + t;
    D = to_gdesc(e->d);
    tl.solve = e->solve;
    tl.step = e->step;
    tl.begin = e->begin;
    tl.end = e->end;
    it = e->it0 + B.q;
  }
}

// The rows of a tile as one thread visits them.  Contiguous tiles (step 0):
// rows begin + t, begin + t + kBlock, ... below end.  Plane-band tiles: the
// chunks begin, begin + step, ... below end of the system's chunk list, each
// {first row, rows}, thread t taking first + t when t < rows.  row() is -1
// where this thread has no row at the current position.
struct TileWalk {
  long long pos, end, step;
  // chunk descriptors as 64-bit words: first row (low half), rows (high half)
  const NS3D_GLOBAL uint64_t* chunks;
  __device__ __forceinline__ bool more() const { return pos < end; }
  __device__ __forceinline__ int64_t row() const {
    if (!step) {
      const long long r = pos + threadIdx.x;
      return r < end ? r : -1;
    }
    const uint64_t c = chunks[pos];
    return threadIdx.x < (unsigned)(c >> 32) ? (int64_t)(uint32_t)c + threadIdx.x : -1;
  }
  __device__ __forceinline__ void next() { pos += step ? step : kBlock; }
};
template <typename T>
__device__ __forceinline__ TileWalk walk_of(const Tile& tl, const GDesc<T>& D) {
  TileWalk w;
  w.pos = tl.begin;
  w.end = tl.end;
  w.step = tl.step;
  w.chunks = tl.step ? (const NS3D_GLOBAL uint64_t*)(D.S.chk + kChunkHeader) : nullptr;
  return w;
}
template <typename T>
__device__ __forceinline__ GDesc<T> solve_of(const Batch<T>& B, int s) {
  return B.single ? to_gdesc(B.one) : load_desc(B.d + s);
}

// Calls body.template operator()<L>() with the solve's layout as a
// compile-time constant (the branch is uniform over the block).
template <typename F>
__device__ __forceinline__ void by_layout(int layout, F&& body) {
  if (layout == ML_PACKED_SYM) body(std::integral_constant<int, ML_PACKED_SYM>{});
  else if (layout == ML_PACKED) body(std::integral_constant<int, ML_PACKED>{});
  else body(std::integral_constant<int, ML_INDEXED>{});
}

// r = b - A x, r0 = r; zero p, v of the first buffer; partials |r|^2, |b|^2.
template <typename T>
__global__ void __launch_bounds__(kBlock)
bicg_init_kernel(Batch<T> B) {
  const int t = block_tile(B.ntiles);
  if (t < 0) return;
  Tile tl;
  GDesc<T> D;
  long long it;
  tile_of(B, t, tl, D, it);
  double v2[2] = {0.0, 0.0};
  by_layout(D.S.layout, [&](auto LC) {
    constexpr int L = decltype(LC)::value;
    const int64_t n = D.S.n;
    for (TileWalk w = walk_of(tl, D); w.more(); w.next()) {
      const int64_t row = w.row();
      if (row < 0) continue;
      const double ax = stencil_row<T, L>(D.S, (int)row, n, [&](int q) { return (double)D.x[q]; });
      const double bv = D.b[row];
      const double r = bv - ax;
      D.w.r[row] = (T)r;
      D.w.r0[row] = (T)r;
      D.w.p[0][row] = (T)0;
      D.w.v[0][row] = (T)0;
      const double rs = (double)(T)r;
      v2[0] += rs * rs;
      v2[1] += bv * bv;
    }
  });
  block_sums_at<2>(v2, D.slot(kSlotI), D.pstride, t - D.t0);
}

// One block per solve.
template <typename T>
__global__ void __launch_bounds__(kBlock)
bicg_init_fin(Batch<T> B) {
  const GDesc<T> D = solve_of(B, blockIdx.x);
  double out[2];
  sum_partials_all<2>(D.slot(kSlotI), D.t1 - D.t0, out, D.pstride);
  if (threadIdx.x != 0) return;
  NS3D_GLOBAL BState* st = D.w.st;
  st->tol = D.tol;
  st->maxit = D.maxit;
  st->eps2 = D.eps2 > 0 ? D.eps2 : kEps * kEps;
  st->r0_sqnorm = out[0];
  st->rhs_sqnorm = out[1];
  st->it_a = 0;
  st->it_c = 0;
  st->ibase = 0;
  st->restarts = 0;
  st->rr = out[0];
  if (out[1] == 0) {
    st->status = BS_ZERO_RHS;
    return;
  }
  st->tol2 = D.tol * D.tol * out[1];
  // the slot of iteration -1 holds Eigen's initial rho = alpha = w = 1
  st->rho[1] = 1;
  st->alpha[1] = 1;
  st->w[1] = 1;
  st->rho_fresh = out[0];  // rho = r0.r = |r0|^2 (no restart possible at i = 0)
  st->fresh_at = 0;
  st->status = (out[0] > st->tol2 && 0 < D.maxit) ? BS_RUNNING : BS_DONE;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
bicg_restart_kernel(Batch<T> B) {
  const int t = block_tile(B.ntiles);
  if (t < 0) return;
  Tile tl;
  GDesc<T> D;
  long long it;
  tile_of(B, t, tl, D, it);
  if (D.w.st->status != BS_RESTART) return;
  double v1[1] = {0.0};
  by_layout(D.S.layout, [&](auto LC) {
    constexpr int L = decltype(LC)::value;
    const int64_t n = D.S.n;
    for (TileWalk w = walk_of(tl, D); w.more(); w.next()) {
      const int64_t row = w.row();
      if (row < 0) continue;
      const double ax = stencil_row<T, L>(D.S, (int)row, n, [&](int q) { return (double)D.x[q]; });
      const double r = (double)D.b[row] - ax;
      D.w.r[row] = (T)r;
      D.w.r0[row] = (T)r;
      const double rs = (double)(T)r;
      v1[0] += rs * rs;
    }
  });
  block_sums_at<1>(v1, D.slot(kSlotI), D.pstride, t - D.t0);
}

// One block per solve.
template <typename T>
__global__ void __launch_bounds__(kBlock)
bicg_restart_fin(Batch<T> B) {
  const GDesc<T> D = solve_of(B, blockIdx.x);
  NS3D_GLOBAL BState* st = D.w.st;
  if (st->status != BS_RESTART) return;
  double out[1];
  sum_partials_all<1>(D.slot(kSlotI), D.t1 - D.t0, out, D.pstride);
  if (threadIdx.x != 0) return;
  // the iteration that found the restart is re-run with the new residual;
  // rho_old stays the pre-restart rho (slot of the previous iteration)
  st->rho_fresh = out[0];
  st->r0_sqnorm = out[0];
  st->fresh_at = st->it_c;
  if (st->restarts++ == 0) st->ibase = st->it_c;
  st->status = BS_RUNNING;
}

// The first row of every thread is fetched before the block sums the
// previous kernel's partials, so the two latencies overlap.  With
// NS3D_SOLVE_PIPE=1 each later row is also fetched one step ahead (the next
// row's loads in flight while the current one is computed); that doubles the
// registers of A and B and measured no faster on MI355X at 128^3, so it is
// off by default.
#ifndef NS3D_SOLVE_PIPE
#define NS3D_SOLVE_PIPE 0
#endif

// The block leader (first thread of the solve's first tile) records the
// scalars for the host.
__device__ __forceinline__ bool leader_of(int t, int t0) { return t == t0 && threadIdx.x == 0; }

// Minimum waves per SIMD asked of the compiler for kernels A and B (it
// then keeps their registers within that occupancy); 0: its own choice.
#ifndef NS3D_A_WAVES
#define NS3D_A_WAVES 0
#endif
#ifndef NS3D_B_WAVES
#define NS3D_B_WAVES 0
#endif
#if NS3D_A_WAVES > 0
#define NS3D_A_ATTR __attribute__((amdgpu_waves_per_eu(NS3D_A_WAVES)))
#else
#define NS3D_A_ATTR
#endif
#if NS3D_B_WAVES > 0
#define NS3D_B_ATTR __attribute__((amdgpu_waves_per_eu(NS3D_B_WAVES)))
#else
#define NS3D_B_ATTR
#endif

template <typename T>
__global__ void __launch_bounds__(kBlock) NS3D_A_ATTR
bicg_A_kernel(Batch<T> B) {
  const int t = block_tile(B);
  if (t < 0) return;
  Tile tl;
  GDesc<T> D;
  long long it;
  tile_of(B, t, tl, D, it);
  NS3D_GLOBAL BState* st = D.w.st;
  // the previous kernel's partials are loaded together with the status;
  // launches queued after the solve stopped exit here, before any row load.
  // The leader may store DONE / RESTART while other blocks of this launch
  // start, so the block takes thread 0's reading: every wave of a block then
  // makes the same decision (and the block reduction below never reads the
  // LDS slot of a wave that left).
  __shared__ int go;
  double pa[2];
  partials_load<2>(D.slot(kSlotC), D.t1 - D.t0, D.pstride, pa);
  if (threadIdx.x == 0) go = st->status == BS_RUNNING && st->it_c == it;
  const bool lead = leader_of(t, D.t0);
  const int sl = (int)(it & 1);
  const int par = sl;
  // (selects, not D.w.p[par]: a run-time index into a local array would
  // put the descriptor on the stack)
  const NS3D_GLOBAL T* __restrict__ pold = par ? D.w.p[1] : D.w.p[0];
  const NS3D_GLOBAL T* __restrict__ vold = par ? D.w.v[1] : D.w.v[0];
  NS3D_GLOBAL T* __restrict__ pnew = par ? D.w.p[0] : D.w.p[1];
  NS3D_GLOBAL T* __restrict__ vnew = par ? D.w.v[0] : D.w.v[1];
  __syncthreads();
  if (!go) return;
  double acc = 0.0;
  bool stop = false;
  by_layout(D.S.layout, [&](auto LC) {
    constexpr int L = decltype(LC)::value;
    const int64_t n = D.S.n;
    TileWalk wk = walk_of(tl, D);
    int64_t row = wk.more() ? wk.row() : -1;
    // first row in flight during the reduction
    AIn<T> cur;
    if (row >= 0) a_fetch<T, L>(D.S, D.w, pold, vold, (int)row, n, cur);
    const long long fresh_at = st->fresh_at;
    const double rho_prev = st->rho[sl ^ 1], alpha_prev = st->alpha[sl ^ 1], w_prev = st->w[sl ^ 1];
    double red[2];
    partials_reduce<2>(pa, red);
    double rho;
    if (fresh_at == it) {
      rho = st->rho_fresh;
    } else {
      const double r0r = red[0], rr2 = red[1];
      if (!(rr2 > st->tol2 && it - st->ibase < st->maxit)) {
        if (lead) { st->rr = rr2; st->status = BS_DONE; }
        stop = true;
        return;
      }
      if (fabs(r0r) < st->eps2 * st->r0_sqnorm) {
        if (lead) { st->rr = rr2; st->status = BS_RESTART; }
        stop = true;
        return;
      }
      rho = r0r;
      if (lead) st->rr = rr2;
    }
    const double beta = (rho / rho_prev) * (alpha_prev / w_prev);
    const double om = w_prev;
    if (lead) {
      st->rho[sl] = rho;
      st->it_a = it;
    }
    while (wk.more()) {
      TileWalk wn = wk;
      wn.next();
      const int64_t nxt = wn.more() ? wn.row() : -1;
      AIn<T> nx;
      if (NS3D_SOLVE_PIPE && nxt >= 0) a_fetch<T, L>(D.S, D.w, pold, vold, (int)nxt, n, nx);
      if (row >= 0) {
        double pn[7];
#pragma unroll
        for (int k = 0; k < 7; ++k)
          pn[k] = (double)(T)((double)cur.r[k] + beta * ((double)cur.p[k] - om * (double)cur.v[k]));
        const double vn = spmv_pts(cur.P, [&](int k) { return (double)cur.m[k] * pn[k]; });
        pnew[row] = (T)pn[3];
        vnew[row] = (T)vn;
        acc += (double)cur.r0 * (double)(T)vn;
      }
      if (!NS3D_SOLVE_PIPE && nxt >= 0) a_fetch<T, L>(D.S, D.w, pold, vold, (int)nxt, n, nx);
      cur = nx;
      row = nxt;
      wk = wn;
    }
  });
  if (stop) return;
  double v1[1] = {acc};
  block_sums_at<1>(v1, D.slot(kSlotA), D.pstride, t - D.t0);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) NS3D_B_ATTR
bicg_B_kernel(Batch<T> B) {
  const int t = block_tile(B);
  if (t < 0) return;
  Tile tl;
  GDesc<T> D;
  long long it;
  tile_of(B, t, tl, D, it);
  NS3D_GLOBAL BState* st = D.w.st;
  const int status = st->status;
  const long long done = st->it_a;
  double pa[1];
  partials_load<1>(D.slot(kSlotA), D.t1 - D.t0, D.pstride, pa);
  const bool lead = leader_of(t, D.t0);
  const int sl = (int)(it & 1);
  const int par = sl ^ 1;
  const NS3D_GLOBAL T* __restrict__ vn = par ? D.w.v[1] : D.w.v[0];
  if (status != BS_RUNNING || done != it) return;
  double a0 = 0.0, a1 = 0.0;
  by_layout(D.S.layout, [&](auto LC) {
    constexpr int L = decltype(LC)::value;
    const int64_t n = D.S.n;
    TileWalk wk = walk_of(tl, D);
    int64_t row = wk.more() ? wk.row() : -1;
    BIn<T> cur;
    if (row >= 0) b_fetch<T, L>(D.S, D.w, vn, (int)row, n, cur);
    const double rho = st->rho[sl];
    double red[1];
    partials_reduce<1>(pa, red);
    const double alpha = rho / red[0];
    if (lead) {
      st->r0v = red[0];
      st->alpha[sl] = alpha;
    }
    while (wk.more()) {
      TileWalk wn = wk;
      wn.next();
      const int64_t nxt = wn.more() ? wn.row() : -1;
      BIn<T> nx;
      if (NS3D_SOLVE_PIPE && nxt >= 0) b_fetch<T, L>(D.S, D.w, vn, (int)nxt, n, nx);
      if (row >= 0) {
        double sv[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) sv[k] = (double)(T)((double)cur.r[k] - alpha * (double)cur.v[k]);
        const double tv = (double)(T)spmv_pts(cur.P, [&](int k) { return (double)cur.m[k] * sv[k]; });
        D.w.t[row] = (T)tv;
        a0 += tv * sv[3];
        a1 += tv * tv;
      }
      if (!NS3D_SOLVE_PIPE && nxt >= 0) b_fetch<T, L>(D.S, D.w, vn, (int)nxt, n, nx);
      cur = nx;
      row = nxt;
      wk = wn;
    }
  });
  double v2[2] = {a0, a1};
  block_sums_at<2>(v2, D.slot(kSlotB), D.pstride, t - D.t0);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
bicg_C_kernel(Batch<T> B) {
  const int t = block_tile(B);
  if (t < 0) return;
  Tile tl;
  GDesc<T> D;
  long long it;
  tile_of(B, t, tl, D, it);
  NS3D_GLOBAL BState* st = D.w.st;
  const int status = st->status;
  const long long done = st->it_a;
  double pa[2];
  partials_load<2>(D.slot(kSlotB), D.t1 - D.t0, D.pstride, pa);
  const bool lead = leader_of(t, D.t0);
  const int sl = (int)(it & 1);
  const int par = sl ^ 1;
  NS3D_GLOBAL T* __restrict__ x = D.x;
  const NS3D_GLOBAL T* __restrict__ invdiag = D.S.invdiag;
  const NS3D_GLOBAL T* __restrict__ pn = par ? D.w.p[1] : D.w.p[0];
  const NS3D_GLOBAL T* __restrict__ vn = par ? D.w.v[1] : D.w.v[0];
  if (status != BS_RUNNING || done != it) return;
  TileWalk wk = walk_of(tl, D);
  int64_t row = wk.more() ? wk.row() : -1;
  CIn<T> cur;
  if (row >= 0) c_fetch<T>(invdiag, x, D.w, pn, vn, row, cur);
  const double alpha = st->alpha[sl];
  double red[2];
  partials_reduce<2>(pa, red);
  const double om = (red[1] > 0.0) ? red[0] / red[1] : 0.0;
  if (lead) {
    st->ts = red[0];
    st->tt = red[1];
    st->w[sl] = om;
    st->it_c = it + 1;
  }
  double a0 = 0.0, a1 = 0.0;
  while (wk.more()) {
    TileWalk wn = wk;
    wn.next();
    const int64_t nxt = wn.more() ? wn.row() : -1;
    CIn<T> nx;
    if (nxt >= 0) c_fetch<T>(invdiag, x, D.w, pn, vn, nxt, nx);
    if (row >= 0) {
      const double md = cur.m;
      const double y = md * (double)cur.p;
      const double s = (double)(T)((double)cur.r - alpha * (double)cur.v);
      const double z = md * s;
      x[row] = (T)((double)cur.x + (alpha * y + om * z));
      const double rn = (double)(T)(s - om * (double)cur.t);
      D.w.r[row] = (T)rn;
      a0 += (double)cur.r0 * rn;
      a1 += rn * rn;
    }
    cur = nx;
    row = nxt;
    wk = wn;
  }
  double v2[2] = {a0, a1};
  block_sums_at<2>(v2, D.slot(kSlotC), D.pstride, t - D.t0);
}

__global__ void __launch_bounds__(kBlock)
invdiag_kernel(SysView S) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= S.n) return;
  const double d = S.coef[row];
  S.invdiag[row] = (d != 0.0) ? 1.0 / d : 1.0;
}

// One word per row: T | B << 2 | N << 4 | S << 15 | E << 26 | W << 45 (bit
// widths kPkTB, kPkNS, kPkEW), each the distance from the row to that
// neighbour (E, N, T above the row, W, S, B below; 0 when absent).
__global__ void __launch_bounds__(kBlock)
pack_neighbours_kernel(const int32_t* __restrict__ nbr, int n, uint64_t* __restrict__ pk, int* bad) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= n) return;
  const int64_t m = n;
  const long long dE = (long long)nbr[DE * m + row] - row, dW = (long long)row - nbr[DW * m + row];
  const long long dN = (long long)nbr[DN * m + row] - row, dS = (long long)row - nbr[DS * m + row];
  const long long dT = (long long)nbr[DT * m + row] - row, dB = (long long)row - nbr[DB * m + row];
  const long long lTB = (1ll << kPkTB), lNS = (1ll << kPkNS), lEW = (1ll << kPkEW);
  const bool ok = dT >= 0 && dT < lTB && dB >= 0 && dB < lTB && dN >= 0 && dN < lNS && dS >= 0 && dS < lNS &&
                  dE >= 0 && dE < lEW && dW >= 0 && dW < lEW;
  if (!ok) {
    atomicOr(bad, 1);
    pk[row] = 0;
    return;
  }
  pk[row] = (uint64_t)dT | ((uint64_t)dB << 2) | ((uint64_t)dN << 4) | ((uint64_t)dS << 15) |
            ((uint64_t)dE << 26) | ((uint64_t)dW << 45);
}

__global__ void __launch_bounds__(kBlock)
coef_to_f32_kernel(const double* __restrict__ coef, const double* __restrict__ invdiag, int64_t n,
                   float* __restrict__ coef32, float* __restrict__ invdiag32) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int q = 0; q < 7; ++q) coef32[q * n + i] = (float)coef[q * n + i];
  invdiag32[i] = (float)invdiag[i];
}

// FP64 residual of the defect correction: r = b - A x, r32 = (float) r.
template <int L, bool TRIAL>
__global__ void __launch_bounds__(kBlock)
residual64_kernel(SMat<double> A, const double* __restrict__ x, const float* __restrict__ d,
                  const double* __restrict__ b, double* __restrict__ r, float* __restrict__ r32, double* partials) {
  const int64_t n = A.n;
  const RowIter rr = block_rows(n);
  double v2[2] = {0.0, 0.0};
  for (int64_t row = rr.begin + threadIdx.x; row < rr.end; row += rr.step) {
    // the trial iterate x + d, as add_f32 would form it
    const double ax = stencil_row<double, L>(A, (int)row, n, [&](int q) {
      if constexpr (TRIAL) return x[q] + 1.0 * (double)d[q];
      else return x[q];
    });
    const double bv = b[row];
    const double rv = bv - ax;
    r[row] = rv;
    r32[row] = (float)rv;
    v2[0] += rv * rv;
    v2[1] += bv * bv;
  }
  const int ops[2] = {RSUM, RSUM};
  block_partials<2>(v2, ops, partials);
}

__global__ void __launch_bounds__(kBlock)
sums_fin(const double* partials, int nb, double* result) {
  const int ops[2] = {RSUM, RSUM};
  double out[2];
  finalize_partials<2>(partials, nb, ops, out);
  if (threadIdx.x == 0) {
    result[0] = out[0];
    result[1] = out[1];
  }
}

__global__ void __launch_bounds__(kBlock)
add_f32_kernel(double* __restrict__ x, const float* __restrict__ d, int64_t n, double a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] += a * (double)d[i];
}

__global__ void __launch_bounds__(kBlock) fill_f32_kernel(float* __restrict__ x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = 0.0f;
}

}  // namespace

// Launch with optional start / stop events taken from the dispatch itself.
template <typename K, typename... Args>
hipError_t launch_timed(K kernel, int nb, hipStream_t s, hipEvent_t* ev, Args... args) {
  if (ev) hipExtLaunchKernelGGL(kernel, dim3(nb), dim3(kBlock), 0, s, ev[0], ev[1], 0, args...);
  else hipLaunchKernelGGL(kernel, dim3(nb), dim3(kBlock), 0, s, args...);
  return hipGetLastError();
}

int tiles_for(int64_t n, bool chunked) {
  int64_t nt = (n + kTileRows - 1) / kTileRows;
  if (nt > kMaxTilesPerSolve) nt = kMaxTilesPerSolve;
  if (nt < 1) nt = 1;
  if (chunked) nt = std::max<int64_t>(kBands, nt / kBands * kBands);
  return (int)nt;
}

void make_tiles(int solve, int64_t n, const int* band_off, std::vector<Tile>& out, int& t0, int& t1) {
  const int nt = tiles_for(n, band_off != nullptr);
  t0 = (int)out.size();
  if (band_off) {
    const int bx = nt / kBands;
    for (int t = 0; t < nt; ++t) {
      const int band = t / bx;
      out.push_back(Tile{solve, bx, (long long)band_off[band] + t % bx, (long long)band_off[band + 1]});
    }
  } else {
    const int64_t per = ((n + nt - 1) / nt + kBlock - 1) / kBlock * kBlock;
    for (int k = 0; k < nt; ++k) {
      const int64_t b = std::min<int64_t>((int64_t)k * per, n);
      const int64_t e = std::min<int64_t>(b + per, n);
      out.push_back(Tile{solve, 0, b, e});
    }
  }
  t1 = (int)out.size();
}

std::vector<int> build_chunks(const int32_t* cell, int64_t n, int ny, int nz, int64_t min_rows) {
  std::vector<int> out;
  if (n < min_rows || n < 1 || ny < kBands) return out;
  // rows of one (plane, band) are contiguous: the rows follow (i, j, k)
  // order and the band is monotone in j
  std::vector<std::vector<int>> band(kBands);
  const int64_t plane = (int64_t)ny * nz;
  int64_t r = 0;
  while (r < n) {
    const int64_t i = cell[r] / plane;
    const int b = (int)(((cell[r] / nz) % ny) * kBands / ny);
    int64_t e = r + 1;
    while (e < n && cell[e] / plane == i && (int)(((cell[e] / nz) % ny) * kBands / ny) == b) ++e;
    for (int64_t c = r; c < e; c += kBlock) {
      band[b].push_back((int)c);
      band[b].push_back((int)std::min<int64_t>(kBlock, e - c));
    }
    r = e;
  }
  out.assign(kChunkHeader, 0);
  int total = 0;
  for (int b = 0; b < kBands; ++b) {
    out[b] = total;
    total += (int)band[b].size() / 2;
  }
  out[kBands] = total;
  for (int b = 0; b < kBands; ++b) out.insert(out.end(), band[b].begin(), band[b].end());
  return out;
}

int tile_grid(int ntiles) { return std::max(kXcd, (ntiles + kXcd - 1) / kXcd * kXcd); }

template <typename T>
Batch<T> single_batch(const SolveDesc<T>& d, long long it0, int q) {
  Batch<T> B{};
  const int nt = tiles_for(d.S.n, d.S.chk != nullptr);
  B.single = 1;
  B.one = d;
  B.one.t0 = 0;
  B.one.t1 = nt;
  B.nsolves = 1;
  B.ntiles = nt;
  B.per = d.S.chk ? 0 : ((d.S.n + nt - 1) / nt + kBlock - 1) / kBlock * kBlock;
  B.one_it0 = it0;
  B.q = q;
  return B;
}

template <typename T>
hipError_t launch_bicg_init(Batch<T> B, hipStream_t s) {
  hipLaunchKernelGGL(bicg_init_kernel<T>, dim3(tile_grid(B.ntiles)), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(bicg_init_fin<T>, dim3(B.nsolves), dim3(kBlock), 0, s, B);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_bicg_restart(Batch<T> B, hipStream_t s) {
  hipLaunchKernelGGL(bicg_restart_kernel<T>, dim3(tile_grid(B.ntiles)), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(bicg_restart_fin<T>, dim3(B.nsolves), dim3(kBlock), 0, s, B);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_bicg_A(Batch<T> B, hipStream_t s, hipEvent_t* ev) {
  return launch_timed(bicg_A_kernel<T>, tile_grid(B.ntiles), s, ev, B);
}

template <typename T>
hipError_t launch_bicg_B(Batch<T> B, hipStream_t s, hipEvent_t* ev) {
  return launch_timed(bicg_B_kernel<T>, tile_grid(B.ntiles), s, ev, B);
}

template <typename T>
hipError_t launch_bicg_C(Batch<T> B, hipStream_t s, hipEvent_t* ev) {
  return launch_timed(bicg_C_kernel<T>, tile_grid(B.ntiles), s, ev, B);
}

template <typename T>
hipError_t graph_add_bicg(hipGraph_t g, int k, const Batch<T>& B, int grid, hipGraphNode_t dep,
                          hipGraphNode_t* out) {
  hipKernelNodeParams prm{};
  prm.func = k == 0 ? (void*)bicg_A_kernel<T> : k == 1 ? (void*)bicg_B_kernel<T> : (void*)bicg_C_kernel<T>;
  prm.gridDim = dim3(grid);
  prm.blockDim = dim3(kBlock);
  prm.sharedMemBytes = 0;
  void* args[1] = {(void*)&B};
  prm.kernelParams = args;
  prm.extra = nullptr;
  return hipGraphAddKernelNode(out, g, dep ? &dep : nullptr, dep ? 1 : 0, &prm);
}

#define NS3D_INSTANTIATE(T)                                                         \
  template hipError_t graph_add_bicg<T>(hipGraph_t, int, const Batch<T>&, int, hipGraphNode_t, hipGraphNode_t*); \
  template Batch<T> single_batch<T>(const SolveDesc<T>&, long long, int);          \
  template hipError_t launch_bicg_init<T>(Batch<T>, hipStream_t);                  \
  template hipError_t launch_bicg_restart<T>(Batch<T>, hipStream_t);               \
  template hipError_t launch_bicg_A<T>(Batch<T>, hipStream_t, hipEvent_t*);        \
  template hipError_t launch_bicg_B<T>(Batch<T>, hipStream_t, hipEvent_t*);        \
  template hipError_t launch_bicg_C<T>(Batch<T>, hipStream_t, hipEvent_t*);
NS3D_INSTANTIATE(double)
NS3D_INSTANTIATE(float)
#undef NS3D_INSTANTIATE

hipError_t launch_invdiag(SysView S, hipStream_t s) {
  if (S.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(invdiag_kernel, dim3((S.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, S);
  return hipGetLastError();
}

hipError_t launch_coef_to_f32(const double* coef, const double* invdiag, int64_t n, float* coef32, float* invdiag32,
                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(coef_to_f32_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, coef,
                     invdiag, n, coef32, invdiag32);
  return hipGetLastError();
}

hipError_t launch_pack_neighbours(const int32_t* nbr, int n, uint64_t* pk, int* bad, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(pack_neighbours_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nbr, n, pk, bad);
  return hipGetLastError();
}

hipError_t launch_residual64(SMat<double> A, const double* x, const double* b, double* r, float* r32, double* partials,
                             double* result, hipStream_t s, const float* d) {
  const int nb = reduce_blocks(A.n);
  auto pick = [&](auto trial) {
    constexpr bool T = decltype(trial)::value;
    return A.layout == ML_PACKED_SYM ? residual64_kernel<ML_PACKED_SYM, T>
         : A.layout == ML_PACKED     ? residual64_kernel<ML_PACKED, T>
                                     : residual64_kernel<ML_INDEXED, T>;
  };
  auto k = d ? pick(std::true_type{}) : pick(std::false_type{});
  hipLaunchKernelGGL(k, dim3(nb), dim3(kBlock), 0, s, A, x, d, b, r, r32, partials);
  hipLaunchKernelGGL(sums_fin, dim3(1), dim3(kBlock), 0, s, partials, nb, result);
  return hipGetLastError();
}

hipError_t launch_add_f32(double* x, const float* d, int64_t n, hipStream_t s, double a) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(add_f32_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, x, d, n, a);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float* x, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, x, n);
  return hipGetLastError();
}

}  // namespace ns3d
