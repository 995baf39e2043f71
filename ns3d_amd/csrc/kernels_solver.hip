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
// One iteration is three kernels, each ending in one grid-wide reduction:
//   A: p' = r + beta (p - w v) and y = M^-1 p' are recomputed on the fly at
//      the row and its six neighbours, v' = A y, partial r0.v'
//   B: s = r - alpha v' and z = M^-1 s on the fly, t = A z, partials t.s, t.t
//   C: x += alpha M^-1 p' + w M^-1 s, r = s - w t, partials r0.r, r.r
// p and v are double-buffered (buffer parity = executed iterations), so the
// recomputation at neighbours never reads a value written in the same kernel.
// All scalars live in device memory (BState); the last block of each
// reduction advances them, so the host only polls the status between chunks
// of queued iterations.  A kernel launched after the solve has stopped exits
// at once.
#include "kernels.h"

namespace ns3d {

namespace {

constexpr double kEps = 2.220446049250313e-16;  // NumTraits<double>::epsilon()

// SpMV row in ascending-column order for monotone numbering:
// W, S, B, diagonal, T, N, E (i-1 < j-1 < k-1 < self < k+1 < j+1 < i+1).
template <typename F>
__device__ __forceinline__ double stencil_row(const SysView& S, int row, int64_t n, F&& y) {
  const int nW = S.nbr[DW * n + row], nS = S.nbr[DS * n + row], nB = S.nbr[DB * n + row];
  const int nT = S.nbr[DT * n + row], nN = S.nbr[DN * n + row], nE = S.nbr[DE * n + row];
  const double cC = S.coef[row];
  const double cE = S.coef[(1 + DE) * n + row], cW = S.coef[(1 + DW) * n + row];
  const double cN = S.coef[(1 + DN) * n + row], cS = S.coef[(1 + DS) * n + row];
  const double cT = S.coef[(1 + DT) * n + row], cB = S.coef[(1 + DB) * n + row];
  double t = cW * y(nW);
  t += cS * y(nS);
  t += cB * y(nB);
  t += cC * y(row);
  t += cT * y(nT);
  t += cN * y(nN);
  t += cE * y(nE);
  return t;
}

// After |r|^2 and r0.r of the new residual are known: loop test, then
// rho_old = rho, rho = r0.r and the restart test (thread 0 of the last block).
__device__ __forceinline__ void loop_head(BState* st, double rr, double r0r) {
  st->rr = rr;
  if (!(rr > st->tol2 && st->i < st->maxit)) {
    st->status = BS_DONE;
    return;
  }
  st->rho_old = st->rho;
  st->rho = r0r;
  if (fabs(st->rho) < kEps * kEps * st->r0_sqnorm) st->status = BS_RESTART;
}

__global__ void __launch_bounds__(kBlock)
bicg_init_kernel(SysView S, const double* __restrict__ x, const double* __restrict__ b, BWork w, double tol,
                 long long maxit) {
  const int64_t n = S.n;
  const int row = logical_block() * kBlock + threadIdx.x;
  double v2[2] = {0.0, 0.0};
  if (row < n) {
    const double ax = stencil_row(S, row, n, [&](int q) { return x[q]; });
    const double r = b[row] - ax;
    w.r[row] = r;
    w.r0[row] = r;
    w.p[0][row] = 0.0;
    w.v[0][row] = 0.0;
    v2[0] = r * r;
    v2[1] = b[row] * b[row];
  }
  const int ops[2] = {RSUM, RSUM};
  double out[2];
  if (grid_reduce<2>(v2, ops, w.partials, &w.st->counter[0], out) && threadIdx.x == 0) {
    BState* st = w.st;
    st->tol = tol;
    st->maxit = maxit;
    st->r0_sqnorm = out[0];
    st->rhs_sqnorm = out[1];
    st->i = 0;
    st->nexec = 0;
    st->restarts = 0;
    st->rr = out[0];
    if (out[1] == 0) {
      st->status = BS_ZERO_RHS;
      return;
    }
    st->status = BS_RUNNING;
    st->tol2 = tol * tol * out[1];
    st->rho = 1;
    st->alpha = 1;
    st->w = 1;
    loop_head(st, out[0], out[0]);
  }
}

__global__ void __launch_bounds__(kBlock)
bicg_restart_kernel(SysView S, const double* __restrict__ x, const double* __restrict__ b, BWork w) {
  if (w.st->status != BS_RESTART) return;
  const int64_t n = S.n;
  const int row = logical_block() * kBlock + threadIdx.x;
  double v1[1] = {0.0};
  if (row < n) {
    const double ax = stencil_row(S, row, n, [&](int q) { return x[q]; });
    const double r = b[row] - ax;
    w.r[row] = r;
    w.r0[row] = r;
    v1[0] = r * r;
  }
  const int ops[1] = {RSUM};
  double out[1];
  if (grid_reduce<1>(v1, ops, w.partials, &w.st->counter[1], out) && threadIdx.x == 0) {
    BState* st = w.st;
    st->rho = out[0];
    st->r0_sqnorm = out[0];
    if (st->restarts++ == 0) st->i = 0;
    st->status = BS_RUNNING;
  }
}

__global__ void __launch_bounds__(kBlock)
bicg_A_kernel(SysView S, BWork w) {
  const BState* st = w.st;
  if (st->status != BS_RUNNING) return;
  const double beta = (st->rho / st->rho_old) * (st->alpha / st->w);
  const double om = st->w;
  const int par = (int)(st->nexec & 1);
  const double* __restrict__ r = w.r;
  const double* __restrict__ pold = w.p[par];
  const double* __restrict__ vold = w.v[par];
  const double* __restrict__ md = S.invdiag;
  const int64_t n = S.n;
  const int row = logical_block() * kBlock + threadIdx.x;
  double v1[1] = {0.0};
  if (row < n) {
    auto y = [&](int q) { return md[q] * (r[q] + beta * (pold[q] - om * vold[q])); };
    const double vn = stencil_row(S, row, n, y);
    w.p[par ^ 1][row] = r[row] + beta * (pold[row] - om * vold[row]);
    w.v[par ^ 1][row] = vn;
    v1[0] = w.r0[row] * vn;
  }
  const int ops[1] = {RSUM};
  double out[1];
  if (grid_reduce<1>(v1, ops, w.partials, &w.st->counter[2], out) && threadIdx.x == 0) {
    w.st->r0v = out[0];
    w.st->alpha = w.st->rho / out[0];
  }
}

__global__ void __launch_bounds__(kBlock)
bicg_B_kernel(SysView S, BWork w) {
  const BState* st = w.st;
  if (st->status != BS_RUNNING) return;
  const double alpha = st->alpha;
  const int par = (int)(st->nexec & 1) ^ 1;
  const double* __restrict__ r = w.r;
  const double* __restrict__ vn = w.v[par];
  const double* __restrict__ md = S.invdiag;
  const int64_t n = S.n;
  const int row = logical_block() * kBlock + threadIdx.x;
  double v2[2] = {0.0, 0.0};
  if (row < n) {
    auto z = [&](int q) { return md[q] * (r[q] - alpha * vn[q]); };
    const double tv = stencil_row(S, row, n, z);
    const double s = r[row] - alpha * vn[row];
    w.t[row] = tv;
    v2[0] = tv * s;
    v2[1] = tv * tv;
  }
  const int ops[2] = {RSUM, RSUM};
  double out[2];
  if (grid_reduce<2>(v2, ops, w.partials, &w.st->counter[3], out) && threadIdx.x == 0) {
    w.st->ts = out[0];
    w.st->tt = out[1];
    w.st->w = (out[1] > 0.0) ? out[0] / out[1] : 0.0;
  }
}

__global__ void __launch_bounds__(kBlock)
bicg_C_kernel(SysView S, double* __restrict__ x, BWork w) {
  const BState* st = w.st;
  if (st->status != BS_RUNNING) return;
  const double alpha = st->alpha, om = st->w;
  const int par = (int)(st->nexec & 1) ^ 1;
  const int64_t n = S.n;
  const int row = logical_block() * kBlock + threadIdx.x;
  double v2[2] = {0.0, 0.0};
  if (row < n) {
    const double md = S.invdiag[row];
    const double y = md * w.p[par][row];
    const double s = w.r[row] - alpha * w.v[par][row];
    const double z = md * s;
    x[row] += alpha * y + om * z;
    const double rn = s - om * w.t[row];
    w.r[row] = rn;
    v2[0] = w.r0[row] * rn;
    v2[1] = rn * rn;
  }
  const int ops[2] = {RSUM, RSUM};
  double out[2];
  if (grid_reduce<2>(v2, ops, w.partials, &w.st->counter[4], out) && threadIdx.x == 0) {
    BState* s2 = w.st;
    s2->i += 1;
    s2->nexec += 1;
    loop_head(s2, out[1], out[0]);
  }
}

__global__ void __launch_bounds__(kBlock)
invdiag_kernel(SysView S) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= S.n) return;
  const double d = S.coef[row];
  S.invdiag[row] = (d != 0.0) ? 1.0 / d : 1.0;
}

}  // namespace

hipError_t launch_bicg_init(SysView S, const double* x, const double* b, BWork w, double tol, long long maxit,
                            hipStream_t s) {
  hipLaunchKernelGGL(bicg_init_kernel, dim3(padded_blocks(S.n)), dim3(kBlock), 0, s, S, x, b, w, tol, maxit);
  return hipGetLastError();
}

hipError_t launch_bicg_restart(SysView S, const double* x, const double* b, BWork w, hipStream_t s) {
  hipLaunchKernelGGL(bicg_restart_kernel, dim3(padded_blocks(S.n)), dim3(kBlock), 0, s, S, x, b, w);
  return hipGetLastError();
}

hipError_t launch_bicg_A(SysView S, BWork w, hipStream_t s) {
  hipLaunchKernelGGL(bicg_A_kernel, dim3(padded_blocks(S.n)), dim3(kBlock), 0, s, S, w);
  return hipGetLastError();
}

hipError_t launch_bicg_B(SysView S, BWork w, hipStream_t s) {
  hipLaunchKernelGGL(bicg_B_kernel, dim3(padded_blocks(S.n)), dim3(kBlock), 0, s, S, w);
  return hipGetLastError();
}

hipError_t launch_bicg_C(SysView S, double* x, BWork w, hipStream_t s) {
  hipLaunchKernelGGL(bicg_C_kernel, dim3(padded_blocks(S.n)), dim3(kBlock), 0, s, S, x, w);
  return hipGetLastError();
}

hipError_t launch_invdiag(SysView S, hipStream_t s) {
  if (S.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(invdiag_kernel, dim3((S.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, S);
  return hipGetLastError();
}

}  // namespace ns3d
