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
// Vectors and coefficients are stored as T (double for the plain path, float
// for the inner solves of the mixed form); every operation is evaluated in
// FP64 registers and reductions accumulate in FP64, so the FP64
// instantiation computes exactly the expressions above.
// All scalars live in device memory (BState).  Each row kernel writes block
// partials and is followed by a one-block finalize kernel that reduces them
// and advances the scalars, so the host only polls the status between chunks
// of queued iterations.  Kernels launched after the solve has stopped exit at
// once.
#include "kernels.h"

namespace ns3d {

namespace {

constexpr double kEps = 2.220446049250313e-16;  // NumTraits<double>::epsilon()

// SpMV row in ascending-column order for monotone numbering:
// W, S, B, diagonal, T, N, E (i-1 < j-1 < k-1 < self < k+1 < j+1 < i+1).
template <typename T, typename F>
__device__ __forceinline__ double stencil_row(const SMat<T>& S, int row, int64_t n, F&& y) {
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
// rho_old = rho, rho = r0.r and the restart test.
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

// r = b - A x, r0 = r; zero p, v of the first buffer; partials |r|^2, |b|^2.
template <typename T>
__global__ void __launch_bounds__(kBlock)
bicg_init_kernel(SMat<T> S, const T* __restrict__ x, const T* __restrict__ b, SWork<T> w) {
  const int64_t n = S.n;
  const RowRange rr = block_rows(n);
  double v2[2] = {0.0, 0.0};
  for (int64_t row = rr.begin + threadIdx.x; row < rr.end; row += kBlock) {
    const double ax = stencil_row(S, (int)row, n, [&](int q) { return (double)x[q]; });
    const double bv = b[row];
    const double r = bv - ax;
    w.r[row] = (T)r;
    w.r0[row] = (T)r;
    w.p[0][row] = (T)0;
    w.v[0][row] = (T)0;
    const double rs = (double)(T)r;
    v2[0] += rs * rs;
    v2[1] += bv * bv;
  }
  const int ops[2] = {RSUM, RSUM};
  block_partials<2>(v2, ops, w.partials);
}

__global__ void __launch_bounds__(kBlock)
bicg_init_fin(double* partials, BState* st, int nb, double tol, long long maxit) {
  const int ops[2] = {RSUM, RSUM};
  double out[2];
  finalize_partials<2>(partials, nb, ops, out);
  if (threadIdx.x != 0) return;
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

template <typename T>
__global__ void __launch_bounds__(kBlock)
bicg_restart_kernel(SMat<T> S, const T* __restrict__ x, const T* __restrict__ b, SWork<T> w) {
  if (w.st->status != BS_RESTART) return;
  const int64_t n = S.n;
  const RowRange rr = block_rows(n);
  double v1[1] = {0.0};
  for (int64_t row = rr.begin + threadIdx.x; row < rr.end; row += kBlock) {
    const double ax = stencil_row(S, (int)row, n, [&](int q) { return (double)x[q]; });
    const double r = (double)b[row] - ax;
    w.r[row] = (T)r;
    w.r0[row] = (T)r;
    const double rs = (double)(T)r;
    v1[0] += rs * rs;
  }
  const int ops[1] = {RSUM};
  block_partials<1>(v1, ops, w.partials);
}

__global__ void __launch_bounds__(kBlock)
bicg_restart_fin(double* partials, BState* st, int nb) {
  if (st->status != BS_RESTART) return;
  const int ops[1] = {RSUM};
  double out[1];
  finalize_partials<1>(partials, nb, ops, out);
  if (threadIdx.x != 0) return;
  st->rho = out[0];
  st->r0_sqnorm = out[0];
  if (st->restarts++ == 0) st->i = 0;
  st->status = BS_RUNNING;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
bicg_A_kernel(SMat<T> S, SWork<T> w) {
  const BState* st = w.st;
  if (st->status != BS_RUNNING) return;
  const double beta = (st->rho / st->rho_old) * (st->alpha / st->w);
  const double om = st->w;
  const int par = (int)(st->nexec & 1);
  const T* __restrict__ r = w.r;
  const T* __restrict__ pold = w.p[par];
  const T* __restrict__ vold = w.v[par];
  T* __restrict__ pnew = w.p[par ^ 1];
  T* __restrict__ vnew = w.v[par ^ 1];
  const T* __restrict__ md = S.invdiag;
  const int64_t n = S.n;
  const RowRange rr = block_rows(n);
  double v1[1] = {0.0};
  for (int64_t row = rr.begin + threadIdx.x; row < rr.end; row += kBlock) {
    auto pn = [&](int q) {
      return (double)(T)((double)r[q] + beta * ((double)pold[q] - om * (double)vold[q]));
    };
    auto y = [&](int q) { return (double)md[q] * pn(q); };
    const double vn = stencil_row(S, (int)row, n, y);
    pnew[row] = (T)pn((int)row);
    vnew[row] = (T)vn;
    v1[0] += (double)w.r0[row] * (double)(T)vn;
  }
  const int ops[1] = {RSUM};
  block_partials<1>(v1, ops, w.partials);
}

__global__ void __launch_bounds__(kBlock)
bicg_A_fin(double* partials, BState* st, int nb) {
  if (st->status != BS_RUNNING) return;
  const int ops[1] = {RSUM};
  double out[1];
  finalize_partials<1>(partials, nb, ops, out);
  if (threadIdx.x != 0) return;
  st->r0v = out[0];
  st->alpha = st->rho / out[0];
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
bicg_B_kernel(SMat<T> S, SWork<T> w) {
  const BState* st = w.st;
  if (st->status != BS_RUNNING) return;
  const double alpha = st->alpha;
  const int par = (int)(st->nexec & 1) ^ 1;
  const T* __restrict__ r = w.r;
  const T* __restrict__ vn = w.v[par];
  const T* __restrict__ md = S.invdiag;
  const int64_t n = S.n;
  const RowRange rr = block_rows(n);
  double v2[2] = {0.0, 0.0};
  for (int64_t row = rr.begin + threadIdx.x; row < rr.end; row += kBlock) {
    auto sv = [&](int q) { return (double)(T)((double)r[q] - alpha * (double)vn[q]); };
    auto z = [&](int q) { return (double)md[q] * sv(q); };
    const double tv = (double)(T)stencil_row(S, (int)row, n, z);
    const double s = sv((int)row);
    w.t[row] = (T)tv;
    v2[0] += tv * s;
    v2[1] += tv * tv;
  }
  const int ops[2] = {RSUM, RSUM};
  block_partials<2>(v2, ops, w.partials);
}

__global__ void __launch_bounds__(kBlock)
bicg_B_fin(double* partials, BState* st, int nb) {
  if (st->status != BS_RUNNING) return;
  const int ops[2] = {RSUM, RSUM};
  double out[2];
  finalize_partials<2>(partials, nb, ops, out);
  if (threadIdx.x != 0) return;
  st->ts = out[0];
  st->tt = out[1];
  st->w = (out[1] > 0.0) ? out[0] / out[1] : 0.0;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
bicg_C_kernel(SMat<T> S, T* __restrict__ x, SWork<T> w) {
  const BState* st = w.st;
  if (st->status != BS_RUNNING) return;
  const double alpha = st->alpha, om = st->w;
  const int par = (int)(st->nexec & 1) ^ 1;
  const int64_t n = S.n;
  const RowRange rr = block_rows(n);
  double v2[2] = {0.0, 0.0};
  for (int64_t row = rr.begin + threadIdx.x; row < rr.end; row += kBlock) {
    const double md = S.invdiag[row];
    const double y = md * (double)w.p[par][row];
    const double s = (double)(T)((double)w.r[row] - alpha * (double)w.v[par][row]);
    const double z = md * s;
    x[row] = (T)((double)x[row] + (alpha * y + om * z));
    const double rn = (double)(T)(s - om * (double)w.t[row]);
    w.r[row] = (T)rn;
    v2[0] += (double)w.r0[row] * rn;
    v2[1] += rn * rn;
  }
  const int ops[2] = {RSUM, RSUM};
  block_partials<2>(v2, ops, w.partials);
}

__global__ void __launch_bounds__(kBlock)
bicg_C_fin(double* partials, BState* st, int nb) {
  if (st->status != BS_RUNNING) return;
  const int ops[2] = {RSUM, RSUM};
  double out[2];
  finalize_partials<2>(partials, nb, ops, out);
  if (threadIdx.x != 0) return;
  st->i += 1;
  st->nexec += 1;
  loop_head(st, out[1], out[0]);
}

__global__ void __launch_bounds__(kBlock)
invdiag_kernel(SysView S) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= S.n) return;
  const double d = S.coef[row];
  S.invdiag[row] = (d != 0.0) ? 1.0 / d : 1.0;
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
__global__ void __launch_bounds__(kBlock)
residual64_kernel(SMat<double> A, const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ r,
                  float* __restrict__ r32, double* partials) {
  const int64_t n = A.n;
  const RowRange rr = block_rows(n);
  double v2[2] = {0.0, 0.0};
  for (int64_t row = rr.begin + threadIdx.x; row < rr.end; row += kBlock) {
    const double ax = stencil_row(A, (int)row, n, [&](int q) { return x[q]; });
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
add_f32_kernel(double* __restrict__ x, const float* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] += (double)d[i];
}

__global__ void __launch_bounds__(kBlock) fill_f32_kernel(float* __restrict__ x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = 0.0f;
}

inline hipError_t record(hipEvent_t e, hipStream_t s) { return e ? hipEventRecord(e, s) : hipSuccess; }

}  // namespace

template <typename T>
hipError_t launch_bicg_init(SMat<T> S, const T* x, const T* b, SWork<T> w, double tol, long long maxit,
                            hipStream_t s) {
  const int nb = reduce_blocks(S.n);
  hipLaunchKernelGGL(bicg_init_kernel<T>, dim3(nb), dim3(kBlock), 0, s, S, x, b, w);
  hipLaunchKernelGGL(bicg_init_fin, dim3(1), dim3(kBlock), 0, s, w.partials, w.st, nb, tol, maxit);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_bicg_restart(SMat<T> S, const T* x, const T* b, SWork<T> w, hipStream_t s) {
  const int nb = reduce_blocks(S.n);
  hipLaunchKernelGGL(bicg_restart_kernel<T>, dim3(nb), dim3(kBlock), 0, s, S, x, b, w);
  hipLaunchKernelGGL(bicg_restart_fin, dim3(1), dim3(kBlock), 0, s, w.partials, w.st, nb);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_bicg_A(SMat<T> S, SWork<T> w, hipStream_t s, hipEvent_t mid) {
  const int nb = reduce_blocks(S.n);
  hipLaunchKernelGGL(bicg_A_kernel<T>, dim3(nb), dim3(kBlock), 0, s, S, w);
  hipError_t e = record(mid, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(bicg_A_fin, dim3(1), dim3(kBlock), 0, s, w.partials, w.st, nb);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_bicg_B(SMat<T> S, SWork<T> w, hipStream_t s, hipEvent_t mid) {
  const int nb = reduce_blocks(S.n);
  hipLaunchKernelGGL(bicg_B_kernel<T>, dim3(nb), dim3(kBlock), 0, s, S, w);
  hipError_t e = record(mid, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(bicg_B_fin, dim3(1), dim3(kBlock), 0, s, w.partials, w.st, nb);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_bicg_C(SMat<T> S, T* x, SWork<T> w, hipStream_t s, hipEvent_t mid) {
  const int nb = reduce_blocks(S.n);
  hipLaunchKernelGGL(bicg_C_kernel<T>, dim3(nb), dim3(kBlock), 0, s, S, x, w);
  hipError_t e = record(mid, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(bicg_C_fin, dim3(1), dim3(kBlock), 0, s, w.partials, w.st, nb);
  return hipGetLastError();
}

#define NS3D_INSTANTIATE(T)                                                                                 \
  template hipError_t launch_bicg_init<T>(SMat<T>, const T*, const T*, SWork<T>, double, long long,       \
                                          hipStream_t);                                                     \
  template hipError_t launch_bicg_restart<T>(SMat<T>, const T*, const T*, SWork<T>, hipStream_t);           \
  template hipError_t launch_bicg_A<T>(SMat<T>, SWork<T>, hipStream_t, hipEvent_t);                        \
  template hipError_t launch_bicg_B<T>(SMat<T>, SWork<T>, hipStream_t, hipEvent_t);                        \
  template hipError_t launch_bicg_C<T>(SMat<T>, T*, SWork<T>, hipStream_t, hipEvent_t);
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

hipError_t launch_residual64(SMat<double> A, const double* x, const double* b, double* r, float* r32, double* partials,
                             double* result, hipStream_t s) {
  const int nb = reduce_blocks(A.n);
  hipLaunchKernelGGL(residual64_kernel, dim3(nb), dim3(kBlock), 0, s, A, x, b, r, r32, partials);
  hipLaunchKernelGGL(sums_fin, dim3(1), dim3(kBlock), 0, s, partials, nb, result);
  return hipGetLastError();
}

hipError_t launch_add_f32(double* x, const float* d, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(add_f32_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, x, d, n);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float* x, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, x, n);
  return hipGetLastError();
}

}  // namespace ns3d
