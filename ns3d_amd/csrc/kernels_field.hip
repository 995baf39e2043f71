// kernels_field.hip -- velocity/pressure corrections, divergence diagnostics,
// the outlet-flux reduction behind the Darcy permeability, and small vector
// kernels.  Reductions are deterministic (fixed block order) and FP64.
#include <algorithm>

#include "kernels.h"

namespace ns3d {

namespace {

__global__ void __launch_bounds__(kBlock) fill_kernel(double* x, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = v;
}

// *bad = 1 if any idx[i] (i < count) is outside [0, n).  Every writer stores
// the same value, so the flag needs no atomic.
__global__ void __launch_bounds__(kBlock) check_index_kernel(const int32_t* idx, int64_t count, int n, double* bad) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < count && (idx[i] < 0 || idx[i] >= n)) *bad = 1.0;
}

// Initpvector, linear profile (NSBase.cxx:2671-2681).
__global__ void __launch_bounds__(kBlock)
init_p_kernel(Geo g, const int32_t* __restrict__ pcell, int np, double* __restrict__ p, double p0, double p1) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= np) return;
  int i, j, k;
  g.decode(pcell[row], i, j, k);
  p[row] = (p1 - p0) * ((double)i + 1.0) / ((double)g.nx) + p0;
}

// Update clamp of ComputeVelocityField (NSBase.cxx:1804-1818).
__global__ void __launch_bounds__(kBlock)
clamp_kernel(double* __restrict__ x, const double* __restrict__ temp, int64_t n, double dvmax) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double delta = x[i] - temp[i];
  if (fabs(delta) > dvmax) {
    if (delta > 0) x[i] = temp[i] + dvmax;
    else if (delta < 0) x[i] = temp[i] - dvmax;
  }
}

__device__ __forceinline__ double clamp_dv(double d, double m) {
  if (d < 0 && fabs(d) > m) d = -m;
  else if (d > 0 && d > m) d = m;
  return d;
}

// CorrectVelocityFieldSIMPLE (kind 0, NSBase.cxx:2082-2172), SIMPLEC (kind 1,
// :2174-2263) and SIMPLER (kind 2, :2265-2335).  One thread per pressure
// node; each face velocity is written by exactly one node (the node on the
// face's + side, or the boundary node for inlet/outlet faces).
__global__ void __launch_bounds__(kBlock)
correct_kernel(Geo g, SysView P, SysView U, SysView V, SysView W, int kind, const double* __restrict__ q,
               double* __restrict__ tu, double* __restrict__ tv, double* __restrict__ tw,
               const double* __restrict__ uh, const double* __restrict__ vh, const double* __restrict__ wh,
               double dvmax, double p_0, double p_1, double dx, double dy, double dz) {
  const int C = logical_block() * kBlock + threadIdx.x;
  if (C >= P.n) return;
  int i, j, k;
  g.decode(P.cell[C], i, j, k);
  const int64_t n = P.n;
  const bool rs = (kind == 1);
  const double* au = rs ? U.rowsum : U.coef;
  const double* av = rs ? V.rowsum : V.coef;
  const double* aw = rs ? W.rowsum : W.coef;
  int o, f;
  if (i > 0) {
    o = P.nbr[DW * n + C];
    if (o != C && (f = g.U(i, j, k)) > -1) {
      if (kind == 2) tu[f] = uh[f] - (q[o] - q[C]) / (U.coef[f] * dx);
      else tu[f] += clamp_dv(-(q[o] - q[C]) / ((au[f]) * dx), dvmax);
    }
  }
  if (j > 0) {
    o = P.nbr[DS * n + C];
    if (o != C && (f = g.V(i, j, k)) > -1) {
      if (kind == 2) tv[f] = vh[f] - (q[o] - q[C]) / (V.coef[f] * dy);
      else tv[f] += clamp_dv(-(q[o] - q[C]) / ((av[f]) * dy), dvmax);
    }
  }
  if (k > 0) {
    o = P.nbr[DB * n + C];
    if (o != C && (f = g.W(i, j, k)) > -1) {
      if (kind == 2) tw[f] = wh[f] - (q[o] - q[C]) / (W.coef[f] * dz);
      else tw[f] += clamp_dv(-(q[o] - q[C]) / ((aw[f]) * dz), dvmax);
    }
  }
  if (i == 0 && (f = g.U(i, j, k)) > -1) {
    if (kind == 2) tu[f] = uh[f] - (p_0 - q[C]) / (U.coef[f] * dx);
    else tu[f] += clamp_dv((q[C]) / ((au[f]) * dx), dvmax);
  }
  if (i == g.nx - 1 && (f = g.U(i + 1, j, k)) > -1) {
    if (kind == 2) tu[f] = uh[f] - (q[C] - p_1) / (U.coef[f] * dx);
    else tu[f] += clamp_dv(-(q[C]) / ((au[f]) * dx), dvmax);
  }
}

// CorrectPressureField (NSBase.cxx:2367-2380): max, min and squared norm of
// pc, then p += beta*pc.  result: max, min, sum of squares.
__global__ void __launch_bounds__(kBlock)
correct_p_kernel(double* __restrict__ p, const double* __restrict__ pc, int64_t n, double beta, double* partials) {
  const RowIter rr = block_rows(n);
  double v[3] = {-__builtin_huge_val(), __builtin_huge_val(), 0.0};
  for (int64_t i = rr.begin + threadIdx.x; i < rr.end; i += rr.step) {
    const double x = pc[i];
    v[0] = x > v[0] ? x : v[0];
    v[1] = x < v[1] ? x : v[1];
    v[2] += x * x;
    p[i] += beta * x;
  }
  const int ops[3] = {RMAX, RMIN, RSUM};
  block_partials<3>(v, ops, partials);
}

__global__ void __launch_bounds__(kBlock) correct_p_fin(const double* partials, int nb, double* result) {
  const int ops[3] = {RMAX, RMIN, RSUM};
  double out[3];
  finalize_partials<3>(partials, nb, ops, out);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) result[q] = out[q];
}

// ComputeDiv (NSBase.cxx:1958-2080).  result: max |sum of outward face
// velocities|, sum div^2, sum |cell velocity|, max |div|.
__global__ void __launch_bounds__(kBlock)
div_kernel(Geo g, SysView P, const double* __restrict__ u, const double* __restrict__ v,
           const double* __restrict__ w, double dx, double dy, double dz, double* partials) {
  const int64_t n = P.n;
  const RowIter rr = block_rows(n);
  double red[4] = {0.0, 0.0, 0.0, 0.0};
  const double idx = 1 / dx, idy = 1 / dy, idz = 1 / dz;
  for (int64_t C64 = rr.begin + threadIdx.x; C64 < rr.end; C64 += rr.step) {
    const int C = (int)C64;
    int i, j, k;
    g.decode(P.cell[C], i, j, k);
    double div = 0, uu = 0, vv = 0, ww = 0, DV = 0;
    int f;
    if (i > 0 && P.nbr[DW * n + C] != C && (f = g.U(i, j, k)) > -1) { div += -u[f] * idx; DV += -u[f]; uu += u[f] * 0.5; }
    if (i < g.nx - 1 && P.nbr[DE * n + C] != C && (f = g.U(i + 1, j, k)) > -1) { div += u[f] * idx; DV += u[f]; uu += u[f] * 0.5; }
    if (j > 0 && P.nbr[DS * n + C] != C && (f = g.V(i, j, k)) > -1) { div += -v[f] * idy; DV += -v[f]; vv += v[f] * 0.5; }
    if (j < g.ny - 1 && P.nbr[DN * n + C] != C && (f = g.V(i, j + 1, k)) > -1) { div += v[f] * idy; DV += v[f]; vv += v[f] * 0.5; }
    if (k > 0 && P.nbr[DB * n + C] != C && (f = g.W(i, j, k)) > -1) { div += -w[f] * idz; DV += -w[f]; ww += w[f] * 0.5; }
    if (k < g.nz - 1 && P.nbr[DT * n + C] != C && (f = g.W(i, j, k + 1)) > -1) { div += w[f] * idz; DV += w[f]; ww += w[f] * 0.5; }
    if (i == 0 && (f = g.U(i, j, k)) > -1) { div += -u[f] * idx; DV += -u[f]; uu += u[f] * 0.5; }
    if (i == g.nx - 1 && (f = g.U(i + 1, j, k)) > -1) { div += u[f] * idx; DV += u[f]; uu += u[f] * 0.5; }
    DV = fabs(DV);
    div = fabs(div);
    red[0] = DV > red[0] ? DV : red[0];
    red[1] += div * div;
    red[2] += sqrt(uu * uu + vv * vv + ww * ww);
    red[3] = div > red[3] ? div : red[3];
  }
  const int ops[4] = {RMAX, RSUM, RSUM, RMAX};
  block_partials<4>(red, ops, partials);
}

__global__ void __launch_bounds__(kBlock) div_fin(const double* partials, int nb, double* result) {
  const int ops[4] = {RMAX, RSUM, RSUM, RMAX};
  double out[4];
  finalize_partials<4>(partials, nb, ops, out);
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; ++q) result[q] = out[q];
}

// ComputePermeability sums (NSBase.cxx:1886-1956), over the (j, k) points
// of the outlet plane.  result: sum of outlet u, sum of p on plane nx-border
// and its node count, sum of p on plane border-1 and its node count.
__global__ void __launch_bounds__(kBlock)
perm_kernel(Geo g, const double* __restrict__ u, const double* __restrict__ p, int border, double* partials) {
  const int64_t m = (int64_t)g.ny * g.nz;
  const RowIter rr = block_rows(m);
  double red[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t q = rr.begin + threadIdx.x; q < rr.end; q += rr.step) {
    const int j = (int)(q / g.nz), k = (int)(q % g.nz);
    if (border <= 0) {
      const int f = g.U(g.nx, j, k);
      if (f > -1) red[0] += u[f];
    } else {
      const int f = g.U(g.nx + 1 - border, j, k);
      if (f > -1) red[0] += u[f];
      int c = g.P(g.nx - border, j, k);
      if (c > -1) { red[1] += p[c]; red[2] += 1.0; }
      c = g.P(border - 1, j, k);
      if (c > -1) { red[3] += p[c]; red[4] += 1.0; }
    }
  }
  const int ops[5] = {RSUM, RSUM, RSUM, RSUM, RSUM};
  block_partials<5>(red, ops, partials);
}

__global__ void __launch_bounds__(kBlock) perm_fin(const double* partials, int nb, double* result) {
  const int ops[5] = {RSUM, RSUM, RSUM, RSUM, RSUM};
  double out[5];
  finalize_partials<5>(partials, nb, ops, out);
  if (threadIdx.x == 0)
    for (int t = 0; t < 5; ++t) result[t] = out[t];
}

// dst = src in 16-byte words, grid-stride (the copy that calibrates HBM
// counters and measures the achievable streaming bandwidth).
__global__ void __launch_bounds__(kBlock) copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                        int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    dst[i] = src[i];
}

}  // namespace

hipError_t launch_copy16(const void* src, void* dst, int64_t words, hipStream_t s) {
  if (words <= 0) return hipSuccess;
  const int64_t nb = std::min<int64_t>((words + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)nb), dim3(kBlock), 0, s, (const uint4*)src, (uint4*)dst, words);
  return hipGetLastError();
}

hipError_t launch_fill(double* x, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, x, n, v);
  return hipGetLastError();
}

hipError_t launch_check_index(const int32_t* idx, int64_t count, int n, double* bad, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(check_index_kernel, dim3((unsigned)((count + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, idx,
                     count, n, bad);
  return hipGetLastError();
}

hipError_t launch_init_pressure_linear(const Geo& g, const int32_t* pcell, int np, double* p, double p0,
                                       double p1, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  hipLaunchKernelGGL(init_p_kernel, dim3((np + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g, pcell, np, p, p0,
                     p1);
  return hipGetLastError();
}

hipError_t launch_clamp(double* x, const double* temp, int64_t n, double dvmax, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(clamp_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, x, temp, n,
                     dvmax);
  return hipGetLastError();
}

hipError_t launch_correct_velocity(const Geo& g, SysView P, SysView U, SysView V, SysView W, int kind,
                                   const double* q, double* tu, double* tv, double* tw, const double* uh,
                                   const double* vh, const double* wh, double dvmax, double p_0, double p_1,
                                   double dx, double dy, double dz, hipStream_t s) {
  if (P.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(correct_kernel, dim3(padded_blocks(P.n)), dim3(kBlock), 0, s, g, P, U, V, W, kind, q, tu,
                     tv, tw, uh, vh, wh, dvmax, p_0, p_1, dx, dy, dz);
  return hipGetLastError();
}

hipError_t launch_correct_pressure(double* p, const double* pc, int64_t n, double beta, double* partials,
                                   double* result, hipStream_t s) {
  const int nb = reduce_blocks(n);
  hipLaunchKernelGGL(correct_p_kernel, dim3(nb), dim3(kBlock), 0, s, p, pc, n, beta, partials);
  hipLaunchKernelGGL(correct_p_fin, dim3(1), dim3(kBlock), 0, s, partials, nb, result);
  return hipGetLastError();
}

hipError_t launch_compute_div(const Geo& g, SysView P, const double* u, const double* v, const double* w,
                              double dx, double dy, double dz, double* partials, double* result, hipStream_t s) {
  const int nb = reduce_blocks(P.n);
  hipLaunchKernelGGL(div_kernel, dim3(nb), dim3(kBlock), 0, s, g, P, u, v, w, dx, dy, dz, partials);
  hipLaunchKernelGGL(div_fin, dim3(1), dim3(kBlock), 0, s, partials, nb, result);
  return hipGetLastError();
}

hipError_t launch_permeability(const Geo& g, const double* u, const double* p, int border, double* partials,
                               double* result, hipStream_t s) {
  const int64_t m = (int64_t)g.ny * g.nz;
  const int nb = reduce_blocks(m);
  hipLaunchKernelGGL(perm_kernel, dim3(nb), dim3(kBlock), 0, s, g, u, p, border, partials);
  hipLaunchKernelGGL(perm_fin, dim3(1), dim3(kBlock), 0, s, partials, nb, result);
  return hipGetLastError();
}

}  // namespace ns3d
