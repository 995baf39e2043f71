// kernels_mg.hip -- aggregation multigrid for the pressure-type systems of
// the mixed form.
//
// The pressure matrices are symmetric and semi-definite on the fluid cells
// of a voxel grid: 7-point couplings -1/(a_f h^2), the diagonal their negated
// sum plus the inlet/outlet terms, where a_f is a momentum diagonal or row
// sum -- negative as the reference assembles the momentum rows, so the
// matrices are negated M-matrices (negative diagonal).  Conjugate gradients
// works for either sign (rho and p.Ap share it, alpha > 0), so the FP32
// inner solve of the FP64 defect correction can be CG preconditioned by one
// multigrid V-cycle instead of Jacobi BiCGSTAB; nothing here assumes a sign.  The hierarchy is geometric
// aggregation: the rows of level l+1 are the 2x2x2 blocks of level l's grid
// holding at least one row (numbered in C order, as the index scan numbers
// the fine rows), prolongation is piecewise constant (P), restriction its
// transpose, and the coarse operator the Galerkin product P^T A P, which is
// again a symmetric 7-point matrix of the same kind, stored like the fine one (diagonal and
// the E, N, T couplings, the W, S, B couplings read from the neighbour).
//
// V-cycle on level l for right-hand side b (x = 0 to start):
//   pre-smoothing   x1 = w D^-1 b                        (one Jacobi step)
//   restriction     b' = P^T (b - A x1)                  (mg_restrict; on the
//                   finest level mg_resid0 + mg_gather0, which admits the iteration)
//   coarse solve    e = V-cycle(l+1, b'), or on the coarsest level a few
//                   Jacobi-preconditioned CG steps in one work-group (FP32)
//   correction and  x2 = x1 + s P e;  x3 = x2 + w D^-1 (b - A x2)
//   post-smoothing                                      (mg_prolong)
// x1 and x2 are never stored: each kernel recomputes them at the seven
// points of a row from b, D^-1, the aggregate map and e.
//
// The Krylov loop is flexible CG (the coarsest solve is not a fixed linear
// map): beta = -alpha_prev (z . q_prev) / rho_prev, the Polak-Ribiere form,
// equal to rho / rho_prev for a fixed symmetric preconditioner.  One
// iteration is the V-cycle, then
//   mg_dir:    p' = z + beta p (recomputed at the seven points), q = A p',
//              partial p'.q
//   mg_update: alpha = rho / p'.q; x += alpha p'; r -= alpha q; partial r.r
// and the next iteration's first restriction sums r.r and stops the solve
// when |r|^2 <= tol^2 |r_0|^2 or the iteration cap is reached.  Every
// kernel reads the iteration it belongs to from the device state (the first
// restriction admits it_c, later kernels check it_a), so a queued chunk of
// iterations is the same sequence of launches every time (a graph can
// replay it) and launches queued after the solve stopped exit at once.
// Reductions are fixed-order (per-block partials summed in block order by
// every block of the next kernel): the result does not depend on timing.
#include <hip/hip_ext.h>

#include <algorithm>

#include "kernels.h"

namespace ns3d {

namespace {

constexpr int kSlotRR = 0, kSlotRZ = 1, kSlotZQ = 2, kSlotPQ = 3;
// the finest prolongation writes r.z and z.q into two consecutive slots
static_assert(kSlotZQ == kSlotRZ + 1, "r.z and z.q partials are adjacent");

// The seven points of a row (W, S, B, self, T, N, E) and its coefficients.
struct MgPts {
  int q[7];
  float c[7];
};

__device__ __forceinline__ void mg_row(const MgLevel& L, int r, MgPts& P) {
  const uint64_t pw = L.pk[r];
  const int dT = (int)(pw & 3u), dB = (int)((pw >> 2) & 3u);
  const int dN = (int)((pw >> 4) & 0x7ffu), dS = (int)((pw >> 15) & 0x7ffu);
  const int dE = (int)((pw >> 26) & 0x7ffffu), dW = (int)((pw >> 45) & 0x7ffffu);
  const int64_t n = L.n;
  P.q[0] = r - dW; P.q[1] = r - dS; P.q[2] = r - dB; P.q[3] = r;
  P.q[4] = r + dT; P.q[5] = r + dN; P.q[6] = r + dE;
  P.c[3] = L.coef[r];
  P.c[6] = L.coef[L.sE * n + r];
  P.c[5] = L.coef[L.sN * n + r];
  P.c[4] = L.coef[L.sT * n + r];
  P.c[0] = dW ? L.coef[L.sE * n + (r - dW)] : 0.f;
  P.c[1] = dS ? L.coef[L.sN * n + (r - dS)] : 0.f;
  P.c[2] = dB ? L.coef[L.sT * n + (r - dB)] : 0.f;
}

// Block-uniform admission of a kernel of the current iteration: thread 0
// reads the state once for the whole block.
__device__ __forceinline__ bool mg_admit(const MgState* st, long long& it) {
  __shared__ int go;
  __shared__ long long sit;
  if (threadIdx.x == 0) {
    go = st->status == BS_RUNNING;
    sit = st->it_a;
  }
  __syncthreads();
  it = sit;
  return go != 0;
}

// ---- hierarchy construction ------------------------------------------------

__global__ void __launch_bounds__(kBlock) mg_fill_i32(int32_t* x, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = v;
}

// flags[block of the row's cell] = 1 (every writer stores the same value)
__global__ void __launch_bounds__(kBlock) mg_mark_kernel(MgLevel F, int cny, int cnz, int32_t* flags) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= F.n) return;
  const int64_t c = F.cell[r];
  const int k = (int)(c % F.nz);
  const int64_t t = c / F.nz;
  const int j = (int)(t % F.ny), i = (int)(t / F.ny);
  flags[((int64_t)(i >> 1) * cny + (j >> 1)) * cnz + (k >> 1)] = 1;
}

constexpr int kMgScanCpt = 16;
constexpr int kMgScanCells = kBlock * kMgScanCpt;

// Exclusive scan of one int per thread over the block; *tot = block total.
__device__ int mg_block_exscan(int v, int* tot) {
  __shared__ int sh[kBlock];
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    const int add = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  const int incl = sh[t];
  *tot = sh[kBlock - 1];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(kBlock) mg_scan_count(const int32_t* flags, int64_t ncell, int32_t* counts) {
  const int64_t base = (int64_t)blockIdx.x * kMgScanCells + (int64_t)threadIdx.x * kMgScanCpt;
  int cnt = 0;
  for (int q = 0; q < kMgScanCpt; ++q) {
    const int64_t c = base + q;
    if (c < ncell) cnt += flags[c] != 0;
  }
  int tot;
  (void)mg_block_exscan(cnt, &tot);
  if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

// One block: exclusive scan of the block counts in place; total at counts[nb].
__global__ void __launch_bounds__(kBlock) mg_scan_offsets(int32_t* counts, int nb) {
  int run = 0;
  for (int b0 = 0; b0 < nb; b0 += kBlock) {
    const int b = b0 + threadIdx.x;
    const int v = b < nb ? counts[b] : 0;
    int tot;
    const int ex = mg_block_exscan(v, &tot);
    if (b < nb) counts[b] = run + ex;
    run += tot;
  }
  if (threadIdx.x == 0) counts[nb] = run;
}

// grid[c] = coarse row of flagged cell c (-1 elsewhere), cell[row] = c.
__global__ void __launch_bounds__(kBlock) mg_scan_write(int32_t* grid, int64_t ncell, const int32_t* offsets,
                                                         int32_t* cell) {
  const int64_t base = (int64_t)blockIdx.x * kMgScanCells + (int64_t)threadIdx.x * kMgScanCpt;
  int cnt = 0;
  for (int q = 0; q < kMgScanCpt; ++q) {
    const int64_t c = base + q;
    if (c < ncell) cnt += grid[c] != 0;
  }
  int tot;
  int row = offsets[blockIdx.x] + mg_block_exscan(cnt, &tot);
  for (int q = 0; q < kMgScanCpt; ++q) {
    const int64_t c = base + q;
    if (c >= ncell) break;
    if (grid[c] != 0) {
      grid[c] = row;
      cell[row] = (int32_t)c;
      ++row;
    } else {
      grid[c] = -1;
    }
  }
}

// agg[r] = coarse row holding fine row r.
__global__ void __launch_bounds__(kBlock) mg_agg_kernel(MgLevel F, MgLevel C, int32_t* agg) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= F.n) return;
  const int64_t c = F.cell[r];
  const int k = (int)(c % F.nz);
  const int64_t t = c / F.nz;
  const int j = (int)(t % F.ny), i = (int)(t / F.ny);
  agg[r] = C.grid[((int64_t)(i >> 1) * C.ny + (j >> 1)) * C.nz + (k >> 1)];
}

// Packed neighbour distances of a coarse level (the fine level's format:
// T | B << 2 | N << 4 | S << 15 | E << 26 | W << 45); *bad = 1 if one does
// not fit.
__global__ void __launch_bounds__(kBlock) mg_pk_kernel(MgLevel C, uint64_t* pk, int* bad) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= C.n) return;
  const int64_t c = C.cell[r];
  const int k = (int)(c % C.nz);
  const int64_t t = c / C.nz;
  const int j = (int)(t % C.ny), i = (int)(t / C.ny);
  const int64_t plane = (int64_t)C.ny * C.nz;
  auto at = [&](bool ok, int64_t cc) -> int64_t { return ok ? (int64_t)C.grid[cc] : -1; };
  const int64_t nE = at(i + 1 < C.nx, c + plane), nW = at(i > 0, c - plane);
  const int64_t nN = at(j + 1 < C.ny, c + C.nz), nS = at(j > 0, c - C.nz);
  const int64_t nT = at(k + 1 < C.nz, c + 1), nB = at(k > 0, c - 1);
  const int64_t dE = nE < 0 ? 0 : nE - r, dW = nW < 0 ? 0 : r - nW;
  const int64_t dN = nN < 0 ? 0 : nN - r, dS = nS < 0 ? 0 : r - nS;
  const int64_t dT = nT < 0 ? 0 : nT - r, dB = nB < 0 ? 0 : r - nB;
  const int64_t lTB = 1ll << kPkTB, lNS = 1ll << kPkNS, lEW = 1ll << kPkEW;
  const bool ok = dT >= 0 && dT < lTB && dB >= 0 && dB < lTB && dN >= 0 && dN < lNS && dS >= 0 && dS < lNS &&
                  dE >= 0 && dE < lEW && dW >= 0 && dW < lEW;
  if (!ok) {
    *bad = 1;
    pk[r] = 0;
    return;
  }
  pk[r] = (uint64_t)dT | ((uint64_t)dB << 2) | ((uint64_t)dN << 4) | ((uint64_t)dS << 15) | ((uint64_t)dE << 26) |
          ((uint64_t)dW << 45);
}

// Galerkin product P^T A P of level F onto level C (one thread per coarse
// row): the diagonal sums the children's diagonals and, twice, the
// couplings between two children; the E (N, T) coupling sums the E (N, T)
// couplings of the children on the block's high side in that direction.
// Sums in FP64 in a fixed order, stored FP32; D^-1 is 0 on an empty row.
__global__ void __launch_bounds__(kBlock) mg_galerkin_kernel(MgLevel F, MgLevel C, float* coef, float* invd) {
  const int64_t R = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (R >= C.n) return;
  const int64_t cc = C.cell[R];
  const int K = (int)(cc % C.nz);
  const int64_t t = cc / C.nz;
  const int J = (int)(t % C.ny), I = (int)(t / C.ny);
  const int64_t n = F.n;
  double d = 0.0, e = 0.0, nn = 0.0, tt = 0.0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const int i = 2 * I + a, j = 2 * J + b, k = 2 * K + c;
        if (i >= F.nx || j >= F.ny || k >= F.nz) continue;
        const int r = F.grid[((int64_t)i * F.ny + j) * F.nz + k];
        if (r < 0) continue;
        const double cE = F.coef[F.sE * n + r], cN = F.coef[F.sN * n + r], cT = F.coef[F.sT * n + r];
        d += (double)F.coef[r];
        if (a == 0) d += 2.0 * cE; else e += cE;
        if (b == 0) d += 2.0 * cN; else nn += cN;
        if (c == 0) d += 2.0 * cT; else tt += cT;
      }
  const int64_t m = C.n;
  coef[R] = (float)d;
  coef[m + R] = (float)e;
  coef[2 * m + R] = (float)nn;
  coef[3 * m + R] = (float)tt;
  invd[R] = d != 0.0 ? (float)(1.0 / d) : 0.0f;
}

// ---- V-cycle ------------------------------------------------------------------

// Admission of an iteration by its first kernel (block-uniform): sums the
// r.r partials of the previous update (or of init) and stops the solve when
// converged or at the cap; false when the solve has stopped.
__device__ __forceinline__ bool mg_admit_first(const MgFcg& G, long long& it) {
  MgState* st = G.st;
  __shared__ int go;
  __shared__ long long sit;
  if (threadIdx.x == 0) {
    go = st->status == BS_RUNNING;
    sit = st->it_c;
  }
  __syncthreads();
  if (!go) return false;
  it = sit;
  double pa[1], rr[1];
  partials_load<1>(G.part + kSlotRR * kMaxRedBlocks, G.nb, kMaxRedBlocks, pa);
  partials_reduce<1>(pa, rr);
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!(rr[0] > st->tol2 && it < st->maxit)) {
    if (lead) {
      st->rr = rr[0];
      st->status = BS_DONE;
    }
    return false;
  }
  if (lead) {
    st->rr = rr[0];
    st->it_a = it;
  }
  return true;
}

// The residual of one fine row after the pre-smoothing step from zero:
// b_r - sum_m c_m (w D^-1 b)_q.
__device__ __forceinline__ float mg_presmoothed_residual(const MgLevel& F, int r, float w) {
  MgPts P;
  mg_row(F, r, P);
  float ax = 0.f;
#pragma unroll
  for (int m = 0; m < 7; ++m) {
    const int q = P.q[m];
    ax += P.c[m] * (w * F.invd[q] * F.b[q]);
  }
  return F.b[r] - ax;
}

// b' = P^T (b - A w D^-1 b) on the coarse rows of C (levels below the
// finest): eight consecutive threads per coarse row, one per child block
// position (a, b, c), each the residual of its fine row (0 where there is
// none), summed over the eight lanes by a fixed butterfly.
__global__ void __launch_bounds__(kBlock) mg_restrict_kernel(MgLevel F, MgLevel C, MgFcg G) {
  long long it;
  if (!mg_admit(G.st, it)) return;
  const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t R = g >> 3;
  const int ch = (int)(g & 7);
  // rows past the end still take part in the butterfly (whole groups of
  // eight lie on one side of C.n, so they only meet each other)
  float res = 0.f;
  if (R < C.n) {
    const int64_t cc = C.cell[R];
    const int K = (int)(cc % C.nz);
    const int64_t t = cc / C.nz;
    const int J = (int)(t % C.ny), I = (int)(t / C.ny);
    const int i = 2 * I + (ch >> 2), j = 2 * J + ((ch >> 1) & 1), k = 2 * K + (ch & 1);
    if (i < F.nx && j < F.ny && k < F.nz) {
      const int r = F.grid[((int64_t)i * F.ny + j) * F.nz + k];
      if (r >= 0) res = mg_presmoothed_residual(F, r, G.omega);
    }
  }
  res += __shfl_xor(res, 1, 64);
  res += __shfl_xor(res, 2, 64);
  res += __shfl_xor(res, 4, 64);
  if (ch == 0 && R < C.n) C.b[R] = res;
}

// The finest level's restriction in two passes: the pre-smoothed residual
// of every fine row into F.x (free until the last prolongation writes z),
// row-parallel over the solve's blocks (this kernel admits the iteration),
// then the children's sum per coarse row in the fixed butterfly order of
// mg_restrict_kernel (same bits).
__global__ void __launch_bounds__(kBlock) mg_resid0_kernel(MgLevel F, MgFcg G) {
  long long it;
  if (!mg_admit_first(G, it)) return;
  const RowIter rr = block_rows(F.n);
  for (int64_t r = rr.begin + threadIdx.x; r < rr.end; r += rr.step)
    F.x[r] = mg_presmoothed_residual(F, (int)r, G.omega);
}

__global__ void __launch_bounds__(kBlock) mg_gather0_kernel(MgLevel F, MgLevel C, MgFcg G) {
  long long it;
  if (!mg_admit(G.st, it)) return;
  const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t R = g >> 3;
  const int ch = (int)(g & 7);
  float res = 0.f;
  if (R < C.n) {
    const int64_t cc = C.cell[R];
    const int K = (int)(cc % C.nz);
    const int64_t t = cc / C.nz;
    const int J = (int)(t % C.ny), I = (int)(t / C.ny);
    const int i = 2 * I + (ch >> 2), j = 2 * J + ((ch >> 1) & 1), k = 2 * K + (ch & 1);
    if (i < F.nx && j < F.ny && k < F.nz) {
      const int r = F.grid[((int64_t)i * F.ny + j) * F.nz + k];
      if (r >= 0) res = F.x[r];
    }
  }
  res += __shfl_xor(res, 1, 64);
  res += __shfl_xor(res, 2, 64);
  res += __shfl_xor(res, 4, 64);
  if (ch == 0 && R < C.n) C.b[R] = res;
}

// Block sum of one float over a kMgCoarseThreads-thread block, valid in
// every thread (fixed order).
__device__ __forceinline__ float mg_bsum(float v) {
  constexpr int kW = kMgCoarseThreads / 64;
  __shared__ float sh[kW];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  float s = sh[0];
#pragma unroll
  for (int q = 1; q < kW; ++q) s += sh[q];
  return s;
}

// Coarsest level: `iters` Jacobi-preconditioned CG steps from x = 0 in one
// work-group, the direction vector in LDS (at most kMgCoarseRows rows).
__global__ void __launch_bounds__(kMgCoarseThreads) mg_coarse_kernel(MgLevel L, MgFcg G, int iters) {
  long long it;
  if (!mg_admit(G.st, it)) return;
  constexpr int kPer = kMgCoarseRows / kMgCoarseThreads;
  __shared__ float ps[kMgCoarseRows];
  const int n = L.n;
  MgPts P[kPer];
  float x[kPer], r[kPer], z[kPer], p[kPer], dg[kPer];
  float rz = 0.f;
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int row = threadIdx.x + kMgCoarseThreads * u;
    x[u] = 0.f;
    if (row < n) {
      mg_row(L, row, P[u]);
      dg[u] = L.invd[row];
      r[u] = L.b[row];
    } else {
      dg[u] = 0.f;
      r[u] = 0.f;
#pragma unroll
      for (int m = 0; m < 7; ++m) { P[u].q[m] = 0; P[u].c[m] = 0.f; }
    }
    z[u] = dg[u] * r[u];
    p[u] = z[u];
    if (row < n) ps[row] = p[u];
    rz += r[u] * z[u];
  }
  rz = mg_bsum(rz);  // (its barriers also publish ps)
  for (int k = 0; k < iters && rz != 0.f; ++k) {
    float q[kPer];
    float pq = 0.f;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      float sv = 0.f;
#pragma unroll
      for (int m = 0; m < 7; ++m) sv += P[u].c[m] * ps[P[u].q[m]];
      q[u] = sv;
      pq += p[u] * sv;
    }
    pq = mg_bsum(pq);
    const float al = rz / pq;
    if (!(al > 0.f && al < __builtin_huge_valf())) break;  // breakdown (or p.q = 0)
    float rzn = 0.f;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      x[u] += al * p[u];
      r[u] -= al * q[u];
      z[u] = dg[u] * r[u];
      rzn += r[u] * z[u];
    }
    rzn = mg_bsum(rzn);
    const float be = rzn / rz;
    rz = rzn;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      p[u] = z[u] + be * p[u];
      const int row = threadIdx.x + kMgCoarseThreads * u;
      if (row < n) ps[row] = p[u];
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int row = threadIdx.x + kMgCoarseThreads * u;
    if (row < n) L.x[row] = x[u];
  }
}

// x3 = x2 + w D^-1 (b - A x2), x2 = w D^-1 b + s P e, on the rows of F (e =
// C.x).  On the finest level (TOP) x3 is the preconditioned residual z, and
// the kernel also writes the partials r.z and z.q (q of the previous
// iteration, zero at the first).
template <bool TOP>
__global__ void __launch_bounds__(kBlock) mg_prolong_kernel(MgLevel F, MgLevel C, MgFcg G) {
  long long it;
  if (!mg_admit(G.st, it)) return;
  const float w = G.omega, s = G.scale;
  auto x2 = [&](int q) { return w * F.invd[q] * F.b[q] + s * C.x[F.agg[q]]; };
  if (!TOP) {
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= F.n) return;
    MgPts P;
    mg_row(F, (int)r, P);
    float xv[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) xv[m] = x2(P.q[m]);
    float ax = P.c[0] * xv[0];
#pragma unroll
    for (int m = 1; m < 7; ++m) ax += P.c[m] * xv[m];
    F.x[r] = xv[3] + w * F.invd[r] * (F.b[r] - ax);
    return;
  }
  double a[2] = {0.0, 0.0};
  const RowIter rr = block_rows(F.n);
  for (int64_t r = rr.begin + threadIdx.x; r < rr.end; r += rr.step) {
    MgPts P;
    mg_row(F, (int)r, P);
    float xv[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) xv[m] = x2(P.q[m]);
    float ax = P.c[0] * xv[0];
#pragma unroll
    for (int m = 1; m < 7; ++m) ax += P.c[m] * xv[m];
    const float bv = F.b[r];
    const float z = xv[3] + w * F.invd[r] * (bv - ax);
    F.x[r] = z;
    a[0] += (double)bv * z;
    a[1] += (double)z * G.q[r];
  }
  block_sums_at<2>(a, G.part + kSlotRZ * kMaxRedBlocks, kMaxRedBlocks, blockIdx.x);
}

// ---- flexible CG -----------------------------------------------------------

// r = rhs, x = 0, p (first buffer) = q = 0; partials |r|^2.
__global__ void __launch_bounds__(kBlock) mg_init_kernel(MgFcg G, const float* __restrict__ rhs) {
  const MgLevel& F = G.f;
  double a[1] = {0.0};
  const RowIter rr = block_rows(F.n);
  for (int64_t r = rr.begin + threadIdx.x; r < rr.end; r += rr.step) {
    const float v = rhs[r];
    F.b[r] = v;
    G.x[r] = 0.f;
    G.p[0][r] = 0.f;
    G.q[r] = 0.f;
    a[0] += (double)v * v;
  }
  block_sums_at<1>(a, G.part + kSlotRR * kMaxRedBlocks, kMaxRedBlocks, blockIdx.x);
}

__global__ void __launch_bounds__(kBlock) mg_init_fin(MgFcg G, double tol, long long maxit) {
  double out[1];
  sum_partials_all<1>(G.part + kSlotRR * kMaxRedBlocks, G.nb, out, kMaxRedBlocks);
  if (threadIdx.x != 0) return;
  MgState* st = G.st;
  st->rho[0] = st->rho[1] = 0.0;
  st->alpha[0] = st->alpha[1] = 0.0;
  st->rr0 = out[0];
  st->rr = out[0];
  st->tol2 = tol * tol * out[0];
  st->maxit = maxit;
  st->it_c = 0;
  st->it_a = -1;
  st->status = out[0] == 0.0 ? BS_ZERO_RHS : BS_RUNNING;
}

// p' = z + beta p at the seven points, q = A p', partial p'.q.
__global__ void __launch_bounds__(kBlock) mg_dir_kernel(MgFcg G) {
  MgState* st = G.st;
  long long it;
  if (!mg_admit(st, it)) return;
  const MgLevel& F = G.f;
  double pa[2], red[2];
  partials_load<2>(G.part + kSlotRZ * kMaxRedBlocks, G.nb, kMaxRedBlocks, pa);
  partials_reduce<2>(pa, red);
  const double rho = red[0], zq = red[1];
  double beta = 0.0;
  if (it > 0) {
    const double ap = st->alpha[(it - 1) & 1], rp = st->rho[(it - 1) & 1];
    if (rp != 0.0) beta = -ap * zq / rp;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->rho[it & 1] = rho;
  const float bf = (float)beta;
  const float* __restrict__ pold = G.p[it & 1];
  float* __restrict__ pnew = G.p[(it & 1) ^ 1];
  const float* __restrict__ z = F.x;
  double a[1] = {0.0};
  const RowIter rr = block_rows(F.n);
  for (int64_t r = rr.begin + threadIdx.x; r < rr.end; r += rr.step) {
    MgPts P;
    mg_row(F, (int)r, P);
    float pv[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) pv[m] = z[P.q[m]] + bf * pold[P.q[m]];
    float qv = P.c[0] * pv[0];
#pragma unroll
    for (int m = 1; m < 7; ++m) qv += P.c[m] * pv[m];
    pnew[r] = pv[3];
    G.q[r] = qv;
    a[0] += (double)pv[3] * qv;
  }
  block_sums_at<1>(a, G.part + kSlotPQ * kMaxRedBlocks, kMaxRedBlocks, blockIdx.x);
}

// alpha = rho / p'.q; x += alpha p'; r -= alpha q; partial r.r.  An alpha
// that is not positive and finite (breakdown: rho and p'.q of different
// signs, or p'.q = 0) stops the solve.
__global__ void __launch_bounds__(kBlock) mg_update_kernel(MgFcg G) {
  MgState* st = G.st;
  long long it;
  if (!mg_admit(st, it)) return;
  const MgLevel& F = G.f;
  double pa[1], red[1];
  partials_load<1>(G.part + kSlotPQ * kMaxRedBlocks, G.nb, kMaxRedBlocks, pa);
  partials_reduce<1>(pa, red);
  const double pq = red[0], rho = st->rho[it & 1];
  const double q0 = rho / pq;
  const bool ok = q0 > 0.0 && q0 < __builtin_huge_val();
  const double al = ok ? q0 : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha[it & 1] = al;
    st->it_c = it + 1;
    if (!ok) st->status = BS_DONE;
  }
  const float af = (float)al;
  const float* __restrict__ pnew = G.p[(it & 1) ^ 1];
  double a[1] = {0.0};
  const RowIter rr = block_rows(F.n);
  for (int64_t r = rr.begin + threadIdx.x; r < rr.end; r += rr.step) {
    G.x[r] += af * pnew[r];
    const float rv = F.b[r] - af * G.q[r];
    F.b[r] = rv;
    a[0] += (double)rv * rv;
  }
  block_sums_at<1>(a, G.part + kSlotRR * kMaxRedBlocks, kMaxRedBlocks, blockIdx.x);
}

inline unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

}  // namespace

hipError_t launch_mg_fill_i32(int32_t* x, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(mg_fill_i32, dim3(blocks_for(n)), dim3(kBlock), 0, s, x, n, v);
  return hipGetLastError();
}

hipError_t launch_mg_mark(const MgLevel& F, const MgLevel& C, int32_t* flags, hipStream_t s) {
  hipLaunchKernelGGL(mg_mark_kernel, dim3(blocks_for(F.n)), dim3(kBlock), 0, s, F, C.ny, C.nz, flags);
  return hipGetLastError();
}

int mg_scan_blocks(int64_t ncell) { return (int)((ncell + kMgScanCells - 1) / kMgScanCells); }

hipError_t launch_mg_scan(int32_t* grid, int64_t ncell, int32_t* counts, int32_t* cell_or_null, hipStream_t s) {
  const int nb = mg_scan_blocks(ncell);
  if (!cell_or_null) {
    hipLaunchKernelGGL(mg_scan_count, dim3(nb), dim3(kBlock), 0, s, grid, ncell, counts);
    hipLaunchKernelGGL(mg_scan_offsets, dim3(1), dim3(kBlock), 0, s, counts, nb);
  } else {
    hipLaunchKernelGGL(mg_scan_write, dim3(nb), dim3(kBlock), 0, s, grid, ncell, counts, cell_or_null);
  }
  return hipGetLastError();
}

hipError_t launch_mg_agg(const MgLevel& F, const MgLevel& C, int32_t* agg, hipStream_t s) {
  hipLaunchKernelGGL(mg_agg_kernel, dim3(blocks_for(F.n)), dim3(kBlock), 0, s, F, C, agg);
  return hipGetLastError();
}

hipError_t launch_mg_pk(const MgLevel& C, uint64_t* pk, int* bad, hipStream_t s) {
  hipLaunchKernelGGL(mg_pk_kernel, dim3(blocks_for(C.n)), dim3(kBlock), 0, s, C, pk, bad);
  return hipGetLastError();
}

hipError_t launch_mg_galerkin(const MgLevel& F, const MgLevel& C, float* coef, float* invd, hipStream_t s) {
  hipLaunchKernelGGL(mg_galerkin_kernel, dim3(blocks_for(C.n)), dim3(kBlock), 0, s, F, C, coef, invd);
  return hipGetLastError();
}

hipError_t launch_mg_init(const MgFcg& G, const float* rhs, double tol, long long maxit, hipStream_t s) {
  hipLaunchKernelGGL(mg_init_kernel, dim3(G.nb), dim3(kBlock), 0, s, G, rhs);
  hipLaunchKernelGGL(mg_init_fin, dim3(1), dim3(kBlock), 0, s, G, tol, maxit);
  return hipGetLastError();
}

// A launch with start / stop events taken from its own dispatch when ev is
// given.
template <typename K, typename... Args>
void mg_launch(K kernel, dim3 grid, dim3 block, hipStream_t s, const hipEvent_t* ev, Args... args) {
  if (ev) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev[0], ev[1], 0, args...);
  else hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

hipError_t launch_mg_iteration(const MgFcg& G, const MgLevel* lv, int nlev, int coarse_iters, hipStream_t s,
                               const hipEvent_t* ev) {
  // down: restrict 0 -> 1 -> ... -> nlev-1
  for (int l = 0; l + 1 < nlev; ++l) {
    const dim3 grid(blocks_for(8 * (int64_t)lv[l + 1].n));
    if (l == 0) {
      mg_launch(mg_resid0_kernel, dim3(G.nb), dim3(kBlock), s, ev, lv[0], G);
      hipLaunchKernelGGL(mg_gather0_kernel, grid, dim3(kBlock), 0, s, lv[0], lv[1], G);
    } else {
      hipLaunchKernelGGL(mg_restrict_kernel, grid, dim3(kBlock), 0, s, lv[l], lv[l + 1], G);
    }
  }
  hipLaunchKernelGGL(mg_coarse_kernel, dim3(1), dim3(kMgCoarseThreads), 0, s, lv[nlev - 1], G, coarse_iters);
  for (int l = nlev - 2; l >= 0; --l) {
    if (l == 0)
      mg_launch(mg_prolong_kernel<true>, dim3(G.nb), dim3(kBlock), s, ev ? ev + 2 : nullptr, lv[0], lv[1], G);
    else
      hipLaunchKernelGGL(mg_prolong_kernel<false>, dim3(blocks_for(lv[l].n)), dim3(kBlock), 0, s, lv[l], lv[l + 1],
                         G);
  }
  mg_launch(mg_dir_kernel, dim3(G.nb), dim3(kBlock), s, ev ? ev + 4 : nullptr, G);
  mg_launch(mg_update_kernel, dim3(G.nb), dim3(kBlock), s, ev ? ev + 6 : nullptr, G);
  return hipGetLastError();
}

}  // namespace ns3d
