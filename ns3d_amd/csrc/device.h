// device.h -- device-side data views and helpers shared by the HIP kernels.
//
// Layout in HBM (one context = one image):
//   * index grids, int32, C order (i slowest, k fastest): p, v, w over
//     (nx, ny, nz); u over (nx+1, ny, nz); -1 = no node;
//   * per linear system (u, v, w, p), structure-of-arrays over rows:
//       cell[n]      row -> linear position in the system's grid
//       nbr[6][n]    neighbour rows in the order E, W, N, S, T, B
//                    (i+1, i-1, j+1, j-1, k+1, k-1); the row itself if absent
//       coef[7][n]   0: diagonal, 1..6: E, W, N, S, T, B off-diagonals
//                    (0 where absent)
//       invdiag[n]   Jacobi preconditioner, rowsum[n] (momentum only)
//   * vectors are dense FP64 over the system's rows.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ns3d {

#ifndef NS3D_BLOCK
#define NS3D_BLOCK 256
#endif
constexpr int kBlock = NS3D_BLOCK;  // threads per block for row kernels (wave64s)
constexpr int kWaves = kBlock / 64;
constexpr int kXcd = 8;       // MI355X: 8 XCDs, workgroups dispatched round-robin

enum Dir { DE = 0, DW = 1, DN = 2, DS = 3, DT = 4, DB = 5 };

struct SysView {
  int n;
  int d0;                 // grid extent along x (nx+1 for u)
  const int32_t* grid;
  const int32_t* cell;
  const int32_t* nbr;     // [6][n]
  double* coef;           // [7][n]
  double* invdiag;
  double* rowsum;
};

struct Geo {
  int nx, ny, nz;
  const int32_t* pg;
  const int32_t* ug;
  const int32_t* vg;
  const int32_t* wg;
  __device__ __forceinline__ int64_t lin(int i, int j, int k) const {
    return ((int64_t)i * ny + j) * nz + k;
  }
  __device__ __forceinline__ bool in3(int i, int j, int k, int d0) const {
    return i >= 0 && i < d0 && j >= 0 && j < ny && k >= 0 && k < nz;
  }
  __device__ __forceinline__ int P(int i, int j, int k) const {
    return in3(i, j, k, nx) ? pg[lin(i, j, k)] : -1;
  }
  __device__ __forceinline__ int U(int i, int j, int k) const {
    return in3(i, j, k, nx + 1) ? ug[lin(i, j, k)] : -1;
  }
  __device__ __forceinline__ int V(int i, int j, int k) const {
    return in3(i, j, k, nx) ? vg[lin(i, j, k)] : -1;
  }
  __device__ __forceinline__ int W(int i, int j, int k) const {
    return in3(i, j, k, nx) ? wg[lin(i, j, k)] : -1;
  }
  __device__ __forceinline__ void decode(int64_t c, int& i, int& j, int& k) const {
    k = (int)(c % nz);
    int64_t t = c / nz;
    j = (int)(t % ny);
    i = (int)(t / ny);
  }
};

// SafeVectorValue (NSBase.cxx:1757-1771)
__device__ __forceinline__ double svv(const double* vec, int idx) { return idx > -1 ? vec[idx] : 0.0; }

// Maps the hardware block index to a logical block so that each XCD sweeps a
// contiguous 1/8 of the rows (blocks are dispatched to XCDs round-robin and
// gridDim.x is padded to a multiple of 8 by the launcher).
__device__ __forceinline__ int logical_block() {
  const int nb = gridDim.x;
  const int per = nb / kXcd;
  const int b = blockIdx.x;
  return (b % kXcd) * per + b / kXcd;
}

inline int padded_blocks(int64_t n) {
  int64_t nb = (n + kBlock - 1) / kBlock;
  if (nb < 1) nb = 1;
  nb = (nb + kXcd - 1) / kXcd * kXcd;
  return (int)nb;
}

enum RedOp { RSUM = 0, RMAX = 1, RMIN = 2 };

template <int OP>
__device__ __forceinline__ double red_apply(double a, double b) {
  if (OP == RSUM) return a + b;
  if (OP == RMAX) return a > b ? a : b;
  return a < b ? a : b;
}

template <int OP>
__device__ __forceinline__ double red_identity() {
  if (OP == RSUM) return 0.0;
  if (OP == RMAX) return -__builtin_huge_val();
  return __builtin_huge_val();
}

template <int OP>
__device__ __forceinline__ double wave_reduce(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = red_apply<OP>(v, __shfl_xor(v, off, 64));
  return v;
}

// Block reduction in a fixed order (deterministic); result valid in thread 0.
template <int OP>
__device__ __forceinline__ double block_reduce(double v, double* sh) {
  v = wave_reduce<OP>(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double r = v;
  if (threadIdx.x == 0) {
    r = sh[0];
#pragma unroll
    for (int q = 1; q < kWaves; ++q) r = red_apply<OP>(r, sh[q]);
  }
  return r;
}

__device__ __forceinline__ double red_rt(int op, double a, double b) {
  return op == RSUM ? a + b : (op == RMAX ? (a > b ? a : b) : (a < b ? a : b));
}
__device__ __forceinline__ double red_id_rt(int op) {
  return op == RSUM ? 0.0 : (op == RMAX ? -__builtin_huge_val() : __builtin_huge_val());
}

__device__ __forceinline__ double block_reduce_rt(int op, double v, double* sh) {
  if (op == RSUM) return block_reduce<RSUM>(v, sh);
  if (op == RMAX) return block_reduce<RMAX>(v, sh);
  return block_reduce<RMIN>(v, sh);
}

// Grid-wide reductions are two-stage: every block of a row kernel stores its
// partials with plain stores (slot q*gridDim.x + blockIdx.x), and a one-block
// finalize kernel launched next on the same stream reduces them in block
// order.  The kernel boundary makes the partials visible; no fences or
// atomics are needed (a device-scope fence per block forces an L2 write-back
// on every XCD and costs far more than the extra launch).  The result is
// deterministic for a given grid size.
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], const int (&ops)[NV], double* partials) {
  __shared__ double sh[NV][kWaves];
  double bv[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) bv[q] = block_reduce_rt(ops[q], v[q], sh[q]);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) partials[(int64_t)q * gridDim.x + blockIdx.x] = bv[q];
  }
}

// In a one-block kernel: reduce nb partials of each of NV values; out[] is
// valid in thread 0.
template <int NV>
__device__ __forceinline__ void finalize_partials(const double* partials, int nb, const int (&ops)[NV],
                                                  double (&out)[NV]) {
  __shared__ double sh[NV][kWaves];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double a = red_id_rt(ops[q]);
    for (int b = threadIdx.x; b < nb; b += kBlock) a = red_rt(ops[q], a, partials[(int64_t)q * nb + b]);
    out[q] = block_reduce_rt(ops[q], a, sh[q]);
  }
}

// Block sums of NV values stored at base[q * qstride + idx] (thread 0).
template <int NV, typename P>
__device__ __forceinline__ void block_sums_at(double (&v)[NV], P base, int64_t qstride, int64_t idx) {
  __shared__ double sh[NV][kWaves];
  double bv[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) bv[q] = block_reduce<RSUM>(v[q], sh[q]);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) base[q * qstride + idx] = bv[q];
  }
}

// In every block of a kernel: all threads get the NV sums of nb partials per
// value (slot q at partials[q * qstride + b], qstride = nb by default).  The order of the additions is
// fixed and the same in every block, so every block computes the same
// bits; this replaces a separate finalize launch between two row kernels.
// In two steps: partials_load (each thread's share, so the loads can be in
// flight with others) and partials_reduce (the block reduction).
template <int NV, typename P>
__device__ __forceinline__ void partials_load(P partials, int nb, int64_t qstride, double (&a)[NV]) {
#pragma unroll
  for (int q = 0; q < NV; ++q) a[q] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) {
#pragma unroll
    for (int q = 0; q < NV; ++q) a[q] += partials[q * qstride + b];
  }
}
template <int NV>
__device__ __forceinline__ void partials_reduce(double (&a)[NV], double (&out)[NV]) {
  __shared__ double sh[NV][kWaves];
#pragma unroll
  for (int q = 0; q < NV; ++q) a[q] = wave_reduce<RSUM>(a[q]);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[q][threadIdx.x >> 6] = a[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double r = sh[q][0];
#pragma unroll
    for (int wv = 1; wv < kWaves; ++wv) r += sh[q][wv];
    out[q] = r;
  }
}
template <int NV, typename P>
__device__ __forceinline__ void sum_partials_all(P partials, int nb, double (&out)[NV], int64_t qstride = -1) {
  if (qstride < 0) qstride = nb;
  __shared__ double sh[NV][kWaves];
  double a[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) a[q] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) {
#pragma unroll
    for (int q = 0; q < NV; ++q) a[q] += partials[q * qstride + b];
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) a[q] = wave_reduce<RSUM>(a[q]);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[q][threadIdx.x >> 6] = a[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double r = sh[q][0];
#pragma unroll
    for (int wv = 1; wv < kWaves; ++wv) r += sh[q][wv];
    out[q] = r;
  }
}

// Row kernels that reduce run on at most kMaxRedBlocks blocks (a multiple of
// 8, so the XCD remap of logical_block() gives each XCD the same number) and
// sweep their rows with a grid-stride loop of a per-block RowIter.  Every XCD
// owns one contiguous eighth of the rows, so neighbour gathers stay in that
// XCD's L2.  Within an XCD (NS3D_ROW_MAP):
//   0: each block owns one contiguous chunk (a multiple of kBlock rows);
//   1: the XCD's blocks sweep its eighth together, block k taking the k-th
//      kBlock-row tile of each window of (blocks per XCD) tiles, so rows
//      visited at the same time are adjacent.
#ifndef NS3D_ROW_MAP
#define NS3D_ROW_MAP 0
#endif
#ifndef NS3D_MAX_RED_BLOCKS
#define NS3D_MAX_RED_BLOCKS 1024
#endif
constexpr int kMaxRedBlocks = NS3D_MAX_RED_BLOCKS;

inline int reduce_blocks(int64_t n) {
  int64_t nb = (n + kBlock - 1) / kBlock;
  if (nb > kMaxRedBlocks) nb = kMaxRedBlocks;
  if (nb < 1) nb = 1;
  nb = (nb + kXcd - 1) / kXcd * kXcd;
  return (int)nb;
}

struct RowIter {
  int64_t begin, end, step;
};

__device__ __forceinline__ RowIter block_rows(int64_t n) {
  const int64_t nb = gridDim.x;
  RowIter r;
#if NS3D_ROW_MAP == 1
  const int64_t per_xcd = nb / kXcd;
  const int64_t xcd = blockIdx.x % kXcd, k = blockIdx.x / kXcd;
  const int64_t span = ((n + kXcd - 1) / kXcd + kBlock - 1) / kBlock * kBlock;
  const int64_t xb = xcd * span;
  const int64_t xe = xb + span < n ? xb + span : n;
  r.begin = xb + k * kBlock < xe ? xb + k * kBlock : xe;
  r.end = xe;
  r.step = per_xcd * kBlock;
#else
  const int64_t per = ((n + nb - 1) / nb + kBlock - 1) / kBlock * kBlock;
  const int64_t b = (int64_t)logical_block() * per;
  r.begin = b < n ? b : n;
  r.end = (b + per) < n ? (b + per) : n;
  r.step = kBlock;
#endif
  return r;
}

}  // namespace ns3d
