// kernels_index.hip -- fluid-index compaction as a two-pass device scan, and
// the neighbour lists of the four staggered systems.
//
// Numbering rule (NSBase.cxx:2495-2573, InitIndexArrays): cells are visited
// in lexicographic (i, j, k) order; at a fluid cell the counters advance in
// this order: v if j>0 and (i,j-1,k) is fluid; w if k>0 and (i,j,k-1) is
// fluid; u(i) if i==0 or (i-1,j,k) is fluid; u(nx) if i==nx-1 (the outlet face
// is numbered inside the last cell's visit); p if any of the six neighbours is
// fluid or i is 0 or nx-1.  Every counter is therefore an exclusive prefix sum
// of a per-cell flag in visit order, which is what the scan computes; the
// result is bit-identical to the serial loop.
#include "kernels.h"

namespace ns3d {

namespace {

enum { F_V = 1, F_W = 2, F_UH = 4, F_UO = 8, F_P = 16 };

__device__ __forceinline__ unsigned cell_flags(const uint8_t* img, int nx, int ny, int nz, int64_t c,
                                               int i, int j, int k) {
  const int64_t sj = nz, si = (int64_t)ny * nz;
  if (img[c] != 0) return 0u;
  unsigned f = 0;
  bool hn = false;
  if (j > 0 && img[c - sj] == 0) { f |= F_V; hn = true; }
  if (j < ny - 1 && img[c + sj] == 0) hn = true;
  if (k > 0 && img[c - 1] == 0) { f |= F_W; hn = true; }
  if (k < nz - 1 && img[c + 1] == 0) hn = true;
  if (i == 0) { f |= F_UH; hn = true; }
  if (i > 0 && img[c - si] == 0) { f |= F_UH; hn = true; }
  if (i == nx - 1) { f |= F_UO; hn = true; }
  if (i < nx - 1 && img[c + si] == 0) hn = true;
  if (hn) f |= F_P;
  return f;
}

__device__ __forceinline__ void counts_of(unsigned f, int& cp, int& cu, int& cv, int& cw) {
  cp += (f >> 4) & 1u;
  cu += ((f >> 2) & 1u) + ((f >> 3) & 1u);
  cv += f & 1u;
  cw += (f >> 1) & 1u;
}

// Exclusive scan over the 256 threads of a block of 4 int counters, in LDS.
__device__ void block_exscan4(int (&x)[4], int (&tot)[4]) {
  __shared__ int sh[4][kScanThreads];
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 4; ++q) sh[q][t] = x[q];
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    int add[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) add[q] = (t >= off) ? sh[q][t - off] : 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) sh[q][t] += add[q];
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    tot[q] = sh[q][kScanThreads - 1];
    x[q] = sh[q][t] - x[q];
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kScanThreads)
scan_count_kernel(const uint8_t* __restrict__ img, int nx, int ny, int nz, int64_t ncell,
                  int32_t* __restrict__ blockcounts, int nblocks) {
  const int64_t base = (int64_t)blockIdx.x * kScanCells + (int64_t)threadIdx.x * kScanCpt;
  int cnt[4] = {0, 0, 0, 0};
  for (int q = 0; q < kScanCpt; ++q) {
    const int64_t c = base + q;
    if (c >= ncell) break;
    const int k = (int)(c % nz);
    const int64_t t = c / nz;
    const int j = (int)(t % ny);
    const int i = (int)(t / ny);
    counts_of(cell_flags(img, nx, ny, nz, c, i, j, k), cnt[0], cnt[1], cnt[2], cnt[3]);
  }
  int tot[4];
  block_exscan4(cnt, tot);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) blockcounts[(int64_t)q * nblocks + blockIdx.x] = tot[q];
  }
}

// One block: exclusive scan of the per-block counts, plus totals.
__global__ void __launch_bounds__(1024)
scan_offsets_kernel(const int32_t* __restrict__ blockcounts, int32_t* __restrict__ blockoffsets,
                    int32_t* __restrict__ totals, int nblocks) {
  __shared__ int sh[4][1024];
  const int t = threadIdx.x;
  const int per = (nblocks + 1023) / 1024;
  const int lo = t * per;
  const int hi = min(lo + per, nblocks);
  int loc[4] = {0, 0, 0, 0};
  for (int b = lo; b < hi; ++b)
#pragma unroll
    for (int q = 0; q < 4; ++q) loc[q] += blockcounts[(int64_t)q * nblocks + b];
#pragma unroll
  for (int q = 0; q < 4; ++q) sh[q][t] = loc[q];
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int add[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) add[q] = (t >= off) ? sh[q][t - off] : 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) sh[q][t] += add[q];
    __syncthreads();
  }
  int run[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) run[q] = sh[q][t] - loc[q];
  for (int b = lo; b < hi; ++b)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      blockoffsets[(int64_t)q * nblocks + b] = run[q];
      run[q] += blockcounts[(int64_t)q * nblocks + b];
    }
  if (t == 1023) {
#pragma unroll
    for (int q = 0; q < 4; ++q) totals[q] = sh[q][1023];
  }
}

__global__ void __launch_bounds__(kScanThreads)
scan_write_kernel(const uint8_t* __restrict__ img, int nx, int ny, int nz, int64_t ncell,
                  const int32_t* __restrict__ blockoffsets, int nblocks, int32_t* __restrict__ pg,
                  int32_t* __restrict__ ug, int32_t* __restrict__ vg, int32_t* __restrict__ wg,
                  int32_t* __restrict__ pcell, int32_t* __restrict__ ucell, int32_t* __restrict__ vcell,
                  int32_t* __restrict__ wcell) {
  const int64_t base = (int64_t)blockIdx.x * kScanCells + (int64_t)threadIdx.x * kScanCpt;
  const int64_t plane = (int64_t)ny * nz;
  unsigned fl[kScanCpt];
  int cnt[4] = {0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < kScanCpt; ++q) {
    const int64_t c = base + q;
    fl[q] = 0;
    if (c < ncell) {
      const int k = (int)(c % nz);
      const int64_t t = c / nz;
      const int j = (int)(t % ny);
      const int i = (int)(t / ny);
      fl[q] = cell_flags(img, nx, ny, nz, c, i, j, k);
      counts_of(fl[q], cnt[0], cnt[1], cnt[2], cnt[3]);
    }
  }
  int tot[4];
  block_exscan4(cnt, tot);
  int ip = cnt[0] + blockoffsets[0 * (int64_t)nblocks + blockIdx.x];
  int iu = cnt[1] + blockoffsets[1 * (int64_t)nblocks + blockIdx.x];
  int iv = cnt[2] + blockoffsets[2 * (int64_t)nblocks + blockIdx.x];
  int iw = cnt[3] + blockoffsets[3 * (int64_t)nblocks + blockIdx.x];
#pragma unroll
  for (int q = 0; q < kScanCpt; ++q) {
    const int64_t c = base + q;
    if (c >= ncell) break;
    const unsigned f = fl[q];
    if (f & F_V) { vg[c] = iv; vcell[iv] = (int32_t)c; ++iv; } else { vg[c] = -1; }
    if (f & F_W) { wg[c] = iw; wcell[iw] = (int32_t)c; ++iw; } else { wg[c] = -1; }
    if (f & F_UH) { ug[c] = iu; ucell[iu] = (int32_t)c; ++iu; } else { ug[c] = -1; }
    if (c / plane == nx - 1) {
      if (f & F_UO) { ug[c + plane] = iu; ucell[iu] = (int32_t)(c + plane); ++iu; }
      else { ug[c + plane] = -1; }
    }
    if (f & F_P) { pg[c] = ip; pcell[ip] = (int32_t)c; ++ip; } else { pg[c] = -1; }
  }
}

// Neighbour rows of each row in its own system's grid, order E W N S T B,
// the row itself where the neighbour node does not exist.
__global__ void __launch_bounds__(kBlock)
build_nbr_kernel(Geo g, int sys, const int32_t* __restrict__ cell, int n, int32_t* __restrict__ nbr,
                 unsigned long long* nnz) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= n) return;
  int i, j, k;
  g.decode(cell[row], i, j, k);
  int e, w, nn, s, t, b;
  if (sys == 0) {
    e = g.U(i + 1, j, k); w = g.U(i - 1, j, k); nn = g.U(i, j + 1, k);
    s = g.U(i, j - 1, k); t = g.U(i, j, k + 1); b = g.U(i, j, k - 1);
  } else if (sys == 1) {
    e = g.V(i + 1, j, k); w = g.V(i - 1, j, k); nn = g.V(i, j + 1, k);
    s = g.V(i, j - 1, k); t = g.V(i, j, k + 1); b = g.V(i, j, k - 1);
  } else if (sys == 2) {
    e = g.W(i + 1, j, k); w = g.W(i - 1, j, k); nn = g.W(i, j + 1, k);
    s = g.W(i, j - 1, k); t = g.W(i, j, k + 1); b = g.W(i, j, k - 1);
  } else {
    e = g.P(i + 1, j, k); w = g.P(i - 1, j, k); nn = g.P(i, j + 1, k);
    s = g.P(i, j - 1, k); t = g.P(i, j, k + 1); b = g.P(i, j, k - 1);
  }
  nbr[(int64_t)DE * n + row] = e > -1 ? e : row;
  nbr[(int64_t)DW * n + row] = w > -1 ? w : row;
  nbr[(int64_t)DN * n + row] = nn > -1 ? nn : row;
  nbr[(int64_t)DS * n + row] = s > -1 ? s : row;
  nbr[(int64_t)DT * n + row] = t > -1 ? t : row;
  nbr[(int64_t)DB * n + row] = b > -1 ? b : row;
  // stored entries of the row: diagonal + existing neighbours
  const unsigned long long cnt = 1ull + (e > -1) + (w > -1) + (nn > -1) + (s > -1) + (t > -1) + (b > -1);
  atomicAdd(nnz, cnt);
}

}  // namespace

hipError_t launch_scan_count(const uint8_t* img, int nx, int ny, int nz, ScanBuffers& sb, hipStream_t s) {
  hipLaunchKernelGGL(scan_count_kernel, dim3(sb.nblocks), dim3(kScanThreads), 0, s, img, nx, ny, nz,
                     sb.ncell, sb.blockcounts, sb.nblocks);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(scan_offsets_kernel, dim3(1), dim3(1024), 0, s, sb.blockcounts, sb.blockoffsets,
                     sb.totals, sb.nblocks);
  return hipGetLastError();
}

hipError_t launch_scan_write(const uint8_t* img, int nx, int ny, int nz, ScanBuffers& sb, int32_t* pg,
                             int32_t* ug, int32_t* vg, int32_t* wg, int32_t* pcell, int32_t* ucell,
                             int32_t* vcell, int32_t* wcell, hipStream_t s) {
  hipLaunchKernelGGL(scan_write_kernel, dim3(sb.nblocks), dim3(kScanThreads), 0, s, img, nx, ny, nz,
                     sb.ncell, sb.blockoffsets, sb.nblocks, pg, ug, vg, wg, pcell, ucell, vcell, wcell);
  return hipGetLastError();
}

hipError_t launch_build_nbr(const Geo& g, int sys, const int32_t* cell, int n, int32_t* nbr,
                            unsigned long long* nnz, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(build_nbr_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g, sys, cell,
                     n, nbr, nnz);
  return hipGetLastError();
}

}  // namespace ns3d
