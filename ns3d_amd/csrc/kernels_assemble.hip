// kernels_assemble.hip -- matrix-free assembly of the staggered momentum and
// pressure(-correction) rows, the momentum right-hand sides and the SIMPLER
// pseudo-velocities.  One thread per row; every coefficient is computed with
// the same expression tree as the reference (compiled without FP contraction)
// so the coefficient arrays are bit-identical to the CPU path's.
#include "kernels.h"

namespace ns3d {

namespace {

// std::min(F, 0.0) / std::max(F, 0.0) with the standard library's operand
// order (matters for NaN and signed zero only).
__device__ __forceinline__ double min0(double F) { return (0.0 < F) ? 0.0 : F; }
__device__ __forceinline__ double max0(double F) { return (F < 0.0) ? 0.0 : F; }

// Face between the row's own node a and the same-component neighbour bn
// ("axial" faces).  NSBase.cxx:917-933, :945-961 (u, x faces), :1130-1175
// (v, y faces), :1365-1409 (w, z faces).  schemefactor = 0 makes the Peclet
// test "F != 0": upwind whenever both velocities share a sign, else centred.
__device__ __forceinline__ void axial(bool plus, double a, double bn, const Phys& ph, double id, double id2,
                                      double h, double& d, double& off) {
  double F;
  if (plus) {
    if (a > 0 && bn > 0) F = -a * ph.rho * id;
    else if (a < 0 && bn < 0) F = -bn * ph.rho * id;
    else F = 0;
  } else {
    if (a < 0 && bn < 0) F = a * ph.rho * id;
    else if (a > 0 && bn > 0) F = bn * ph.rho * id;
    else F = 0;
  }
  if (fabs(F) * h * h > 2 * ph.mu * ph.sf) {
    d += -ph.mu * id2 + min0(F);
    off = ph.mu * id2 + max0(F);
  } else {
    const double m = (a + bn) * 0.5;
    const double Fd = plus ? -m * ph.rho * id : m * ph.rho * id;
    d += -ph.mu * id2 + Fd * 0.5;
    off = ph.mu * id2 + Fd * 0.5;
  }
}

// Face with a transverse advecting velocity m (average of the two velocities
// of the other component on the face).  rho_first selects the reference's two
// spellings of the Peclet test, |m|*rho*h or |m|*h*rho.
__device__ __forceinline__ void trans(bool plus, double m, const Phys& ph, double id, double id2, double h,
                                      bool rho_first, double& d, double& off) {
  const double F = plus ? -m * ph.rho * id : m * ph.rho * id;
  const double test = rho_first ? fabs(m) * ph.rho * h : fabs(m) * h * ph.rho;
  if (test > 2 * ph.mu * ph.sf) {
    d += -ph.mu * id2 + min0(F);
    off = ph.mu * id2 + max0(F);
  } else {
    d += -ph.mu * id2 + F * 0.5;
    off = ph.mu * id2 + F * 0.5;
  }
}

// Sum of the stored entries of a row in ascending column order, starting from
// 0 (Eigen's row(r).sum() on a RowMajor matrix with sorted inner indices).
// Absent entries carry a key larger than any row and value 0.
__device__ __forceinline__ double rowsum_sorted(int (&key)[7], double (&val)[7]) {
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = 0; b < 6 - a; ++b) {
      const bool sw = key[b] > key[b + 1];
      const int k0 = sw ? key[b + 1] : key[b];
      const int k1 = sw ? key[b] : key[b + 1];
      const double v0 = sw ? val[b + 1] : val[b];
      const double v1 = sw ? val[b] : val[b + 1];
      key[b] = k0; key[b + 1] = k1; val[b] = v0; val[b + 1] = v1;
    }
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 7; ++q) s += val[q];
  return s;
}

__device__ __forceinline__ void store_row(SysView& S, int row, double d, const double (&off)[6],
                                          const bool (&has)[6], const int (&nb)[6]) {
  const int64_t n = S.n;
  S.coef[row] = d;
#pragma unroll
  for (int q = 0; q < 6; ++q) S.coef[(1 + q) * n + row] = off[q];
  S.invdiag[row] = (d != 0.0) ? 1.0 / d : 1.0;
  if (S.rowsum) {
    int key[7];
    double val[7];
    key[0] = row;
    val[0] = d;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      key[q + 1] = has[q] ? nb[q] : 0x7fffffff;
      val[q + 1] = has[q] ? off[q] : 0.0;
    }
    S.rowsum[row] = rowsum_sorted(key, val);
  }
}

// AssembleMomentum_x (NSBase.cxx:886-1057): rows of u.
__global__ void __launch_bounds__(kBlock)
mom_x_kernel(Geo g, SysView S, Vel3 vc, Phys ph) {
  const int row = logical_block() * kBlock + threadIdx.x;
  if (row >= S.n) return;
  int i, j, k;
  g.decode(S.cell[row], i, j, k);
  const int64_t n = S.n;
  int nb[6];
  bool has[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) { nb[q] = S.nbr[q * n + row]; has[q] = nb[q] != row; }
  const double idx2 = 1 / (ph.dx * ph.dx), idy2 = 1 / (ph.dy * ph.dy), idz2 = 1 / (ph.dz * ph.dz);
  const double idx = 1 / ph.dx, idy = 1 / ph.dy, idz = 1 / ph.dz;
  const double* uc = vc.u;
  const double* vcv = vc.v;
  const double* wc = vc.w;
  const double a = uc[row];
  double d = 0.0;
  double off[6] = {0, 0, 0, 0, 0, 0};
  // E
  if (i + 1 < g.nx + 1) {
    if (has[DE]) axial(true, a, uc[nb[DE]], ph, idx, idx2, ph.dx, d, off[DE]);
    else d += -ph.mu * idx2 - (0.25 * ph.rho * idx) * (a);
  } else {
    d += -(ph.rho * idx) * (a);
  }
  // W
  if (i - 1 >= 0) {
    if (has[DW]) axial(false, a, uc[nb[DW]], ph, idx, idx2, ph.dx, d, off[DW]);
    else d += -ph.mu * idx2 + (0.25 * ph.rho * idx) * (a);
  } else {
    d += (ph.rho * idx) * (a);
  }
  // N
  if (j + 1 < g.ny) {
    if (has[DN]) {
      const double vm = (svv(vcv, g.V(i - 1, j + 1, k)) + svv(vcv, g.V(i, j + 1, k))) * 0.5;
      trans(true, vm, ph, idy, idy2, ph.dy, true, d, off[DN]);
    } else {
      d += -2 * ph.mu * idy2;
    }
  } else {
    d += -2 * ph.mu * idy2;
  }
  // S
  if (j - 1 >= 0) {
    if (has[DS]) {
      const double vm = (svv(vcv, g.V(i - 1, j, k)) + svv(vcv, g.V(i, j, k))) * 0.5;
      trans(false, vm, ph, idy, idy2, ph.dy, false, d, off[DS]);
    } else {
      d += -2 * ph.mu * idy2;
    }
  } else {
    d += -2 * ph.mu * idy2;
  }
  if (g.nz > 1) {
    // T
    if (k + 1 < g.nz) {
      if (has[DT]) {
        const double wm = (svv(wc, g.W(i - 1, j, k + 1)) + svv(wc, g.W(i, j, k + 1))) * 0.5;
        trans(true, wm, ph, idz, idz2, ph.dz, true, d, off[DT]);
      } else {
        d += -2 * ph.mu * idz2;
      }
    } else {
      d += -2 * ph.mu * idz2;
    }
    // B
    if (k - 1 >= 0) {
      if (has[DB]) {
        const double wm = (svv(wc, g.W(i - 1, j, k)) + svv(wc, g.W(i, j, k))) * 0.5;
        trans(false, wm, ph, idz, idz2, ph.dz, true, d, off[DB]);
      } else {
        d += -2 * ph.mu * idz2;
      }
    } else {
      d += -2 * ph.mu * idz2;
    }
  } else {
    has[DT] = has[DB] = false;
  }
  d /= ph.alpha;
  store_row(S, row, d, off, has, nb);
}

// AssembleMomentum_y (NSBase.cxx:1059-1250): rows of v.
__global__ void __launch_bounds__(kBlock)
mom_y_kernel(Geo g, SysView S, Vel3 vc, Phys ph) {
  const int row = logical_block() * kBlock + threadIdx.x;
  if (row >= S.n) return;
  int i, j, k;
  g.decode(S.cell[row], i, j, k);
  const int64_t n = S.n;
  int nb[6];
  bool has[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) { nb[q] = S.nbr[q * n + row]; has[q] = nb[q] != row; }
  const double idx2 = 1 / (ph.dx * ph.dx), idy2 = 1 / (ph.dy * ph.dy), idz2 = 1 / (ph.dz * ph.dz);
  const double idx = 1 / ph.dx, idy = 1 / ph.dy, idz = 1 / ph.dz;
  const double* uc = vc.u;
  const double* vcv = vc.v;
  const double* wc = vc.w;
  const double a = vcv[row];
  double d = 0.0;
  double off[6] = {0, 0, 0, 0, 0, 0};
  // E
  if (i + 1 < g.nx) {
    if (has[DE]) {
      const double um = (svv(uc, g.U(i + 1, j - 1, k)) + svv(uc, g.U(i + 1, j, k))) * 0.5;
      trans(true, um, ph, idx, idx2, ph.dx, false, d, off[DE]);
    } else {
      d += -2 * ph.mu * idx2;
    }
  } else {
    d += -(0.5 * ph.rho * idx) * (svv(uc, g.U(i + 1, j - 1, k)) + svv(uc, g.U(i + 1, j, k)));
  }
  // W
  if (i - 1 >= 0) {
    if (has[DW]) {
      const double um = (svv(uc, g.U(i, j - 1, k)) + svv(uc, g.U(i, j, k))) * 0.5;
      trans(false, um, ph, idx, idx2, ph.dx, false, d, off[DW]);
    } else {
      d += -2 * ph.mu * idx2;
    }
  } else {
    d += -ph.mu * idx2 + (0.25 * ph.rho * idx) * (svv(uc, g.U(i, j - 1, k)) + svv(uc, g.U(i, j, k)));
  }
  // N
  if (j + 1 < g.ny) {
    if (has[DN]) axial(true, a, vcv[nb[DN]], ph, idy, idy2, ph.dy, d, off[DN]);
    else d += -ph.mu * idy2 - (0.25 * ph.rho * idy) * (a);
  } else {
    d += -ph.mu * idy2 - (0.25 * ph.rho * idy) * (a);
  }
  // S
  if (j - 1 >= 0) {
    if (has[DS]) axial(false, a, vcv[nb[DS]], ph, idy, idy2, ph.dy, d, off[DS]);
    else d += -ph.mu * idy2 + (0.25 * ph.rho * idy) * (a);
  } else {
    d += -ph.mu * idy2 + (0.25 * ph.rho * idy) * (a);
  }
  if (g.nz > 1) {
    if (k + 1 < g.nz) {
      if (has[DT]) {
        const double wm = (svv(wc, g.W(i, j - 1, k + 1)) + svv(wc, g.W(i, j, k + 1))) * 0.5;
        trans(true, wm, ph, idz, idz2, ph.dz, true, d, off[DT]);
      } else {
        d += -2 * ph.mu * idz2;
      }
    } else {
      d += -2 * ph.mu * idz2;
    }
    if (k - 1 >= 0) {
      if (has[DB]) {
        const double wm = (svv(wc, g.W(i, j - 1, k)) + svv(wc, g.W(i, j, k))) * 0.5;
        trans(false, wm, ph, idz, idz2, ph.dz, true, d, off[DB]);
      } else {
        d += -2 * ph.mu * idz2;
      }
    } else {
      d += -2 * ph.mu * idz2;
    }
  } else {
    has[DT] = has[DB] = false;
  }
  d /= ph.alpha;
  store_row(S, row, d, off, has, nb);
}

// AssembleMomentum_z (NSBase.cxx:1252-1442): rows of w (no nz>1 guard).
__global__ void __launch_bounds__(kBlock)
mom_z_kernel(Geo g, SysView S, Vel3 vc, Phys ph) {
  const int row = logical_block() * kBlock + threadIdx.x;
  if (row >= S.n) return;
  int i, j, k;
  g.decode(S.cell[row], i, j, k);
  const int64_t n = S.n;
  int nb[6];
  bool has[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) { nb[q] = S.nbr[q * n + row]; has[q] = nb[q] != row; }
  const double idx2 = 1 / (ph.dx * ph.dx), idy2 = 1 / (ph.dy * ph.dy), idz2 = 1 / (ph.dz * ph.dz);
  const double idx = 1 / ph.dx, idy = 1 / ph.dy, idz = 1 / ph.dz;
  const double* uc = vc.u;
  const double* vcv = vc.v;
  const double* wc = vc.w;
  const double a = wc[row];
  double d = 0.0;
  double off[6] = {0, 0, 0, 0, 0, 0};
  // E
  if (i + 1 < g.nx) {
    if (has[DE]) {
      const double um = (svv(uc, g.U(i + 1, j, k - 1)) + svv(uc, g.U(i + 1, j, k))) * 0.5;
      trans(true, um, ph, idx, idx2, ph.dx, false, d, off[DE]);
    } else {
      d += -2 * ph.mu * idx2;
    }
  } else {
    d += -(0.5 * ph.rho * idx) * (svv(uc, g.U(i + 1, j, k - 1)) + svv(uc, g.U(i + 1, j, k)));
  }
  // W
  if (i - 1 >= 0) {
    if (has[DW]) {
      const double um = (svv(uc, g.U(i, j, k - 1)) + svv(uc, g.U(i, j, k))) * 0.5;
      trans(false, um, ph, idx, idx2, ph.dx, false, d, off[DW]);
    } else {
      d += -2 * ph.mu * idx2;
    }
  } else {
    d += -ph.mu * idx2 + (0.25 * ph.rho * idx) * (svv(uc, g.U(i, j, k - 1)) + svv(uc, g.U(i, j, k)));
  }
  // N
  if (j + 1 < g.ny) {
    if (has[DN]) {
      const double vm = (svv(vcv, g.V(i, j + 1, k - 1)) + svv(vcv, g.V(i, j + 1, k))) * 0.5;
      trans(true, vm, ph, idy, idy2, ph.dy, true, d, off[DN]);
    } else {
      d += -2 * ph.mu * idy2;
    }
  } else {
    d += -2 * ph.mu * idy2;
  }
  // S
  if (j - 1 >= 0) {
    if (has[DS]) {
      const double vm = (svv(vcv, g.V(i, j, k - 1)) + svv(vcv, g.V(i, j, k))) * 0.5;
      trans(false, vm, ph, idy, idy2, ph.dy, true, d, off[DS]);
    } else {
      d += -2 * ph.mu * idy2;
    }
  } else {
    d += -2 * ph.mu * idy2;
  }
  // T
  if (k + 1 < g.nz) {
    if (has[DT]) axial(true, a, wc[nb[DT]], ph, idz, idz2, ph.dz, d, off[DT]);
    else d += -ph.mu * idz2 - (0.25 * ph.rho * idz) * (a);
  } else {
    d += -ph.mu * idz2 - (0.25 * ph.rho * idz) * (a);
  }
  // B
  if (k - 1 >= 0) {
    if (has[DB]) axial(false, a, wc[nb[DB]], ph, idz, idz2, ph.dz, d, off[DB]);
    else d += -ph.mu * idz2 + (0.25 * ph.rho * idz) * (a);
  } else {
    d += -ph.mu * idz2 + (0.25 * ph.rho * idz) * (a);
  }
  d /= ph.alpha;
  store_row(S, row, d, off, has, nb);
}

// Pressure-type rows: AssemblePressure (kind 0, NSBase.cxx:478-610),
// AssemblePressureCorrectionSIMPLE (kind 1, :734-851) and SIMPLEC (kind 2,
// :612-732).  a_f is the face row's diagonal (kinds 0, 1) or its row sum
// (kind 2).  Accumulation order: x-, x+, y-, y+, z-, z+, inlet, outlet.
__global__ void __launch_bounds__(kBlock)
pressure_kernel(Geo g, SysView P, SysView U, SysView V, SysView W, Vel3 vc, const double* __restrict__ pold,
                double* __restrict__ b, int kind, int relax, Phys ph) {
  const int row = logical_block() * kBlock + threadIdx.x;
  if (row >= P.n) return;
  int i, j, k;
  g.decode(P.cell[row], i, j, k);
  const int64_t n = P.n;
  const double dx2 = ph.dx * ph.dx, dy2 = ph.dy * ph.dy, dz2 = ph.dz * ph.dz;
  const double idx = 1 / ph.dx, idy = 1 / ph.dy, idz = 1 / ph.dz;
  const bool rs = (kind == 2);
  const double* au = rs ? U.rowsum : U.coef;
  const double* av = rs ? V.rowsum : V.coef;
  const double* aw = rs ? W.rowsum : W.coef;
  double d = 0.0, bb = 0.0;
  double off[6] = {0, 0, 0, 0, 0, 0};
  int q, f;
  if (i > 0) {
    q = P.nbr[DW * n + row]; f = g.U(i, j, k);
    if (q != row && f > -1) {
      d += 1 / ((au[f]) * dx2);
      off[DW] = -1 / ((au[f]) * dx2);
      bb -= vc.u[f] * idx;
    }
  }
  if (i < g.nx - 1) {
    q = P.nbr[DE * n + row]; f = g.U(i + 1, j, k);
    if (q != row && f > -1) {
      d += 1 / ((au[f]) * dx2);
      off[DE] = -1 / ((au[f]) * dx2);
      bb += vc.u[f] * idx;
    }
  }
  if (j > 0) {
    q = P.nbr[DS * n + row]; f = g.V(i, j, k);
    if (q != row && f > -1) {
      d += 1 / ((av[f]) * dy2);
      off[DS] = -1 / ((av[f]) * dy2);
      bb -= vc.v[f] * idy;
    }
  }
  if (j < g.ny - 1) {
    q = P.nbr[DN * n + row]; f = g.V(i, j + 1, k);
    if (q != row && f > -1) {
      d += 1 / ((av[f]) * dy2);
      off[DN] = -1 / ((av[f]) * dy2);
      bb += vc.v[f] * idy;
    }
  }
  if (k > 0 && g.nz > 0) {
    q = P.nbr[DB * n + row]; f = g.W(i, j, k);
    if (q != row && f > -1) {
      d += 1 / ((aw[f]) * dz2);
      off[DB] = -1 / ((aw[f]) * dz2);
      bb -= vc.w[f] * idz;
    }
  }
  if (k < g.nz - 1 && g.nz > 0) {
    q = P.nbr[DT * n + row]; f = g.W(i, j, k + 1);
    if (q != row && f > -1) {
      d += 1 / ((aw[f]) * dz2);
      off[DT] = -1 / ((aw[f]) * dz2);
      bb += vc.w[f] * idz;
    }
  }
  if (i == 0 && (f = g.U(i, j, k)) > -1) {
    d += 1 / ((au[f]) * dx2);
    if (kind == 0) bb -= vc.u[f] * idx - ph.p0 / (U.coef[f] * dx2);
    else bb -= vc.u[f] * idx;
  }
  if (i == g.nx - 1 && (f = g.U(i + 1, j, k)) > -1) {
    d += 1 / ((au[f]) * dx2);
    if (kind == 0) bb += vc.u[f] * idx + ph.p1 / (U.coef[f] * dx2);
    else bb += vc.u[f] * idx;
  }
  if (kind == 0 && relax) {
    d /= ph.beta;
    bb += pold[row] * d * (1 - ph.beta);
  }
  b[row] = bb;
  P.coef[row] = d;
#pragma unroll
  for (int qq = 0; qq < 6; ++qq) P.coef[(1 + qq) * n + row] = off[qq];
  P.invdiag[row] = (d != 0.0) ? 1.0 / d : 1.0;
}

// Assemble_RHS_u/v/w (NSBase.cxx:1445-1546).
__global__ void __launch_bounds__(kBlock)
rhs_kernel(int sys, Geo g, SysView S, const double* __restrict__ p, const double* __restrict__ xold,
           double* __restrict__ b, Phys ph) {
  const int row = logical_block() * kBlock + threadIdx.x;
  if (row >= S.n) return;
  int i, j, k;
  g.decode(S.cell[row], i, j, k);
  double val;
  if (sys == 0) {
    const double idx = 1 / ph.dx;
    if (i == 0) val = (p[g.P(i, j, k)] - ph.p0) * idx;
    else if (i == g.nx) val = (ph.p1 - p[g.P(i - 1, j, k)]) * idx;
    else val = (p[g.P(i, j, k)] - p[g.P(i - 1, j, k)]) * idx;
    val += xold[row] * S.coef[row] * (1 - ph.alpha);
  } else if (sys == 1) {
    const double idy = 1 / ph.dy;
    val = (p[g.P(i, j, k)] - p[g.P(i, j - 1, k)]) * idy + xold[row] * S.coef[row] * (1 - ph.alpha);
  } else {
    const double idz = 1 / ph.dz;
    val = (p[g.P(i, j, k)] - p[g.P(i, j, k - 1)]) * idz + xold[row] * S.coef[row] * (1 - ph.alpha);
  }
  b[row] = val;
}

// Compute_u_hat / v_hat / w_hat (NSBase.cxx:1548-1755): neighbour order E, W,
// N, S, T, B over the existing neighbours (z only when nz > 1 for u and v).
__global__ void __launch_bounds__(kBlock)
hat_kernel(int sys, int nz, SysView S, const double* __restrict__ xc, const double* __restrict__ xold,
           double* __restrict__ hat, double alpha) {
  const int row = logical_block() * kBlock + threadIdx.x;
  if (row >= S.n) return;
  const int64_t n = S.n;
  const bool zok = (sys == 2) || (nz > 1);
  double h = 0.0;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    if (q >= DT && !zok) break;
    const int nb = S.nbr[q * n + row];
    if (nb != row) h -= S.coef[(1 + q) * n + row] * xc[nb];
  }
  hat[row] = h / (S.coef[row]) + xold[row] * (1 - alpha);
}

}  // namespace

hipError_t launch_assemble_momentum(int sys, const Geo& g, SysView S, Vel3 vc, Phys ph, hipStream_t s) {
  if (S.n <= 0) return hipSuccess;
  const int nb = padded_blocks(S.n);
  if (sys == 0) hipLaunchKernelGGL(mom_x_kernel, dim3(nb), dim3(kBlock), 0, s, g, S, vc, ph);
  else if (sys == 1) hipLaunchKernelGGL(mom_y_kernel, dim3(nb), dim3(kBlock), 0, s, g, S, vc, ph);
  else hipLaunchKernelGGL(mom_z_kernel, dim3(nb), dim3(kBlock), 0, s, g, S, vc, ph);
  return hipGetLastError();
}

hipError_t launch_assemble_pressure(const Geo& g, SysView P, SysView U, SysView V, SysView W, Vel3 vc,
                                    const double* pold, double* b, int kind, int relax, Phys ph,
                                    hipStream_t s) {
  if (P.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(pressure_kernel, dim3(padded_blocks(P.n)), dim3(kBlock), 0, s, g, P, U, V, W, vc, pold,
                     b, kind, relax, ph);
  return hipGetLastError();
}

hipError_t launch_rhs(int sys, const Geo& g, SysView S, const double* p, const double* xold, double* b,
                      Phys ph, hipStream_t s) {
  if (S.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(rhs_kernel, dim3(padded_blocks(S.n)), dim3(kBlock), 0, s, sys, g, S, p, xold, b, ph);
  return hipGetLastError();
}

hipError_t launch_hat(int sys, const Geo& g, SysView S, const double* xc, const double* xold, double* hat,
                      Phys ph, hipStream_t s) {
  if (S.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(hat_kernel, dim3(padded_blocks(S.n)), dim3(kBlock), 0, s, sys, g.nz, S, xc, xold, hat,
                     ph.alpha);
  return hipGetLastError();
}

}  // namespace ns3d
