// kernels.h -- host-callable launchers of the HIP kernels (one per .hip file
// group) and the device-side BiCGSTAB state.  Included by the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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
struct BState {
  double rho, rho_old, alpha, w;
  double r0_sqnorm, rhs_sqnorm, tol2, tol;
  double rr, r0v, ts, tt;
  long long i, nexec, maxit, restarts;
  int status;
  int pad;
  unsigned int counter[8];
};
struct BWork {
  double* r;
  double* r0;
  double* p[2];
  double* v[2];
  double* t;
  double* partials;   // [4][max blocks]
  BState* st;
};
// Launchers take the matrix view S (coef/nbr/invdiag) and the work buffers.
hipError_t launch_bicg_init(SysView S, const double* x, const double* b, BWork w, double tol,
                            long long maxit, hipStream_t s);
hipError_t launch_bicg_restart(SysView S, const double* x, const double* b, BWork w, hipStream_t s);
hipError_t launch_bicg_A(SysView S, BWork w, hipStream_t s);
hipError_t launch_bicg_B(SysView S, BWork w, hipStream_t s);
hipError_t launch_bicg_C(SysView S, double* x, BWork w, hipStream_t s);
hipError_t launch_invdiag(SysView S, hipStream_t s);

// ---- fields, corrections, diagnostics (kernels_field.hip) -------------------
hipError_t launch_fill(double* x, int64_t n, double v, hipStream_t s);
hipError_t launch_init_pressure_linear(const Geo& g, const int32_t* pcell, int np, double* p, double p0,
                                       double p1, hipStream_t s);
hipError_t launch_clamp(double* x, const double* temp, int64_t n, double dvmax, hipStream_t s);
hipError_t launch_correct_velocity(const Geo& g, SysView P, SysView U, SysView V, SysView W, int kind,
                                   const double* q, double* tu, double* tv, double* tw, const double* uh,
                                   const double* vh, const double* wh, double dvmax, double p_0, double p_1,
                                   double dx, double dy, double dz, hipStream_t s);
hipError_t launch_correct_pressure(double* p, const double* pc, int64_t n, double beta, double* partials,
                                   unsigned int* counter, double* result, hipStream_t s);
hipError_t launch_compute_div(const Geo& g, SysView P, const double* u, const double* v, const double* w,
                              double dx, double dy, double dz, double* partials, unsigned int* counter,
                              double* result, hipStream_t s);
hipError_t launch_permeability(const Geo& g, const double* u, const double* p, int border, double* partials,
                               unsigned int* counter, double* result, hipStream_t s);

}  // namespace ns3d
