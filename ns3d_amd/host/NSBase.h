// NSBase.h -- host facade of the reference's NavierStokesBase
// (src/NSBase.h:63-594) over the GPU context of include/ns3d.h.
//
// The members and methods keep the reference's names and meaning; fields and
// matrices live on the GPU (the drivers' u/u_temp/u_hat, pvector, pc and b are
// the context's velocity sets and vectors), so the methods take set
// selectors where the reference took Eigen vectors.  Scalar state (alpha,
// beta, RMR history, solver reports) lives here, as in the reference.
#pragma once

#include <chrono>
#include <cstdint>
#include <fstream>
#include <list>
#include <string>
#include <vector>

#include "../../include/ns3d.h"

namespace ns3d_host {

// Eigen::BiCGSTAB's observable state: iterations(), info(), error().
struct SolverReport {
  int64_t it = -1;  // -1: never ran (the reference then reads garbage)
  int info_ = 0;
  double err = 0;
  int64_t total = 0;  // summed iterations over the run
  int64_t work = 0;   // summed executed iterations (zero right-hand sides do none)
  int64_t iterations() const { return it; }
  int info() const { return info_; }
  double error() const { return err; }
  void set(const ns3d_solve_info& s) {
    it = s.iterations;
    info_ = s.info;
    err = s.error;
    total += s.iterations;
    work += s.work;
  }
};

// Per-phase wall clock of the outer loop (all device calls are synchronous).
enum Phase { PH_MOMENTUM = 0, PH_VELOCITY, PH_HAT, PH_PASM, PH_PSOLVE, PH_CORRECT, PH_DIAG, PH_OTHER, PH_COUNT };

class NavierStokesBase {
 public:
  // input parameters (NSBase.h:68-72)
  int maxitsolver, porevalue, gradient_direction, save_step, border;
  double dx, dy, dz, gradp, alpha0, beta0, alpha, beta, mu, rho, p1, p0, p_outlet, p_inlet;
  double vsolvertol, psolvertol, pcsolvertol, deltaRMR;
  // global state (NSBase.h:74-82)
  int nx, ny, nz;
  int64_t np, nu, nv, nw;
  double schemefactor;
  double pcmax, pcmin, pcnorm;
  double min_alpha, min_beta;
  double relax, maxconvdelta, RMR, old_RMR;
  std::list<double> avgRMR;
  SolverReport usolver, vsolver, wsolver, psolver;
  std::string outfilename, info_filename, log_filename, outdir;
  std::string pfieldfilename, vxfilename, vyfilename, vzfilename;
  // in-memory warm start (node-set vectors; empty: not given), used like the
  // field files by Initpvector / InitVelocity
  std::vector<double> warm_p, warm_u, warm_v, warm_w;
  std::ofstream info_file, log_file;
  bool verbose = true;        // reference-style progress on stdout
  bool progress = false;      // one line per outer iteration on stderr
  bool write_files = true;
  ns3d_ctx* ctx = nullptr;
  double phase[PH_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};

  // img: raw image (voxel == porevalue is fluid), C order, x slowest.
  NavierStokesBase(const std::vector<uint8_t>& raw, int nx_, int ny_, int nz_, const std::string& imagename,
                   const std::string& pfield, const std::string& vxf, const std::string& vyf,
                   const std::string& vzf, double dx, double dy, double dz, double vsolvertol, double psolvertol,
                   double pcsolvertol, int maxitsolver, double alpha, double beta, double p0, double gradp,
                   double mu, double rho, int porevalue, int gradient_direction, double deltaRMR, int save_step,
                   int border, int device, int verbose, bool write_files, const std::string& outdir);
  ~NavierStokesBase();
  NavierStokesBase(const NavierStokesBase&) = delete;
  NavierStokesBase& operator=(const NavierStokesBase&) = delete;

  void AssembleAllMomentum();
  void ComputeVelocityField(bool limit, double delta_v_max);
  void Compute_hat(int uc_set, int uold_set);
  void AssemblePressure(bool relax_p);
  void AssemblePressureCorrectionSIMPLEC();
  void AssemblePressureCorrectionSIMPLE();
  void ComputePressureCorrection();
  void ComputePressureField();
  void CorrectVelocityFieldSIMPLE(double delta_v_max);
  void CorrectVelocityFieldSIMPLEC(double delta_v_max);
  void CorrectVelocityFieldSIMPLER(int target_set, double p_0, double p_1);
  void CorrectPressureField();
  void ComputeDiv(double& MaxDeltaV, double& normdiv, double& velocityNorm, double& maxdiv);
  void ComputePermeability(double& Q, double& v_moy, double& K);
  void Initpvector();
  void InitVelocity();
  void RelaxControl(bool beta_cst);
  bool CheckConvergence(double res, double tol);
  void FieldsToImage(const std::string& filename);
  void CopyVelocity(int dst_set, int src_set);
  void ZeroFields();

  // fields on the node sets (host copies)
  std::vector<double> vector(int which) const;
  void say(const char* fmt, ...) const;
  void note(const char* fmt, ...) const;

 private:
  std::chrono::steady_clock::time_point t0_;
  void tick(Phase p);
};

// What RelaxControl reads and writes (the NavierStokesBase members of the
// same names), so that the step can run without a GPU context.
struct RelaxState {
  double alpha, beta, alpha0, min_alpha, min_beta, relax, maxconvdelta, deltaRMR, RMR, old_RMR;
  std::list<double> avgRMR;
};
// RelaxControl's events (a bit set): the residual jumped above maxconvdelta x
// its best; the window of 30 residuals was flat ("oscillating"); or rising.
enum RelaxEvent { RELAX_JUMP = 1, RELAX_FLAT = 2, RELAX_RISING = 4 };
// One RelaxControl step on s (NSBase.cxx:196-301); returns the events.
// avg1/avg2/delta of the window test are stored in win[3] when it ran.
int relax_control_step(RelaxState& s, bool beta_cst, double win[3], bool* window_ran);

// Throws std::runtime_error with ns3d_last_error() when st != NS3D_OK.
void check(ns3d_status st, const char* what);

}  // namespace ns3d_host
