// drivers.cpp -- SIMPLEC (SIMPLEC.cxx:73-232), SIMPLER (SIMPLER.cxx:75-270),
// SIMPLERR (SIMPLERR.cxx:75-280) and IDEAL (IDEAL.cxx:75-281) outer loops.
//
// The four algorithms share their initialisation, diagnostics and RES-row
// logic; what differs is the body of one outer iteration, written out per
// algorithm below in the reference's call order.
#include "drivers.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>

namespace ns3d_host {

namespace {

const char* kResHeader =
    "iteration, Q (m³/s), <u_out> (m/s), Max Rel Mass Res., K (m²), gradP (Pa/m), ||Div_u||, "
    "Lref*(max DIV)/<u_out>, <||u||>, max DIV, it_u,it_v,it_w,it_pc";

struct Diag {
  double Q = 0, v_moy = 0, K = 0, normdiv = 0, reldiv = 0, velocityNorm = 0, maxdiv = 0, MaxDeltaV = 0;
};

// Solver iteration counts as the RES row prints them; a solver that has not
// run yet (the reference prints uninitialised memory) reads 0.
double its(const SolverReport& s) { return s.iterations() < 0 ? 0.0 : (double)s.iterations(); }

// it, Q, v_moy, RMR, K, gradP, normdiv, reldiv, <||u||>, maxdiv, it_u..it_pc
// (e.g. SIMPLEC.cxx:173): default stream precision, ", " before maxdiv and it_u.
void res_row(NavierStokesBase& b, RunLog& log, int it, const Diag& d) {
  const double gradP = (b.p_inlet - b.p_outlet) / (b.dx * (b.nx - 2 * b.border));
  b.log_file << it << ',' << d.Q << ',' << d.v_moy << ',' << b.RMR << ',' << d.K << ',' << gradP << ','
             << d.normdiv << ',' << d.reldiv << ',' << d.velocityNorm << ", " << d.maxdiv << ", "
             << (int64_t)its(b.usolver) << ',' << (int64_t)its(b.vsolver) << ',' << (int64_t)its(b.wsolver) << ','
             << (int64_t)its(b.psolver) << std::endl;
  log.rows.push_back(ResRow{(double)it, d.Q, d.v_moy, b.RMR, d.K, gradP, d.normdiv, d.reldiv, d.velocityNorm,
                            d.maxdiv, its(b.usolver), its(b.vsolver), its(b.wsolver), its(b.psolver)});
}

void init_common(NavierStokesBase& b, const char* name, const char* itname, int maxit, double simpletol) {
  b.say("Initializing %s algorithm\n", name);
  b.info_file << "Initializing " << name << " algorithm" << std::endl;
  b.ZeroFields();
  b.Initpvector();
  b.InitVelocity();
  b.info_file << "Max number of " << itname << " iterations: " << maxit << std::endl;
  b.say("Stopping criterion for relative residual: %g\n", simpletol);
  b.log_file << kResHeader << std::endl;
  b.say("Computing initial velocity field\n");
}

void diagnostics(NavierStokesBase& b, Diag& d) {
  b.ComputePermeability(d.Q, d.v_moy, d.K);
  b.ComputeDiv(d.MaxDeltaV, d.normdiv, d.velocityNorm, d.maxdiv);
  b.RMR = d.MaxDeltaV / std::abs(d.v_moy);
  d.reldiv = (b.nx * b.dx + b.ny * b.dy + b.nz * b.dz) * d.maxdiv / (3 * d.v_moy);
  b.say("Permeability K (m²): %g\n", d.K);
  b.info_file << "Permeability K (m²): " << d.K << std::endl;
  b.say("Maximum value of the velocity field divergence : %g\n", d.maxdiv);
  b.say("Relative Mass Residual, Lref*(max DIV)/<u_out>, ||Div u|| %g, %g, %g\n", b.RMR, d.reldiv, d.normdiv);
  b.note("  outer it: K %.10e RMR %.3e its u/v/w/p %lld/%lld/%lld/%lld alpha %g beta %g\n", d.K, b.RMR,
         (long long)b.usolver.it, (long long)b.vsolver.it, (long long)b.wsolver.it, (long long)b.psolver.it, b.alpha,
         b.beta);
}

void track_old_rmr(NavierStokesBase& b, int it) {
  if (it > 1) {
    if (b.RMR < b.old_RMR) b.old_RMR = b.RMR;
  } else {
    b.old_RMR = b.RMR;
  }
}

bool converged(NavierStokesBase& b, int it, double simpletol, const Diag& d) {
  if (it > 1 && b.CheckConvergence(b.RMR, simpletol)) {
    b.say("Permeability K (m²): %g\n", d.K);
    b.info_file << "Permeability K (m²): " << d.K << std::endl;
    return true;
  }
  return false;
}

void save_snapshot(NavierStokesBase& b, int it, int save_step, int& savecounter) {
  if (it % save_step == 0 && it > 0) {
    savecounter += 1;
    if (savecounter > 3) savecounter = 1;
    const std::string stem = b.outfilename.substr(0, b.outfilename.length() - 4);
    b.FieldsToImage(stem + "-SAVE-" + std::to_string(savecounter) + ".vti");
  }
}

void finish(NavierStokesBase& b, RunLog& log, const char* name, int it, time_t start, const Diag& d,
            double seconds) {
  b.info_file << "Norm of the velocity field divergence: " << d.normdiv << std::endl;
  b.say("Norm of the velocity field divergence: %g\n", d.normdiv);
  b.info_file << "Relative divergence of the velocity field: " << d.reldiv << std::endl;
  b.say("Relative divergence of the velocity field: %g\n", d.reldiv);
  b.FieldsToImage(b.outfilename.substr(0, b.outfilename.length() - 4) + ".vti");
  const time_t end = time(0);
  b.say("%s terminated in %d iteration and %g seconds\n", name, it, difftime(end, start));
  b.info_file << name << " terminated in " << it << " iteration and " << difftime(end, start) << " seconds"
              << std::endl;
  log.iterations = it;
  log.K = d.K; log.Q = d.Q; log.v_moy = d.v_moy; log.RMR = b.RMR;
  log.normdiv = d.normdiv; log.reldiv = d.reldiv; log.maxdiv = d.maxdiv;
  log.seconds = seconds;
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

SIMPLECSOLVER::SIMPLECSOLVER(double tol, int mx) : maxit(mx), simpletol(tol) { start = time(0); }
SIMPLERSOLVER::SIMPLERSOLVER(double tol, int mx, int n1) : N1(n1), maxit(mx), simpletol(tol) { start = time(0); }
SIMPLERRSOLVER::SIMPLERRSOLVER(double tol, int mx, int n1) : N1(n1), maxit(mx), simpletol(tol) { start = time(0); }
IDEALSOLVER::IDEALSOLVER(double tol, int mx, int n1, int n2) : N1(n1), N2(n2), maxit(mx), simpletol(tol) {
  start = time(0);
}

// SIMPLEC.cxx:73-232
void SIMPLECSOLVER::Solve(NavierStokesBase& B) {
  const double t0 = now();
  Diag d;
  save_step = B.save_step;
  int savecounter = 0;
  init_common(B, "SIMPLEC", "SIMPLE", maxit, simpletol);
  B.alpha = B.alpha0;
  B.AssembleAllMomentum();
  B.ComputeVelocityField(false, 0);
  B.ComputePermeability(d.Q, d.v_moy, d.K);
  B.say("Initial Permeability K (m²): %g\n", d.K);
  B.info_file << "Initial Permeability K (m²): " << d.K << std::endl;
  int it;
  for (it = 0; it < maxit; ++it) {
    B.say("----------------------------------\nIteration %d  ---------------------\n", it);
    B.info_file << "Iteration " << it << "  ------------" << std::endl;
    if (it > 0) {
      B.AssembleAllMomentum();
      B.ComputeVelocityField(true, d.v_moy * std::min(1., B.alpha * 1.5));
    }
    diagnostics(B, d);
    track_old_rmr(B, it);
    res_row(B, log, it, d);
    if (converged(B, it, simpletol, d)) { log.converged = true; break; }
    B.AssemblePressureCorrectionSIMPLEC();
    B.ComputePressureCorrection();
    B.CorrectVelocityFieldSIMPLEC(d.v_moy * B.alpha);
    B.CorrectPressureField();
    B.ComputeDiv(d.MaxDeltaV, d.normdiv, d.velocityNorm, d.maxdiv);
    B.say("Max div after correction: %g\n", d.maxdiv);
    save_snapshot(B, it, save_step, savecounter);
    B.RelaxControl(true);
  }
  finish(B, log, "SIMPLEC", it, start, d, now() - t0);
}

// Step 1 of SIMPLER/SIMPLERR/IDEAL (SIMPLER.cxx:161-181): N1 pressure
// predictions from the pseudo-velocities, correcting the temp velocities.
static void pressure_prediction(NavierStokesBase& B, int N1) {
  B.CopyVelocity(NS3D_SET_TEMP, NS3D_SET_VEL);
  B.say("STEP 1 -- Updating pressure field\n");
  for (int sub1 = 0; sub1 < N1; sub1++) {
    B.say("inner iteration %d\n", sub1);
    B.Compute_hat(NS3D_SET_TEMP, NS3D_SET_VEL);
    B.AssemblePressure(true);
    B.ComputePressureField();
    B.CorrectVelocityFieldSIMPLER(NS3D_SET_TEMP, B.p0, B.p1);
  }
}

// SIMPLER.cxx:75-270
void SIMPLERSOLVER::Solve(NavierStokesBase& B) {
  const double t0 = now();
  Diag d;
  save_step = B.save_step;
  int savecounter = 0;
  init_common(B, "SIMPLER", "SIMPLE", maxit, simpletol);
  B.AssembleAllMomentum();
  B.ComputeVelocityField(false, 0);
  B.ComputePermeability(d.Q, d.v_moy, d.K);
  B.say("Initial Permeability K (m²): %g\n", d.K);
  B.info_file << "Initial Permeability K (m²): " << d.K << std::endl;
  B.alpha = B.alpha0;
  int it;
  for (it = 0; it < maxit; ++it) {
    B.say("----------------------------------\nIteration %d  ---------------------\n", it);
    B.info_file << "Iteration " << it << "  ------------" << std::endl;
    B.AssembleAllMomentum();
    pressure_prediction(B, N1);
    B.say("STEP 2 -- Updating velocity field\n");
    B.ComputeVelocityField(true, d.v_moy * std::min(1., B.alpha * 1.5));
    diagnostics(B, d);
    track_old_rmr(B, it);
    res_row(B, log, it, d);
    if (converged(B, it, simpletol, d)) { log.converged = true; break; }
    B.say("STEP 3 -- Correcting velocity field\n");
    B.AssemblePressureCorrectionSIMPLE();
    B.ComputePressureCorrection();
    B.CorrectVelocityFieldSIMPLE(d.v_moy * 100);
    B.ComputeDiv(d.MaxDeltaV, d.normdiv, d.velocityNorm, d.maxdiv);
    B.say("Max div after correction: %g\n", d.maxdiv);
    save_snapshot(B, it, save_step, savecounter);
    B.RelaxControl(true);
  }
  finish(B, log, "SIMPLER", it, start, d, now() - t0);
}

// SIMPLERR.cxx:75-280: SIMPLER plus a pressure refresh after the correction;
// the RES row is written at the end of the iteration and beta adapts too.
void SIMPLERRSOLVER::Solve(NavierStokesBase& B) {
  const double t0 = now();
  Diag d;
  save_step = B.save_step;
  int savecounter = 0;
  init_common(B, "SIMPLER-R", "SIMPLE-R", maxit, simpletol);
  B.AssembleAllMomentum();
  B.ComputeVelocityField(false, 0);
  B.ComputePermeability(d.Q, d.v_moy, d.K);
  B.say("Initial Permeability K (m²): %g\n", d.K);
  B.info_file << "Initial Permeability K (m²): " << d.K << std::endl;
  B.alpha = B.alpha0;
  int it;
  for (it = 0; it < maxit; ++it) {
    B.say("----------------------------------\nIteration %d  ---------------------\n", it);
    B.info_file << "Iteration " << it << "  ------------" << std::endl;
    B.AssembleAllMomentum();
    pressure_prediction(B, N1);
    B.say("STEP 2 -- Updating velocity field\n");
    B.ComputeVelocityField(true, d.v_moy * std::min(1., B.alpha * 1.5));
    diagnostics(B, d);
    track_old_rmr(B, it);
    if (converged(B, it, simpletol, d)) { log.converged = true; break; }
    B.say("STEP 3 -- Correcting pressure & velocity field\n");
    B.AssemblePressureCorrectionSIMPLE();
    B.ComputePressureCorrection();
    B.CorrectVelocityFieldSIMPLE(d.v_moy * 100);
    B.Compute_hat(NS3D_SET_VEL, NS3D_SET_VEL);
    B.AssemblePressure(true);
    B.ComputePressureField();
    B.ComputeDiv(d.MaxDeltaV, d.normdiv, d.velocityNorm, d.maxdiv);
    B.say("Max div after correction: %g\n", d.maxdiv);
    res_row(B, log, it, d);
    B.RelaxControl(false);
    save_snapshot(B, it, save_step, savecounter);
  }
  finish(B, log, "SIMPLER-R", it, start, d, now() - t0);
}

// IDEAL.cxx:75-281: SIMPLER's step 1, then N2 pressure/velocity updates of the
// velocity itself; no RES rows (header only, IDEAL.cxx:138).
void IDEALSOLVER::Solve(NavierStokesBase& B) {
  const double t0 = now();
  Diag d;
  save_step = B.save_step;
  int savecounter = 0;
  init_common(B, "IDEAL", "IDEAL", maxit, simpletol);
  B.AssembleAllMomentum();
  B.ComputeVelocityField(false, 0);
  B.ComputePermeability(d.Q, d.v_moy, d.K);
  B.say("Initial Permeability K (m²): %g\n", d.K);
  B.info_file << "Initial Permeability K (m²): " << d.K << std::endl;
  B.alpha = B.alpha0;
  int it;
  for (it = 0; it < maxit; ++it) {
    B.say("----------------------------------\nIteration %d  ---------------------\n", it);
    B.info_file << "Iteration " << it << "  ------------" << std::endl;
    B.AssembleAllMomentum();
    pressure_prediction(B, N1);
    B.say("STEP 2 -- Updating velocity field\n");
    B.ComputeVelocityField(true, d.v_moy * std::min(1., B.alpha * 1.5));
    diagnostics(B, d);
    track_old_rmr(B, it);
    log.rows.push_back(ResRow{(double)it, d.Q, d.v_moy, B.RMR, d.K,
                              (B.p_inlet - B.p_outlet) / (B.dx * (B.nx - 2 * B.border)), d.normdiv, d.reldiv,
                              d.velocityNorm, d.maxdiv, its(B.usolver), its(B.vsolver), its(B.wsolver),
                              its(B.psolver)});  // in-memory only: IDEAL writes no RES rows
    if (converged(B, it, simpletol, d)) { log.converged = true; break; }
    B.say("STEP 3 -- Updating pressure & velocity field\n");
    B.CopyVelocity(NS3D_SET_TEMP, NS3D_SET_VEL);
    for (int sub1 = 0; sub1 < N2; sub1++) {
      B.say("inner iteration %d\n", sub1);
      B.Compute_hat(NS3D_SET_VEL, NS3D_SET_TEMP);
      B.AssemblePressure(false);
      B.ComputePressureField();
      B.CorrectVelocityFieldSIMPLER(NS3D_SET_VEL, B.p0, B.p1);
    }
    B.ComputeDiv(d.MaxDeltaV, d.normdiv, d.velocityNorm, d.maxdiv);
    B.say("Max div after correction: %g\n", d.maxdiv);
    save_snapshot(B, it, save_step, savecounter);
    B.RelaxControl(true);
  }
  finish(B, log, "IDEAL", it, start, d, now() - t0);
}

}  // namespace ns3d_host
