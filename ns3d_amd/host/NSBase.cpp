// NSBase.cpp -- host facade of NavierStokesBase over the GPU context.
// Each method cites the reference function it replaces.
#include "NSBase.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <stdexcept>

#include "vtk_io.h"

namespace ns3d_host {

void check(ns3d_status st, const char* what) {
  if (st != NS3D_OK) throw std::runtime_error(std::string(what) + ": " + ns3d_last_error());
}

void NavierStokesBase::say(const char* fmt, ...) const {
  if (!verbose) return;
  va_list ap;
  va_start(ap, fmt);
  vprintf(fmt, ap);
  va_end(ap);
}

void NavierStokesBase::note(const char* fmt, ...) const {
  if (!progress) return;
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fflush(stderr);
}

void NavierStokesBase::tick(Phase p) {
  const auto t = std::chrono::steady_clock::now();
  phase[p] += std::chrono::duration<double>(t - t0_).count();
  t0_ = t;
}

// NSBase.cxx:50-194 (constructor).
NavierStokesBase::NavierStokesBase(const std::vector<uint8_t>& raw, int nx_, int ny_, int nz_,
                                   const std::string& imagename, const std::string& pfield,
                                   const std::string& vxf, const std::string& vyf, const std::string& vzf,
                                   double dx_, double dy_, double dz_, double vsolvertol_, double psolvertol_,
                                   double pcsolvertol_, int maxitsolver_, double alpha_, double beta_, double p0_,
                                   double gradp_, double mu_, double rho_, int porevalue_, int direction_,
                                   double deltaRMR_, int save_step_, int border_, int device, int verbose_,
                                   bool write_files_, const std::string& outdir_)
    : maxitsolver(maxitsolver_), porevalue(porevalue_), gradient_direction(direction_), save_step(save_step_),
      border(border_), dx(dx_), dy(dy_), dz(dz_), gradp(gradp_), alpha0(alpha_), beta0(beta_), alpha(alpha_),
      beta(beta_), mu(mu_), rho(rho_), p1(0), p0(p0_), p_outlet(0), p_inlet(0), vsolvertol(vsolvertol_),
      psolvertol(psolvertol_), pcsolvertol(pcsolvertol_), deltaRMR(deltaRMR_), nx(0), ny(0), nz(0), np(0), nu(0),
      nv(0), nw(0), schemefactor(0), pcmax(0), pcmin(0), pcnorm(0), min_alpha(0.001), min_beta(0.001),
      relax(0.05), maxconvdelta(2.5), RMR(0), old_RMR(0), outdir(outdir_), pfieldfilename(pfield),
      vxfilename(vxf), vyfilename(vyf), vzfilename(vzf), verbose(verbose_ == 1), progress(verbose_ >= 2), write_files(write_files_) {
  say("Low Reynolds Steady Navier-Stokes simulation (MI355X)\n");
  say("--- Input parameters ---\n");
  say("Image %s\n", imagename.c_str());
  say("dx, dy, dz = %g, %g, %g\n", dx, dy, dz);
  say("viscosity (Pas.s) = %g\n", mu);
  say("density (kg/m^3) = %g\n", rho);
  say("Relaxation parameter for velocity correction: %g\n", alpha);
  say("Relaxation parameter for pressure correction: %g\n", beta);
  say("Tolerance for linear system resolution (pressure & velocity & pressure correction): %g, %g, %g\n",
      psolvertol, vsolvertol, pcsolvertol);
  say("Max number of iteration for the iterative linear solver: %d\n", maxitsolver);
  say("Pore voxel value = %d\n", porevalue);
  say("Direction of the pressure gradient: %d (1:x, 2:y, 3:z)\n", gradient_direction);
  say("Pressure gradient (Pa/m): %g\n", gradp);
  say("Save step (snapshot every N outer iterations): %d\n", save_step);
  say("Border (x2): %d\n", border);

  // output names: basename of the image minus its 4-character extension
  outfilename = imagename;
  const size_t slash = outfilename.find_last_of('/');
  if (slash != std::string::npos) outfilename.erase(0, slash + 1);
  const std::string stem = outfilename.size() >= 4 ? outfilename.substr(0, outfilename.size() - 4) : outfilename;
  info_filename = "INFO-" + std::to_string(gradient_direction) + "-" + stem + ".log";
  log_filename = "RES-" + std::to_string(gradient_direction) + "-" + stem + ".csv";
  if (write_files) {
    const std::string pre = outdir.empty() ? "" : outdir + "/";
    info_file.open(pre + info_filename);
    log_file.open(pre + log_filename);
    if (!info_file || !log_file) throw std::runtime_error("cannot open the INFO/RES files in '" + outdir + "'");
  }
  info_file.precision(15);

  // binarisation (vtkio.cxx:139-142), rotation and border (NSBase.cxx:131-136)
  std::vector<uint8_t> img(raw.size());
  for (size_t i = 0; i < raw.size(); ++i) img[i] = (raw[i] == (uint8_t)porevalue) ? 0 : 1;
  int dims[3];
  check(ns3d_prepare_image(img.data(), nx_, ny_, nz_, gradient_direction, border, nullptr, dims), "prepare");
  std::vector<uint8_t> prep((size_t)dims[0] * dims[1] * dims[2]);
  check(ns3d_prepare_image(img.data(), nx_, ny_, nz_, gradient_direction, border, prep.data(), dims), "prepare");
  if (border > 0) say("Adding a border of %d voxels on each side\n", border);
  nx = dims[0]; ny = dims[1]; nz = dims[2];
  info_file << "Size " << nx << "," << ny << "," << nz << std::endl;
  say("Size %d, %d, %d\n", nx, ny, nz);

  check(ns3d_create(prep.data(), nx, ny, nz, device, &ctx), "ns3d_create");
  int64_t cnt[4];
  check(ns3d_counts(ctx, cnt), "ns3d_counts");
  np = cnt[0]; nu = cnt[1]; nv = cnt[2]; nw = cnt[3];

  // NSBase.cxx:155-161
  p1 = -gradp * (dx * nx) + p0;
  if (p1 < 0) {
    p0 -= p1;
    p1 = 0;
  }
  ns3d_params prm;
  prm.dx = dx; prm.dy = dy; prm.dz = dz; prm.mu = mu; prm.rho = rho; prm.p0 = p0; prm.p1 = p1;
  prm.border = border; prm.reserved = 0;
  check(ns3d_set_params(ctx, &prm), "ns3d_set_params");

  info_file << "Porosity: " << float(np) / float((int64_t)nx * ny * nz) << std::endl;
  info_file << "Nodes for the pressure field: " << np << std::endl;
  info_file << "Nodes for the u velocity field: " << nu << std::endl;
  info_file << "Nodes for the v velocity field: " << nv << std::endl;
  info_file << "Nodes for the w velocity field: " << nw << std::endl;
  info_file << "Delta RMR min : " << deltaRMR << std::endl;
  say("Porosity: %g\n", float(np) / float((int64_t)nx * ny * nz));
  say("Nodes for the pressure field: %lld\n", (long long)np);
  say("Nodes for the u velocity field: %lld\n", (long long)nu);
  say("Nodes for the v velocity field: %lld\n", (long long)nv);
  say("Nodes for the w velocity field: %lld\n", (long long)nw);
  say("Delta RMR min : %g\n", deltaRMR);
  t0_ = std::chrono::steady_clock::now();
}

NavierStokesBase::~NavierStokesBase() { ns3d_destroy(ctx); }

void NavierStokesBase::ZeroFields() { check(ns3d_zero_fields(ctx), "ns3d_zero_fields"); }

void NavierStokesBase::CopyVelocity(int dst, int src) {
  check(ns3d_copy_velocity(ctx, dst, src), "ns3d_copy_velocity");
  tick(PH_OTHER);
}

// NSBase.cxx:853-884 (AssembleAllMomentum, with compute()).
void NavierStokesBase::AssembleAllMomentum() {
  tick(PH_OTHER);
  check(ns3d_assemble_momentum(ctx, alpha), "AssembleAllMomentum");
  tick(PH_MOMENTUM);
}

// NSBase.cxx:1773-1884 (ComputeVelocityField).
void NavierStokesBase::ComputeVelocityField(bool limit, double delta_v_max) {
  tick(PH_OTHER);
  ns3d_solve_info inf[3];
  check(ns3d_compute_velocity_field(ctx, alpha, limit ? 1 : 0, delta_v_max, vsolvertol, maxitsolver, inf),
        "ComputeVelocityField");
  tick(PH_VELOCITY);
  usolver.set(inf[0]);
  vsolver.set(inf[1]);
  say("Solving for component u in %d iterations (status %d)\n", (int)inf[0].iterations, inf[0].info);
  info_file << "Solving for component u in " << inf[0].iterations << " iterations (status " << inf[0].info << ")"
            << std::endl;
  say("Solving for component v in %d iterations (status %d)\n", (int)inf[1].iterations, inf[1].info);
  info_file << "Solving for component v in " << inf[1].iterations << " iterations (status " << inf[1].info << ")"
            << std::endl;
  if (nw > 0) {
    wsolver.set(inf[2]);
    say("Solving for component w in %d iterations (status %d)\n", (int)inf[2].iterations, inf[2].info);
    info_file << "Solving for component w in " << inf[2].iterations << " iterations (status " << inf[2].info
              << ")" << std::endl;
  }
}

// NSBase.cxx:1548-1755 (Compute_u_hat, Compute_v_hat, Compute_w_hat).
void NavierStokesBase::Compute_hat(int uc_set, int uold_set) {
  tick(PH_OTHER);
  check(ns3d_compute_hat(ctx, alpha, uc_set, uold_set), "Compute_hat");
  tick(PH_HAT);
}

// NSBase.cxx:478-610 (AssemblePressure on the pseudo-velocities).
void NavierStokesBase::AssemblePressure(bool relax_p) {
  tick(PH_OTHER);
  check(ns3d_assemble_pressure(ctx, NS3D_PRESSURE_FULL, relax_p ? 1 : 0, beta, NS3D_SET_HAT), "AssemblePressure");
  tick(PH_PASM);
}

// NSBase.cxx:612-732
void NavierStokesBase::AssemblePressureCorrectionSIMPLEC() {
  tick(PH_OTHER);
  say("Assembling the linear system matrix for the pressure correction...OK\n");
  check(ns3d_assemble_pressure(ctx, NS3D_PC_SIMPLEC, 0, beta, NS3D_SET_VEL), "AssemblePressureCorrectionSIMPLEC");
  tick(PH_PASM);
}

// NSBase.cxx:734-851
void NavierStokesBase::AssemblePressureCorrectionSIMPLE() {
  tick(PH_OTHER);
  check(ns3d_assemble_pressure(ctx, NS3D_PC_SIMPLE, 0, beta, NS3D_SET_VEL), "AssemblePressureCorrectionSIMPLE");
  tick(PH_PASM);
}

static void report_psolve(NavierStokesBase& b, const ns3d_solve_info& s, const char* what) {
  if (s.info == 0) {
    b.say("%d iterations (status %d)\n", (int)s.iterations, s.info);
    b.info_file << s.iterations << " iterations (status " << s.info << ")" << std::endl;
  } else {
    const char* m = s.info == 1 ? "Numerical Issue" : s.info == 2 ? "No convergence" : "Invalid input";
    b.say("%s (iterations: %d). Residual: %g\n", m, (int)s.iterations, s.error);
    b.info_file << m << " (iterations:" << s.iterations << "). Residual: " << s.error << std::endl;
  }
  (void)what;
}

// NSBase.cxx:2337-2365 (ComputePressureCorrection)
void NavierStokesBase::ComputePressureCorrection() {
  tick(PH_OTHER);
  say("Solving for pressure correction: ");
  info_file << "Solving for pressure correction : ";
  ns3d_solve_info s;
  check(ns3d_compute_pressure_correction(ctx, pcsolvertol, maxitsolver, &s), "ComputePressureCorrection");
  tick(PH_PSOLVE);
  psolver.set(s);
  report_psolve(*this, s, "pc");
}

// NSBase.cxx:2383-2423 (ComputePressureField)
void NavierStokesBase::ComputePressureField() {
  tick(PH_OTHER);
  say("Solving for pressure: ");
  info_file << "Solving for pressure in : ";
  ns3d_solve_info s;
  check(ns3d_compute_pressure_field(ctx, psolvertol, maxitsolver, &s), "ComputePressureField");
  tick(PH_PSOLVE);
  psolver.set(s);
  report_psolve(*this, s, "p");
}

// NSBase.cxx:2082-2172
void NavierStokesBase::CorrectVelocityFieldSIMPLE(double delta_v_max) {
  tick(PH_OTHER);
  say("Correcting Velocity\n");
  check(ns3d_correct_velocity(ctx, NS3D_CORR_SIMPLE, delta_v_max, NS3D_SET_VEL, 0, 0), "CorrectVelocityFieldSIMPLE");
  tick(PH_CORRECT);
}

// NSBase.cxx:2174-2263
void NavierStokesBase::CorrectVelocityFieldSIMPLEC(double delta_v_max) {
  tick(PH_OTHER);
  say("Correcting Velocity\n");
  check(ns3d_correct_velocity(ctx, NS3D_CORR_SIMPLEC, delta_v_max, NS3D_SET_VEL, 0, 0),
        "CorrectVelocityFieldSIMPLEC");
  tick(PH_CORRECT);
}

// NSBase.cxx:2265-2335
void NavierStokesBase::CorrectVelocityFieldSIMPLER(int target_set, double p_0, double p_1) {
  tick(PH_OTHER);
  say("Correcting Velocity\n");
  check(ns3d_correct_velocity(ctx, NS3D_CORR_SIMPLER, 0.0, target_set, p_0, p_1), "CorrectVelocityFieldSIMPLER");
  tick(PH_CORRECT);
}

// NSBase.cxx:2367-2380
void NavierStokesBase::CorrectPressureField() {
  tick(PH_OTHER);
  double out[3];
  check(ns3d_correct_pressure(ctx, beta, out), "CorrectPressureField");
  tick(PH_CORRECT);
  pcmax = out[0];
  pcmin = out[1];
  pcnorm = out[2];
  say("Max and min pressure correction: %g, %g\n", pcmax, pcmin);
  info_file << "Correction de pression maximale et minimale: " << pcmax << ", " << pcmin << std::endl;
  say("Norm of the pressure correction: %g\n", pcnorm);
  info_file << "Norme de la correction de pression : " << pcnorm << std::endl;
}

// NSBase.cxx:1958-2080
void NavierStokesBase::ComputeDiv(double& MaxDeltaV, double& normdiv, double& velocityNorm, double& maxdiv) {
  tick(PH_OTHER);
  double out[4];
  check(ns3d_compute_div(ctx, NS3D_SET_VEL, out), "ComputeDiv");
  tick(PH_DIAG);
  MaxDeltaV = out[0];
  normdiv = out[1];
  velocityNorm = out[2];
  maxdiv = out[3];
}

// NSBase.cxx:1886-1956
void NavierStokesBase::ComputePermeability(double& Q, double& v_moy, double& K) {
  tick(PH_OTHER);
  double out[5];
  check(ns3d_compute_permeability(ctx, out), "ComputePermeability");
  tick(PH_DIAG);
  Q = out[0];
  v_moy = out[1];
  K = out[2];
  p_inlet = out[3];
  p_outlet = out[4];
  if (K == 0.0 && !(std::fabs(p_inlet - p_outlet) > 0.0))
    say("WARNING: zero pressure difference (p_inlet = p_outlet), K set to 0\n");
}

// NSBase.cxx:2645-2690 (Initpvector): field from file or the linear profile.
void NavierStokesBase::Initpvector() {
  if (!warm_p.empty()) {
    check(ns3d_set_vector(ctx, NS3D_VEC_P, warm_p.data(), np), "ns3d_set_vector");
    say("Importing pressure field (warm start)\n");
  } else if (pfieldfilename != "none") {
    Image im = read_image(pfieldfilename);
    if (im.nx != nx || im.ny != ny || im.nz != nz) throw std::runtime_error("pressure field size mismatch");
    std::vector<int32_t> pidx((size_t)nx * ny * nz);
    check(ns3d_get_index(ctx, pidx.data(), nullptr, nullptr, nullptr), "ns3d_get_index");
    std::vector<double> p((size_t)np, 0.0);
    for (size_t c = 0; c < pidx.size(); ++c)
      if (pidx[c] != -1) p[pidx[c]] = im.data[c * im.ncomp];
    check(ns3d_set_vector(ctx, NS3D_VEC_P, p.data(), np), "ns3d_set_vector");
    say("Importing pressure field %s\n", pfieldfilename.c_str());
  } else {
    check(ns3d_init_pressure_linear(ctx), "Initpvector");
  }
  say("p0, p1 (Pa)= %g, %g\n", p0, p1);
  info_file << "p0, p1 (Pa)= " << p0 << "," << p1 << std::endl;
  say("Pressure gradient (Pa/m)= %g\n", (p0 - p1) / (dx * nx));
  info_file << "Pressure gradient (Pa/m)= " << (p0 - p1) / (dx * nx) << std::endl;
  info_file << "inlet/outlet area (m²)= " << (dy * dz * ny * nz) << std::endl;
}

// NSBase.cxx:2575-2642 (InitVelocity): warm start from Float64 fields.
void NavierStokesBase::InitVelocity() {
  const std::string* names[3] = {&vxfilename, &vyfilename, &vzfilename};
  const int vecs[3] = {NS3D_VEC_U, NS3D_VEC_V, NS3D_VEC_W};
  const int64_t lens[3] = {nu, nv, nw};
  const std::vector<double>* warm[3] = {&warm_u, &warm_v, &warm_w};
  std::vector<int32_t> idx[3];
  for (int q = 0; q < 3; ++q) {
    if (!warm[q]->empty()) {
      check(ns3d_set_vector(ctx, vecs[q], warm[q]->data(), lens[q]), "ns3d_set_vector");
      say("Importing velocity field %c (warm start)\n", "xyz"[q]);
      continue;
    }
    if (*names[q] == "none") continue;
    if (idx[0].empty()) {
      idx[0].resize((size_t)(nx + 1) * ny * nz);
      idx[1].resize((size_t)nx * ny * nz);
      idx[2].resize((size_t)nx * ny * nz);
      check(ns3d_get_index(ctx, nullptr, idx[0].data(), idx[1].data(), idx[2].data()), "ns3d_get_index");
    }
    Image im = read_image(*names[q]);
    const int ex = (q == 0) ? nx + 1 : nx;
    if (im.nx != ex || im.ny != ny || im.nz != nz) throw std::runtime_error("velocity field size mismatch");
    std::vector<double> x((size_t)lens[q], 0.0);
    for (size_t c = 0; c < idx[q].size(); ++c)
      if (idx[q][c] != -1) x[idx[q][c]] = im.data[c * im.ncomp];
    check(ns3d_set_vector(ctx, vecs[q], x.data(), lens[q]), "ns3d_set_vector");
    say("Importing velocity field %c from %s\n", "xyz"[q], names[q]->c_str());
  }
}

// NSBase.cxx:196-301 (RelaxControl).  The reference pops from a possibly
// empty history on the first growth test; here an empty history is left
// alone.
int relax_control_step(RelaxState& s, bool beta_cst, double win[3], bool* window_ran) {
  const size_t size = 30;
  int ev = 0;
  *window_ran = false;
  if (s.RMR > s.old_RMR * s.maxconvdelta) {
    s.alpha = std::max(s.min_alpha, s.alpha - s.relax);
    if (!beta_cst) s.beta = std::max(s.min_beta, s.beta - s.relax);
    s.old_RMR *= 1.1;
    if (!s.avgRMR.empty()) s.avgRMR.pop_front();
    ev |= RELAX_JUMP;
  }
  const bool isDecreasing = !s.avgRMR.empty() && s.avgRMR.back() > s.RMR;
  s.avgRMR.push_back(s.RMR);
  if (s.avgRMR.size() <= size) return ev;
  s.avgRMR.pop_front();
  // first and second half of the window (the middle entry goes to the second)
  const double half = s.avgRMR.size() * 0.5;
  double sum[2] = {0, 0};
  int cnt[2] = {0, 0};
  int i = 0;
  for (double x : s.avgRMR) {
    const int h = ++i < half ? 0 : 1;
    sum[h] += x;
    cnt[h] += 1;
  }
  const double avg1 = sum[0] / cnt[0], avg2 = sum[1] / cnt[1];
  const double delta = (avg2 - avg1) / std::max(avg1, avg2);
  win[0] = avg1; win[1] = avg2; win[2] = delta;
  *window_ran = true;
  if (isDecreasing) return ev;
  const bool flat = delta <= 0 && std::fabs(delta) < s.deltaRMR;
  if (!flat && !(delta > 0)) return ev;
  for (int q = 0; q < 15 && !s.avgRMR.empty(); ++q) s.avgRMR.pop_front();
  if (flat) {
    // no lower clamp on alpha in this branch (NSBase.cxx:263)
    s.alpha = std::min(s.alpha0, s.alpha - s.relax * 0.25);
    if (!beta_cst) s.beta = std::max(s.min_beta, s.beta - s.relax * 0.25);
    return ev | RELAX_FLAT;
  }
  s.alpha = std::max(s.min_alpha, s.alpha - s.relax);
  if (!beta_cst) s.beta = std::max(s.min_beta, s.beta - s.relax);
  return ev | RELAX_RISING;
}

void NavierStokesBase::RelaxControl(bool beta_cst) {
  RelaxState s{alpha, beta, alpha0, min_alpha, min_beta, relax, maxconvdelta, deltaRMR, RMR, old_RMR, {}};
  s.avgRMR.swap(avgRMR);
  const size_t before = s.avgRMR.size();
  const double last = before ? s.avgRMR.back() : 0.0;
  double win[3];
  bool ran = false;
  const int ev = relax_control_step(s, beta_cst, win, &ran);
  alpha = s.alpha;
  beta = s.beta;
  old_RMR = s.old_RMR;
  avgRMR.swap(s.avgRMR);
  auto tell = [&](const char* what) {
    say("%s -> new relaxation parameter alpha, beta: %g, %g\n", what, alpha, beta);
    info_file << what << " -> new relaxation parameter alpha, beta: " << alpha << ", " << beta << std::endl;
  };
  if (ev & RELAX_JUMP) tell("Relative Mass Residual increases too much");
  // against the history's last entry, if one is left after a jump's pop
  if (before > ((ev & RELAX_JUMP) ? 1u : 0u)) say("delta RMR = %6.3e \n", RMR - last);
  if (ran)
    say("avg RMRold = %f, avg RMRnew = %f, delta avgRMR / avgRMR_max = %f %%\n", win[0], win[1], win[2] * 100);
  if (ev & RELAX_FLAT) tell("Simulation is oscillating");
  if (ev & RELAX_RISING) tell("Simulation is diverging");
}

// NSBase.cxx:304-328 (CheckConvergence); nz > 0 always holds.
bool NavierStokesBase::CheckConvergence(double res, double tol) {
  if (res < tol) {
    say("Converged ! RMR=: %g\n", res);
    return true;
  }
  if (usolver.iterations() == 0 && vsolver.iterations() == 0 && wsolver.iterations() == 0 &&
      psolver.iterations() == 0) {
    say("Converged !\nVelocity and pressure field stabilized\n");
    return true;
  }
  return false;
}

std::vector<double> NavierStokesBase::vector(int which) const {
  int64_t n = np;
  if (which == NS3D_VEC_U || which == NS3D_VEC_UT || which == NS3D_VEC_UH) n = nu;
  if (which == NS3D_VEC_V || which == NS3D_VEC_VT || which == NS3D_VEC_VH) n = nv;
  if (which == NS3D_VEC_W || which == NS3D_VEC_WT || which == NS3D_VEC_WH) n = nw;
  std::vector<double> x((size_t)n);
  check(ns3d_get_vector(ctx, which, x.data(), n), "ns3d_get_vector");
  return x;
}

// NSBase.cxx:330-476 (FieldsToImage): scatter to grids, face-averaged cell
// velocity, write Vx/Vy/Vz/V/P .vti files.
void NavierStokesBase::FieldsToImage(const std::string& filename) {
  if (!write_files) return;
  const std::string pre = outdir.empty() ? "" : outdir + "/";
  const std::string dir = std::to_string(gradient_direction);
  const size_t nc = (size_t)nx * ny * nz, nuc = (size_t)(nx + 1) * ny * nz;
  std::vector<int32_t> pi(nc), ui(nuc), vi(nc), wi(nc);
  check(ns3d_get_index(ctx, pi.data(), ui.data(), vi.data(), wi.data()), "ns3d_get_index");
  const std::vector<double> p = vector(NS3D_VEC_P), u = vector(NS3D_VEC_U), v = vector(NS3D_VEC_V),
                            w = vector(NS3D_VEC_W);
  std::vector<double> uf(nuc, 0.0), vf(nc, 0.0), wf(nc, 0.0), pf(nc, 0.0);
  for (size_t c = 0; c < nuc; ++c)
    if (ui[c] > -1) uf[c] = u[ui[c]];
  for (size_t c = 0; c < nc; ++c) {
    if (vi[c] > -1) vf[c] = v[vi[c]];
    if (wi[c] > -1) wf[c] = w[wi[c]];
    if (pi[c] > -1) pf[c] = p[pi[c]];
  }
  write_scalar(pre + "Vx-" + dir + "-" + filename, uf.data(), nx + 1, ny, nz);
  write_scalar(pre + "Vy-" + dir + "-" + filename, vf.data(), nx, ny, nz);
  write_scalar(pre + "Vz-" + dir + "-" + filename, wf.data(), nx, ny, nz);
  std::vector<double> au(nc, 0.0), av(nc, 0.0), aw(nc, 0.0);
  const size_t sj = nz, si = (size_t)ny * nz;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const size_t c = (size_t)i * si + (size_t)j * sj + k;
        const int a = ui[c + si], b = ui[c];
        if (a > -1 && b > -1) au[c] = (u[a] + u[b]) * 0.50;
        else if (a > -1) au[c] = (u[a]) * 0.50;
        else if (b > -1) au[c] = (u[b]) * 0.50;
        if (j < ny - 1) {
          const int e = vi[c + sj], f = vi[c];
          if (e > -1 && f > -1) av[c] = (v[e] + v[f]) * 0.50;
          else if (e > -1) av[c] = (v[e]) * 0.50;
          else if (f > -1) av[c] = (v[f]) * 0.50;
        } else if (vi[c] > -1) {
          av[c] = (v[vi[c]]) * 0.50;
        }
        if (k < nz - 1) {
          const int e = wi[c + 1], f = wi[c];
          if (e > -1 && f > -1) aw[c] = (w[e] + w[f]) * 0.50;
          else if (e > -1) aw[c] = (w[e]) * 0.50;
          else if (f > -1) aw[c] = (w[f]) * 0.50;
        } else if (wi[c] > -1) {
          aw[c] = (w[wi[c]]) * 0.50;
        }
      }
  write_scalar(pre + "P-" + dir + "-" + filename, pf.data(), nx, ny, nz);
  write_vector(pre + "V-" + dir + "-" + filename, au.data(), av.data(), aw.data(), nx, ny, nz);
  for (const char* t : {"Vx-", "Vy-", "Vz-", "V-", "P-"}) info_file << "Writing " << t << dir << "-" << filename << std::endl;
}

}  // namespace ns3d_host
