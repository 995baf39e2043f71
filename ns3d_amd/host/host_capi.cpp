// host_capi.cpp -- C ABI of libns3d_host.so (include/ns3d_host.h): one
// NavierStokesBase + one driver per handle.
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/ns3d_host.h"
#include "NSBase.h"
#include "drivers.h"
#include "vtk_io.h"

using namespace ns3d_host;

namespace {
thread_local std::string g_host_err;
}

struct ns3d_solver {
  ns3d_run_params prm;
  std::unique_ptr<NavierStokesBase> base;
  RunLog log;
};

extern "C" {

const char* ns3d_host_last_error(void) { return g_host_err.c_str(); }

int ns3d_read_image(const char* path, int porevalue, uint8_t* out, int dims[3]) {
  if (!path || !dims) {
    g_host_err = "ns3d_read_image: bad arguments";
    return 1;
  }
  try {
    int nx = 0, ny = 0, nz = 0;
    const std::vector<uint8_t> img = read_binary_image(path, porevalue, nx, ny, nz);
    dims[0] = nx; dims[1] = ny; dims[2] = nz;
    if (out) std::memcpy(out, img.data(), img.size());
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return 1;
  }
  return 0;
}

int ns3d_relax_replay(const double* rmr, int n, int beta_cst, double alpha0, double beta0, double deltaRMR,
                      double* alpha_out, double* beta_out, int* events) {
  if (!rmr || n < 0 || !alpha_out || !beta_out || !events) {
    g_host_err = "ns3d_relax_replay: bad arguments";
    return 1;
  }
  // the constants of the NavierStokesBase constructor (NSBase.cxx:171-179)
  RelaxState s{alpha0, beta0, alpha0, 0.001, 0.001, 0.05, 2.5, deltaRMR, 0.0, 0.0, {}};
  for (int it = 0; it < n; ++it) {
    s.RMR = rmr[it];
    // the drivers' best-residual tracking (SIMPLEC.cxx:155-166)
    if (it > 1) {
      if (s.RMR < s.old_RMR) s.old_RMR = s.RMR;
    } else {
      s.old_RMR = s.RMR;
    }
    double win[3];
    bool ran = false;
    events[it] = relax_control_step(s, beta_cst != 0, win, &ran);
    alpha_out[it] = s.alpha;
    beta_out[it] = s.beta;
  }
  return 0;
}

int ns3d_solver_create(const ns3d_run_params* p, const uint8_t* img, int nx, int ny, int nz, ns3d_solver** out) {
  if (!p || !img || !out || nx <= 0 || ny <= 0 || nz <= 0) {
    g_host_err = "ns3d_solver_create: bad arguments";
    return 1;
  }
  *out = nullptr;
  try {
    auto s = std::make_unique<ns3d_solver>();
    s->prm = *p;
    const double dy = p->dy <= 0 ? p->dx : p->dy, dz = p->dz <= 0 ? p->dx : p->dz;
    std::vector<uint8_t> raw(img, img + (size_t)nx * ny * nz);
    s->base = std::make_unique<NavierStokesBase>(
        raw, nx, ny, nz, p->name ? p->name : "image.vti", "none", "none", "none", "none", p->dx, dy, dz,
        p->vsolvertol, p->psolvertol, p->pcsolvertol, p->maxitsolver, p->alpha, p->beta, p->p0, p->gradp, p->mu,
        p->rho, p->porevalue, p->direction, p->deltaRMR, p->save_step, p->border, p->device, p->verbose,
        p->write_files != 0, p->outdir ? p->outdir : "");
    if (p->precision != 0) {
      check(ns3d_set_precision(s->base->ctx, p->precision, p->inner_tol), "ns3d_set_precision");
      check(ns3d_set_mixed_min_rows(s->base->ctx, p->mixed_min_rows), "ns3d_set_mixed_min_rows");
      check(ns3d_set_multigrid(s->base->ctx, p->mg_min_rows), "ns3d_set_multigrid");
    }
    *out = s.release();
    return 0;
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return 2;
  }
}

int ns3d_solver_info(ns3d_solver* s, int dims[3], int64_t counts[4]) {
  if (!s) return 1;
  dims[0] = s->base->nx; dims[1] = s->base->ny; dims[2] = s->base->nz;
  counts[0] = s->base->np; counts[1] = s->base->nu; counts[2] = s->base->nv; counts[3] = s->base->nw;
  return 0;
}

int ns3d_solver_run(ns3d_solver* s, double* rows, int max_rows, int* nrows, ns3d_run_result* res) {
  if (!s) return 1;
  try {
    NavierStokesBase& B = *s->base;
    const ns3d_run_params& p = s->prm;
    if (p.method == NS3D_METHOD_SIMPLEC) {
      SIMPLECSOLVER d(p.simpletol, p.maxitSIMPLE);
      d.Solve(B);
      s->log = d.log;
    } else if (p.method == NS3D_METHOD_SIMPLER) {
      SIMPLERSOLVER d(p.simpletol, p.maxitSIMPLE, p.N1);
      d.Solve(B);
      s->log = d.log;
    } else if (p.method == NS3D_METHOD_SIMPLERR) {
      SIMPLERRSOLVER d(p.simpletol, p.maxitSIMPLE, p.N1);
      d.Solve(B);
      s->log = d.log;
    } else if (p.method == NS3D_METHOD_IDEAL) {
      IDEALSOLVER d(p.simpletol, p.maxitSIMPLE, p.N1, p.N2);
      d.Solve(B);
      s->log = d.log;
    } else {
      throw std::runtime_error("unknown method");
    }
    const RunLog& L = s->log;
    if (nrows) *nrows = (int)L.rows.size();
    if (rows)
      for (int r = 0; r < (int)L.rows.size() && r < max_rows; ++r)
        std::memcpy(rows + 14 * r, L.rows[r].data(), 14 * sizeof(double));
    if (res) {
      std::memset(res, 0, sizeof(*res));
      res->iterations = L.iterations;
      res->converged = L.converged ? 1 : 0;
      res->K = L.K; res->Q = L.Q; res->v_moy = L.v_moy; res->RMR = L.RMR;
      res->normdiv = L.normdiv; res->reldiv = L.reldiv; res->maxdiv = L.maxdiv;
      res->seconds = L.seconds;
      res->np = B.np; res->nu = B.nu; res->nv = B.nv; res->nw = B.nw;
      res->solver_iterations[0] = B.usolver.total;
      res->solver_iterations[1] = B.vsolver.total;
      res->solver_iterations[2] = B.wsolver.total;
      res->solver_iterations[3] = B.psolver.total;
      res->solver_work[0] = B.usolver.work;
      res->solver_work[1] = B.vsolver.work;
      res->solver_work[2] = B.wsolver.work;
      res->solver_work[3] = B.psolver.work;
      for (int q = 0; q < PH_COUNT; ++q) res->phase_seconds[q] = B.phase[q];
    }
    return 0;
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return 2;
  }
}

int ns3d_solver_fields(ns3d_solver* s, double* p, double* u, double* v, double* w) {
  if (!s) return 1;
  try {
    NavierStokesBase& B = *s->base;
    auto cp = [&](double* dst, int which) {
      if (!dst) return;
      const std::vector<double> x = B.vector(which);
      std::memcpy(dst, x.data(), x.size() * sizeof(double));
    };
    cp(p, NS3D_VEC_P);
    cp(u, NS3D_VEC_U);
    cp(v, NS3D_VEC_V);
    cp(w, NS3D_VEC_W);
    return 0;
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return 2;
  }
}

int ns3d_solver_warm_start(ns3d_solver* s, const double* p, const double* u, const double* v, const double* w) {
  if (!s) return 1;
  NavierStokesBase& B = *s->base;
  auto put = [](std::vector<double>& dst, const double* src, int64_t n) {
    if (src) dst.assign(src, src + n);
    else dst.clear();
  };
  put(B.warm_p, p, B.np);
  put(B.warm_u, u, B.nu);
  put(B.warm_v, v, B.nv);
  put(B.warm_w, w, B.nw);
  return 0;
}

void ns3d_solver_destroy(ns3d_solver* s) { delete s; }

ns3d_ctx* ns3d_solver_context(ns3d_solver* s) { return s ? s->base->ctx : nullptr; }

}  // extern "C"
