// server.cpp -- permeability solves served as they arrive (no counterpart in
// the reference, which runs one image per process).
//
// Requests queue in arrival order.  A pool of worker threads (the slots)
// takes the first queued request (in arrival order, or the largest first)
// whose projected device footprint fits in what the budget has left
// (ns3d_footprint, an upper bound of the context's allocations) and, if a
// cap is set, whose rows fit under the rows in flight.  It runs the
// request's driver with the context attached to the server's one solve
// engine, and frees the memory and the slot on return.  Because all
// running requests hand their BiCGSTAB solves to the same engine, solves of
// requests that overlap in time run in the same launches, and a request that
// arrives while others are on the device joins them at the next chunk
// boundary (continuous batching).  A mixed-precision request whose run fails
// is run again in plain FP64 (fallback >= 1); with fallback 2 one that ends
// unconverged at its outer-iteration cap is also continued in FP64.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ns3d.h"
#include "../../include/ns3d_host.h"

namespace {

thread_local std::string g_srv_err;

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct ns3d_server {
  struct Job {
    int64_t id = 0;
    ns3d_run_params p{};
    std::string outdir, name;
    std::vector<uint8_t> img;
    int nx = 0, ny = 0, nz = 0;
    double footprint = 0;
    double rows = 0;  // fluid cells of the (rotated, bordered) image: the size of each of its systems
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    ns3d_run_result res{};
    double t_submit = 0, t_start = 0, t_end = 0;
    std::string err;
  };

  int device = 0;
  double budget = 0;
  int fallback = 1;
  ns3d_engine* eng = nullptr;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job*> queue;
  std::unordered_map<int64_t, std::unique_ptr<Job>> jobs;
  int64_t next_id = 1;
  double in_use = 0, peak_in_use = 0;
  // admission policy: order 0 takes requests in arrival order, 1 the largest
  // first; both skip to the first queued request that fits.  max_rows > 0
  // caps the rows in flight (enough to fill the GPU; more only stretches
  // every request's iteration cadence).
  int order = 0;
  double max_rows = 0, rows_in_flight = 0;
  int running = 0, peak_running = 0;
  double n_done = 0, n_failed = 0, n_fallback = 0, n_submitted = 0;
  bool stopping = false;
  std::vector<std::thread> workers;

  void worker();
  void run_job(Job& j);
  // The queued request to start now, or queue.end() (call with mu held).
  std::deque<Job*>::iterator pick() {
    auto best = queue.end();
    for (auto it = queue.begin(); it != queue.end(); ++it) {
      Job* j = *it;
      const bool fits = running == 0 || (in_use + j->footprint <= budget &&
                                         (max_rows <= 0 || rows_in_flight + j->rows <= max_rows));
      if (!fits) continue;
      if (order == 0) return it;
      if (best == queue.end() || j->rows > (*best)->rows) best = it;
    }
    return best;
  }
};

void ns3d_server::run_job(Job& j) {
  // fields: in (warm start, may be empty) and out (final fields, when asked)
  auto attempt = [&](ns3d_run_params p, ns3d_run_result& res, std::vector<double>* warm, std::vector<double>* keep) {
    ns3d_solver* s = nullptr;
    if (ns3d_solver_create(&p, j.img.data(), j.nx, j.ny, j.nz, &s) != 0) throw std::string(ns3d_host_last_error());
    if (ns3d_attach_engine(ns3d_solver_context(s), eng) != NS3D_OK) {
      const std::string m = ns3d_last_error();
      ns3d_solver_destroy(s);
      throw m;
    }
    if (warm && !warm[0].empty())
      ns3d_solver_warm_start(s, warm[0].data(), warm[1].data(), warm[2].data(), warm[3].empty() ? nullptr : warm[3].data());
    int rc = ns3d_solver_run(s, nullptr, 0, nullptr, &res);
    std::string m = rc ? ns3d_host_last_error() : "";
    if (!rc && keep) {
      int dims[3];
      int64_t cnt[4];
      ns3d_solver_info(s, dims, cnt);
      keep[0].resize(cnt[0]); keep[1].resize(cnt[1]); keep[2].resize(cnt[2]); keep[3].resize(cnt[3]);
      rc = ns3d_solver_fields(s, keep[0].data(), keep[1].data(), keep[2].data(), keep[3].data());
      if (rc) m = ns3d_host_last_error();
    }
    ns3d_solver_destroy(s);
    if (rc) throw m;
  };
  ns3d_run_params p = j.p;
  p.device = device;
  p.outdir = j.outdir.empty() ? nullptr : j.outdir.c_str();
  p.name = j.name.empty() ? nullptr : j.name.c_str();
  const bool mixed = p.precision != NS3D_PRECISION_FP64;
  const bool continue_unconverged = mixed && fallback >= NS3D_FALLBACK_UNCONVERGED;
  std::vector<double> fields[4];
  std::string failure;
  try {
    attempt(p, j.res, nullptr, continue_unconverged ? fields : nullptr);
  } catch (const std::string& m) {
    if (!mixed || fallback < NS3D_FALLBACK_FAILED) throw;
    failure = m.empty() ? "mixed-precision run failed" : m;
  }
  const bool broken = !failure.empty() || !std::isfinite(j.res.K) || !std::isfinite(j.res.RMR);
  if (mixed && fallback >= NS3D_FALLBACK_FAILED && broken) {
    // the mixed run failed or produced non-finite values: run it again in
    // FP64 from the start
    std::fprintf(stderr, "ns3d server: request %lld (%dx%dx%d): mixed-precision run %s; running it again in FP64\n",
                 (long long)j.id, j.nx, j.ny, j.nz,
                 failure.empty() ? (std::isfinite(j.res.K) ? "gave a non-finite residual" : "gave a non-finite K")
                                 : ("failed: " + failure).c_str());
    p.precision = NS3D_PRECISION_FP64;
    attempt(p, j.res, nullptr, nullptr);
    j.res.fallback = 1;
  } else if (continue_unconverged && !j.res.converged) {
    // continue in FP64 from the mixed run's fields, with a quarter of the
    // outer-iteration cap: where FP64 can converge it does so quickly from
    // there, and a request that stalls in any precision is not paid twice
    const int total = j.res.iterations;
    p.precision = NS3D_PRECISION_FP64;
    p.maxitSIMPLE = std::max(10, p.maxitSIMPLE / 4);
    attempt(p, j.res, fields, nullptr);
    j.res.fallback = 1;
    j.res.iterations += total;
  }
}

void ns3d_server::worker() {
  for (;;) {
    Job* j = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      std::deque<Job*>::iterator at;
      cv_work.wait(lk, [&] {
        if (stopping && queue.empty()) return true;
        at = pick();
        return at != queue.end();
      });
      if (queue.empty()) return;  // stopping
      j = *at;
      queue.erase(at);
      rows_in_flight += j->rows;
      j->state = 1;
      j->t_start = now_s(t0);
      in_use += j->footprint;
      peak_in_use = std::max(peak_in_use, in_use);
      running += 1;
      peak_running = std::max(peak_running, running);
    }
    std::string err;
    try {
      run_job(*j);
    } catch (const std::string& m) {
      err = m.empty() ? "request failed" : m;
    } catch (const std::exception& e) {
      err = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      j->t_end = now_s(t0);
      j->state = err.empty() ? 2 : 3;
      j->err = err;
      j->img.clear();
      j->img.shrink_to_fit();
      in_use -= j->footprint;
      rows_in_flight -= j->rows;
      running -= 1;
      if (err.empty()) n_done += 1;
      else n_failed += 1;
      if (j->res.fallback) n_fallback += 1;
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

extern "C" {

const char* ns3d_server_last_error(void) { return g_srv_err.c_str(); }

int ns3d_server_create(int device, double budget_bytes, int max_slots, int chunk, int fallback, ns3d_server** out) {
  if (!out || max_slots < 1 || budget_bytes <= 0) {
    g_srv_err = "ns3d_server_create: bad arguments";
    return 1;
  }
  *out = nullptr;
  auto s = std::make_unique<ns3d_server>();
  s->device = device < 0 ? 0 : device;
  s->budget = budget_bytes;
  s->fallback = fallback < 0 ? 0 : (fallback > NS3D_FALLBACK_UNCONVERGED ? NS3D_FALLBACK_UNCONVERGED : fallback);
  s->t0 = std::chrono::steady_clock::now();
  if (ns3d_engine_create(s->device, chunk, &s->eng) != NS3D_OK) {
    g_srv_err = ns3d_last_error();
    return 2;
  }
  ns3d_server* raw = s.release();
  for (int q = 0; q < max_slots; ++q) raw->workers.emplace_back([raw] { raw->worker(); });
  *out = raw;
  return 0;
}

int ns3d_server_submit(ns3d_server* s, const ns3d_run_params* p, const uint8_t* img, int nx, int ny, int nz,
                       int64_t* ticket) {
  if (!s || !p || !img || !ticket || nx <= 0 || ny <= 0 || nz <= 0) {
    g_srv_err = "ns3d_server_submit: bad arguments";
    return 1;
  }
  if (p->direction < 1 || p->direction > 3 || p->border < 0) {
    g_srv_err = "ns3d_server_submit: bad direction or border";
    return 1;
  }
  auto j = std::make_unique<ns3d_server::Job>();
  j->p = *p;
  if (p->outdir) j->outdir = p->outdir;
  if (p->name) j->name = p->name;
  const size_t n = (size_t)nx * ny * nz;
  j->img.assign(img, img + n);
  j->nx = nx; j->ny = ny; j->nz = nz;
  int64_t nfluid = 0;
  for (size_t q = 0; q < n; ++q) nfluid += (img[q] == (uint8_t)p->porevalue);
  int d[3] = {nx, ny, nz};  // the frame after rotation and border
  if (p->direction == 2) { d[0] = ny; d[1] = nx; }
  if (p->direction == 3) { d[0] = nz; d[2] = nx; }
  d[0] += 2 * p->border;
  nfluid += (int64_t)2 * p->border * d[1] * d[2];
  j->rows = (double)nfluid;
  if (ns3d_footprint(d[0], d[1], d[2], nfluid, p->precision != 0, 3, &j->footprint) != NS3D_OK) {
    g_srv_err = ns3d_last_error();
    return 2;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stopping) {
    g_srv_err = "ns3d_server_submit: server is shutting down";
    return 2;
  }
  if (j->footprint > s->budget) {
    g_srv_err = "ns3d_server_submit: request needs more device memory than the whole budget";
    return 3;
  }
  j->id = s->next_id++;
  j->t_submit = now_s(s->t0);
  *ticket = j->id;
  s->queue.push_back(j.get());
  s->jobs[j->id] = std::move(j);
  s->n_submitted += 1;
  s->cv_work.notify_all();
  return 0;
}

static int fetch(ns3d_server* s, int64_t ticket, ns3d_run_result* res, double times[4], bool block) {
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(ticket);
  if (it == s->jobs.end()) {
    g_srv_err = "unknown ticket";
    return -1;
  }
  ns3d_server::Job& j = *it->second;
  if (block) s->cv_done.wait(lk, [&] { return j.state >= 2; });
  if (j.state < 2) return 1;
  if (times) {
    times[0] = j.t_submit; times[1] = j.t_start; times[2] = j.t_end; times[3] = j.footprint;
  }
  if (j.state == 3) {
    g_srv_err = j.err;
    return -2;
  }
  if (res) *res = j.res;
  return 0;
}

int ns3d_server_poll(ns3d_server* s, int64_t ticket, ns3d_run_result* res, double times[4]) {
  if (!s) return -1;
  return fetch(s, ticket, res, times, false);
}

int ns3d_server_wait(ns3d_server* s, int64_t ticket, ns3d_run_result* res, double times[4]) {
  if (!s) return -1;
  return fetch(s, ticket, res, times, true);
}

int ns3d_server_stats(ns3d_server* s, double out[16]) {
  if (!s || !out) return 1;
  std::lock_guard<std::mutex> lk(s->mu);
  double v[16] = {s->n_submitted, (double)s->queue.size(), (double)s->running, s->n_done, s->n_failed,
                  s->n_fallback, s->in_use, s->peak_in_use, (double)s->peak_running, s->budget, now_s(s->t0)};
  std::memcpy(out, v, sizeof(v));
  return 0;
}

ns3d_engine* ns3d_server_engine(ns3d_server* s) { return s ? s->eng : nullptr; }

int ns3d_server_set_policy(ns3d_server* s, int order, double max_rows) {
  if (!s || order < 0 || order > 1 || max_rows < 0) {
    g_srv_err = "ns3d_server_set_policy: bad arguments";
    return 1;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->order = order;
    s->max_rows = max_rows;
  }
  s->cv_work.notify_all();
  return 0;
}

void ns3d_server_destroy(ns3d_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers)
    if (t.joinable()) t.join();
  ns3d_engine_destroy(s->eng);
  delete s;
}

}  // extern "C"
