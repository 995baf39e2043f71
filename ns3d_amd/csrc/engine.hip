// engine.hip -- the batched solve engine (see engine.h).
//
// The dispatcher loop, per round:
//   1. admit the pending requests that fit (tiles and solve slots): wait for
//      each request's ready event, launch init over a table holding only the
//      new solves, and add them to the active set of their precision;
//   2. rebuild the descriptor and tile tables of a precision group whose
//      membership changed (a few KB copied from pinned memory);
//   3. queue `chunk` iterations, each A, B, C for the FP64 group and the FP32
//      group (one launch per kernel per group, over all active solves);
//   4. gather every active solve's BState into one array (one kernel, one
//      copy), synchronise, retire the solves that stopped and wake their
//      callers, and queue restart kernels for those that asked for one.
// Kernels of solves that stopped exit at once, so a chunk never computes
// past a solve's end.  The host images of the tables are only rewritten
// between rounds, after the stream has drained, so no copy in flight reads a
// buffer that changed.
#include "engine.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>

using namespace ns3d;

namespace {

constexpr int kEngMaxSolves = 4096;
constexpr int kEngMaxTiles = 1 << 17;

#define ECHK(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

__global__ void gather_states_kernel(BState* const* __restrict__ src, BState* __restrict__ dst, int n) {
  constexpr int kWords = (int)(sizeof(BState) / 8);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * kWords) return;
  const int s = i / kWords, w = i % kWords;
  reinterpret_cast<uint64_t*>(dst + s)[w] = reinterpret_cast<const uint64_t*>(src[s])[w];
}

double kernel_bytes_eng(int64_t n, int64_t nnz, int k, int vb) {
  static const double streams[3] = {7.0, 4.0, 9.0};
  const double mat = (k < 2) ? (double)nnz * (vb + 4.0) : 0.0;
  return mat + streams[k] * vb * (double)n;
}

template <typename T>
SolveDesc<T>& desc_of(EngineRequest* r);
template <>
SolveDesc<double>& desc_of<double>(EngineRequest* r) { return r->d64; }
template <>
SolveDesc<float>& desc_of<float>(EngineRequest* r) { return r->d32; }

}  // namespace

ns3d_engine::ns3d_engine(int dev, int ch) : device(dev), chunk(ch > 0 ? ch : 8) {
  ECHK(hipSetDevice(device));
  ECHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  auto dalloc = [&](size_t bytes) {
    void* p = nullptr;
    ECHK(hipMalloc(&p, bytes));
    allocs.push_back(p);
    return p;
  };
  const size_t dsz = std::max(sizeof(SolveDesc<double>), sizeof(SolveDesc<float>));
  for (Group& g : grp) {
    g.d_desc = dalloc(dsz * kEngMaxSolves);
    g.d_idesc = dalloc(dsz * kEngMaxSolves);
    g.d_tiles = (Tile*)dalloc(sizeof(Tile) * kEngMaxTiles);
    g.d_itiles = (Tile*)dalloc(sizeof(Tile) * kEngMaxTiles);
    g.d_it0 = (long long*)dalloc(sizeof(long long) * kEngMaxSolves);
    for (int b = 0; b < 2; ++b) {
      g.h_it0[b].resize(kEngMaxSolves);
      g.h_desc[b].resize(dsz * kEngMaxSolves);
      g.h_idesc[b].resize(dsz * kEngMaxSolves);
    }
  }
  for (int b = 0; b < 2; ++b) {
    d_stptr[b] = (BState**)dalloc(sizeof(BState*) * 2 * kEngMaxSolves);
    d_states[b] = (BState*)dalloc(sizeof(BState) * 2 * kEngMaxSolves);
    h_stptr[b].resize(2 * kEngMaxSolves);
    ECHK(hipHostMalloc((void**)&h_states[b], sizeof(BState) * 2 * kEngMaxSolves, hipHostMallocDefault));
    ECHK(hipEventCreateWithFlags(&round_done[b], hipEventDisableTiming));
    for (auto& e : tev[b]) ECHK(hipEventCreate(&e));
  }
  thr = std::thread([this] { run(); });
}

ns3d_engine::~ns3d_engine() {
  {
    std::lock_guard<std::mutex> lk(mu);
    stopping = true;
  }
  cv_work.notify_all();
  if (thr.joinable()) thr.join();
  if (stream) (void)hipStreamSynchronize(stream);
  for (int b = 0; b < 2; ++b) {
    if (h_states[b]) (void)hipHostFree(h_states[b]);
    if (round_done[b]) (void)hipEventDestroy(round_done[b]);
    for (auto& e : tev[b])
      if (e) (void)hipEventDestroy(e);
  }
  for (void* p : allocs) (void)hipFree(p);
  if (stream) (void)hipStreamDestroy(stream);
}

void ns3d_engine::solve(EngineRequest& req) {
  EngineRequest* r = &req;
  solve_many(&r, 1);
}

void ns3d_engine::solve_many(EngineRequest* const* reqs, int n) {
  std::unique_lock<std::mutex> lk(mu);
  if (!fatal.empty()) throw std::runtime_error("solve engine failed: " + fatal);
  if (stopping) throw std::runtime_error("solve engine is shutting down");
  for (int i = 0; i < n; ++i) {
    reqs[i]->done = false;
    pending.push_back(reqs[i]);
  }
  cv_work.notify_one();
  auto all_done = [&] {
    for (int i = 0; i < n; ++i)
      if (!reqs[i]->done) return false;
    return true;
  };
  cv_done.wait(lk, [&] { return all_done() || !fatal.empty(); });
  if (!all_done()) throw std::runtime_error("solve engine failed: " + fatal);
}

void ns3d_engine::fence() {
  std::unique_lock<std::mutex> lk(mu);
  const long long target = rounds_queued;
  cv_done.wait(lk, [&] { return rounds_done >= target || !fatal.empty(); });
}

void ns3d_engine::stats(double* out, int count) {
  std::lock_guard<std::mutex> lk(mu);
  double v[32] = {};
  v[0] = n_chunks; v[1] = n_launch; v[2] = n_done; v[3] = max_active; v[4] = rows_x_iters;
  for (int p = 0; p < 2; ++p)
    for (int k = 0; k < 3; ++k) {
      v[5 + 9 * p + 3 * k] = ms[p][k];
      v[6 + 9 * p + 3 * k] = launches[p][k];
      v[7 + 9 * p + 3 * k] = bytes[p][k];
    }
  for (int q = 0; q < count && q < 32; ++q) out[q] = v[q];
}

void ns3d_engine::fail(const std::string& msg) {
  std::lock_guard<std::mutex> lk(mu);
  fatal = msg;
  cv_done.notify_all();
}

template <typename T>
Batch<T> ns3d_engine::batch(int g, int q) const {
  const Group& G = grp[g];
  Batch<T> B{};
  B.d = (const SolveDesc<T>*)G.d_desc;
  B.tiles = G.d_tiles;
  B.it0 = G.d_it0;
  B.nsolves = (int)G.active.size();
  B.ntiles = G.ntiles;
  B.q = q;
  return B;
}

// Host images of the tables (index buf) are rewritten only for a round queued
// after the one that last used them has been processed, so no copy still in
// flight reads a buffer that changed.
void ns3d_engine::admit(std::vector<EngineRequest*>& reqs, int buf) {
  for (EngineRequest* r : reqs)
    if (r->ready) ECHK(hipStreamWaitEvent(stream, r->ready, 0));
  for (int g = 0; g < 2; ++g) {
    Group& G = grp[g];
    std::vector<Tile> tiles;
    int ns = 0;
    auto init = [&](auto tag) {
      using T = decltype(tag);
      SolveDesc<T>* hd = (SolveDesc<T>*)G.h_idesc[buf].data();
      for (EngineRequest* r : reqs) {
        if (r->prec != g) continue;
        SolveDesc<T> d = desc_of<T>(r);
        make_tiles(ns, d.S.n, tiles, d.t0, d.t1);
        hd[ns++] = d;
      }
      if (ns == 0) return;
      G.h_itiles[buf] = tiles;
      ECHK(hipMemcpyAsync(G.d_idesc, hd, ns * sizeof(SolveDesc<T>), hipMemcpyHostToDevice, stream));
      ECHK(hipMemcpyAsync(G.d_itiles, G.h_itiles[buf].data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice,
                          stream));
      Batch<T> B{};
      B.d = (const SolveDesc<T>*)G.d_idesc;
      B.tiles = G.d_itiles;
      B.nsolves = ns;
      B.ntiles = (int)tiles.size();
      ECHK(launch_bicg_init<T>(B, stream));
    };
    if (g == 0) init(double{});
    else init(float{});
    for (EngineRequest* r : reqs)
      if (r->prec == g) {
        r->serial = next_serial++;
        r->it_next = 0;
        r->it_queued = -1;
        r->restart_queued = false;
        live[r->serial] = r;
        G.active.push_back(r);
      }
    if (ns) G.dirty = true;
  }
}

// Iteration bases of the group's active solves, in table order, for the
// round being queued (launch q of the chunk is iteration it0 + q); each
// solve is then expected to be a chunk further on.
void ns3d_engine::upload_it0(int g, int buf, bool restarted) {
  Group& G = grp[g];
  if (G.active.empty()) return;
  for (size_t i = 0; i < G.active.size(); ++i) {
    EngineRequest* r = G.active[i];
    G.h_it0[buf][i] = r->it_next;
    r->it_queued = r->it_next;
    r->restart_queued = restarted;
    r->it_next += chunk;
  }
  ECHK(hipMemcpyAsync(G.d_it0, G.h_it0[buf].data(), G.active.size() * sizeof(long long), hipMemcpyHostToDevice,
                      stream));
}

void ns3d_engine::rebuild(int g, int buf) {
  Group& G = grp[g];
  std::vector<Tile> tiles;
  auto build = [&](auto tag) {
    using T = decltype(tag);
    SolveDesc<T>* hd = (SolveDesc<T>*)G.h_desc[buf].data();
    for (size_t i = 0; i < G.active.size(); ++i) {
      SolveDesc<T> d = desc_of<T>(G.active[i]);
      make_tiles((int)i, d.S.n, tiles, d.t0, d.t1);
      hd[i] = d;
    }
    G.ntiles = (int)tiles.size();
    if (G.active.empty()) return;
    G.h_tiles[buf] = tiles;
    ECHK(hipMemcpyAsync(G.d_desc, hd, G.active.size() * sizeof(SolveDesc<T>), hipMemcpyHostToDevice, stream));
    ECHK(hipMemcpyAsync(G.d_tiles, G.h_tiles[buf].data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice,
                        stream));
  };
  if (g == 0) build(double{});
  else build(float{});
  G.dirty = false;
}

// Tables, restarts, one chunk, and the gather of the states of every solve
// in the round, into buffer buf.
void ns3d_engine::queue_round(int buf, bool restart[2]) {
  for (int g = 0; g < 2; ++g) {
    if (grp[g].dirty) rebuild(g, buf);
    // the restart kernel acts only on solves whose status asks for it
    const bool rs = restart[g] && !grp[g].active.empty();
    if (rs) {
      if (g == 0) ECHK(launch_bicg_restart<double>(batch<double>(0, 0), stream));
      else ECHK(launch_bicg_restart<float>(batch<float>(1, 0), stream));
    }
    restart[g] = false;
    upload_it0(g, buf, rs);
  }
  std::vector<long long>& ord = order[buf];
  ord.clear();
  int nact = 0;
  round_bytes[buf].assign(6, 0.0);
  for (int g = 0; g < 2; ++g)
    for (EngineRequest* r : grp[g].active) {
      ord.push_back(r->serial);
      h_stptr[buf][nact++] = r->prec == 0 ? r->d64.w.st : r->d32.w.st;
      for (int k = 0; k < 3; ++k) round_bytes[buf][3 * r->prec + k] += kernel_bytes_eng(r->n, r->nnz, k, r->prec ? 4 : 8);
    }
  max_active = std::max(max_active, (double)nact);
  ECHK(hipMemcpyAsync(d_stptr[buf], h_stptr[buf].data(), nact * sizeof(BState*), hipMemcpyHostToDevice, stream));
  // one chunk of iterations; the middle one carries dispatch events
  const int qt = chunk / 2;
  hipEvent_t* ev = tev[buf];
  const bool g0 = !grp[0].active.empty(), g1 = !grp[1].active.empty();
  for (int q = 0; q < chunk; ++q) {
    const bool e = (q == qt);
    if (g0) ECHK(launch_bicg_A<double>(batch<double>(0, q), stream, e ? ev : nullptr));
    if (g1) ECHK(launch_bicg_A<float>(batch<float>(1, q), stream, e ? ev + 6 : nullptr));
    if (g0) ECHK(launch_bicg_B<double>(batch<double>(0, q), stream, e ? ev + 2 : nullptr));
    if (g1) ECHK(launch_bicg_B<float>(batch<float>(1, q), stream, e ? ev + 8 : nullptr));
    if (g0) ECHK(launch_bicg_C<double>(batch<double>(0, q), stream, e ? ev + 4 : nullptr));
    if (g1) ECHK(launch_bicg_C<float>(batch<float>(1, q), stream, e ? ev + 10 : nullptr));
  }
  n_launch += (double)chunk * 3 * ((g0 ? 1 : 0) + (g1 ? 1 : 0));
  n_chunks += 1;
  const int words = nact * (int)(sizeof(BState) / 8);
  hipLaunchKernelGGL(gather_states_kernel, dim3((words + 255) / 256), dim3(256), 0, stream, d_stptr[buf],
                     d_states[buf], nact);
  ECHK(hipGetLastError());
  ECHK(hipMemcpyAsync(h_states[buf], d_states[buf], nact * sizeof(BState), hipMemcpyDeviceToHost, stream));
  ECHK(hipEventRecord(round_done[buf], stream));
  std::lock_guard<std::mutex> lk(mu);
  rounds_queued += 1;
}

// After round buf has completed: sampled kernel times, then every solve of the
// round that is still live.  The round queued since (with base it_queued)
// advances a running solve by a chunk if that base equals its count, and
// does nothing to it otherwise (its kernels exit at their iteration check);
// a solve that asked for a restart is restarted by that round only if the
// round restarts its group, else it is flagged for the next one.  Stopped
// solves are retired and their callers woken.  Returns whether any was.
bool ns3d_engine::process_round(int buf, bool restart[2]) {
  ECHK(hipEventSynchronize(round_done[buf]));
  for (int g = 0; g < 2; ++g) {
    if (round_bytes[buf][3 * g] == 0) continue;
    for (int k = 0; k < 3; ++k) {
      float e = 0.f;
      if (hipEventElapsedTime(&e, tev[buf][6 * g + 2 * k], tev[buf][6 * g + 2 * k + 1]) != hipSuccess) continue;
      ms[g][k] += e;
      launches[g][k] += 1;
      bytes[g][k] += round_bytes[buf][3 * g + k];
    }
  }
  bool any_done = false;
  std::lock_guard<std::mutex> lk(mu);
  const std::vector<long long>& ord = order[buf];
  for (size_t i = 0; i < ord.size(); ++i) {
    auto lv = live.find(ord[i]);
    if (lv == live.end()) continue;  // retired by the previous round's results
    EngineRequest* r = lv->second;
    const BState& st = h_states[buf][i];
    const bool next_runs = r->it_queued == st.it_c;
    if (st.status == BS_RUNNING) {
      r->it_next = next_runs ? st.it_c + chunk : st.it_c;
      continue;
    }
    if (st.status == BS_RESTART) {
      if (next_runs && r->restart_queued) {
        r->it_next = st.it_c + chunk;
      } else {
        restart[r->prec] = true;
        r->it_next = st.it_c;
      }
      continue;
    }
    live.erase(lv);
    r->result = st;
    r->done = true;
    rows_x_iters += (double)r->n * (double)st.it_c;
    n_done += 1;
    any_done = true;
    auto& act = grp[r->prec].active;
    auto it = std::find(act.begin(), act.end(), r);
    if (it != act.end()) {
      act.erase(it);
      grp[r->prec].dirty = true;
    }
  }
  rounds_done += 1;
  return any_done;
}

// Rounds are pipelined: round k+1 is queued before the states of round k are
// processed, so the device keeps working while the host retires, admits and
// rebuilds.  Round k+1 was queued before round k's stops were known: its
// kernels of those solves exit at their status check, and their callers were
// already woken -- fence() is what a context calls before freeing memory.
void ns3d_engine::run() {
  try {
    ECHK(hipSetDevice(device));
    bool restart[2] = {false, false};
    int qbuf = 0;       // buffer of the next round to queue
    int inflight = -1;  // buffer of the queued round not yet processed
    for (;;) {
      std::vector<EngineRequest*> admitted;
      {
        std::unique_lock<std::mutex> lk(mu);
        if (inflight < 0) cv_work.wait(lk, [&] { return stopping || !pending.empty() || any_active(); });
        if (stopping && pending.empty() && !any_active() && inflight < 0) break;
        int tiles[2] = {grp[0].ntiles, grp[1].ntiles};
        int solves[2] = {(int)grp[0].active.size(), (int)grp[1].active.size()};
        while (!pending.empty()) {
          EngineRequest* r = pending.front();
          const int64_t n = r->prec == 0 ? r->d64.S.n : r->d32.S.n;
          const int nt = tiles_for(n);
          if (solves[r->prec] > 0 && (tiles[r->prec] + nt > kEngMaxTiles || solves[r->prec] + 1 > kEngMaxSolves))
            break;  // waits for room (first-come first-served)
          tiles[r->prec] += nt;
          solves[r->prec] += 1;
          admitted.push_back(r);
          pending.pop_front();
        }
      }
      bool queued = false;
      if (!admitted.empty()) admit(admitted, qbuf);
      if (any_active()) {
        queue_round(qbuf, restart);
        queued = true;
      }
      bool woke = false;
      if (inflight >= 0) woke = process_round(inflight, restart);
      if (woke || inflight >= 0) cv_done.notify_all();
      inflight = queued ? qbuf : -1;
      if (queued) qbuf ^= 1;
    }
  } catch (const std::exception& e) {
    fail(e.what());
  }
}
