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
    g.h_it0.resize(kEngMaxSolves);
    g.h_desc.resize(dsz * kEngMaxSolves);
    g.h_idesc.resize(dsz * kEngMaxSolves);
  }
  d_stptr = (BState**)dalloc(sizeof(BState*) * 2 * kEngMaxSolves);
  d_states = (BState*)dalloc(sizeof(BState) * 2 * kEngMaxSolves);
  h_stptr.resize(2 * kEngMaxSolves);
  ECHK(hipHostMalloc((void**)&h_states, sizeof(BState) * 2 * kEngMaxSolves, hipHostMallocDefault));
  for (auto& e : tev) ECHK(hipEventCreate(&e));
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
  if (h_states) (void)hipHostFree(h_states);
  for (void* p : allocs) (void)hipFree(p);
  for (auto& e : tev)
    if (e) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

void ns3d_engine::solve(EngineRequest& req) {
  std::unique_lock<std::mutex> lk(mu);
  if (!fatal.empty()) throw std::runtime_error("solve engine failed: " + fatal);
  if (stopping) throw std::runtime_error("solve engine is shutting down");
  req.done = false;
  pending.push_back(&req);
  cv_work.notify_one();
  cv_done.wait(lk, [&] { return req.done || !fatal.empty(); });
  if (!req.done) throw std::runtime_error("solve engine failed: " + fatal);
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

void ns3d_engine::admit(std::vector<EngineRequest*>& reqs) {
  for (EngineRequest* r : reqs)
    if (r->ready) ECHK(hipStreamWaitEvent(stream, r->ready, 0));
  for (int g = 0; g < 2; ++g) {
    Group& G = grp[g];
    std::vector<Tile> tiles;
    int ns = 0;
    auto init = [&](auto tag) {
      using T = decltype(tag);
      SolveDesc<T>* hd = (SolveDesc<T>*)G.h_idesc.data();
      for (EngineRequest* r : reqs) {
        if (r->prec != g) continue;
        SolveDesc<T> d = desc_of<T>(r);
        make_tiles(ns, d.S.n, tiles, d.t0, d.t1);
        hd[ns++] = d;
      }
      if (ns == 0) return;
      G.h_itiles = tiles;
      ECHK(hipMemcpyAsync(G.d_idesc, hd, ns * sizeof(SolveDesc<T>), hipMemcpyHostToDevice, stream));
      ECHK(hipMemcpyAsync(G.d_itiles, G.h_itiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, stream));
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
        r->it_c = 0;
        G.active.push_back(r);
      }
    if (ns) G.dirty = true;
  }
}

// Executed iterations of the group's active solves, in table order, for the
// row kernels of the next chunk (launch q of the chunk is iteration it0 + q).
void ns3d_engine::upload_it0(int g) {
  Group& G = grp[g];
  if (G.active.empty()) return;
  for (size_t i = 0; i < G.active.size(); ++i) G.h_it0[i] = G.active[i]->it_c;
  ECHK(hipMemcpyAsync(G.d_it0, G.h_it0.data(), G.active.size() * sizeof(long long), hipMemcpyHostToDevice, stream));
}

void ns3d_engine::rebuild(int g) {
  Group& G = grp[g];
  std::vector<Tile> tiles;
  auto build = [&](auto tag) {
    using T = decltype(tag);
    SolveDesc<T>* hd = (SolveDesc<T>*)G.h_desc.data();
    for (size_t i = 0; i < G.active.size(); ++i) {
      SolveDesc<T> d = desc_of<T>(G.active[i]);
      make_tiles((int)i, d.S.n, tiles, d.t0, d.t1);
      hd[i] = d;
    }
    G.ntiles = (int)tiles.size();
    if (G.active.empty()) return;
    G.h_tiles = tiles;
    ECHK(hipMemcpyAsync(G.d_desc, hd, G.active.size() * sizeof(SolveDesc<T>), hipMemcpyHostToDevice, stream));
    ECHK(hipMemcpyAsync(G.d_tiles, G.h_tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, stream));
  };
  if (g == 0) build(double{});
  else build(float{});
  G.dirty = false;
}

void ns3d_engine::run() {
  try {
    ECHK(hipSetDevice(device));
    bool restart[2] = {false, false};
    for (;;) {
      std::vector<EngineRequest*> admitted;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stopping || !pending.empty() || any_active(); });
        if (stopping && pending.empty() && !any_active()) break;
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
      if (!admitted.empty()) admit(admitted);
      // tables, iteration bases (each pinned staging buffer is written once
      // per round, after the previous round's sync), restarts
      for (int g = 0; g < 2; ++g) {
        if (grp[g].dirty) rebuild(g);
        upload_it0(g);
        if (restart[g] && !grp[g].active.empty()) {
          if (g == 0) ECHK(launch_bicg_restart<double>(batch<double>(0, 0), stream));
          else ECHK(launch_bicg_restart<float>(batch<float>(1, 0), stream));
        }
        restart[g] = false;
      }
      // state pointers in solve order (group 0 then group 1)
      order.clear();
      for (int g = 0; g < 2; ++g)
        for (EngineRequest* r : grp[g].active) order.push_back(r);
      const int nact = (int)order.size();
      max_active = std::max(max_active, (double)nact);
      for (int i = 0; i < nact; ++i) h_stptr[i] = order[i]->prec == 0 ? order[i]->d64.w.st : order[i]->d32.w.st;
      ECHK(hipMemcpyAsync(d_stptr, h_stptr.data(), nact * sizeof(BState*), hipMemcpyHostToDevice, stream));
      // one chunk of iterations; the middle one carries dispatch events
      const int qt = chunk / 2;
      for (int q = 0; q < chunk; ++q) {
        const bool ev = (q == qt);
        if (!grp[0].active.empty()) ECHK(launch_bicg_A<double>(batch<double>(0, q), stream, ev ? tev : nullptr));
        if (!grp[1].active.empty()) ECHK(launch_bicg_A<float>(batch<float>(1, q), stream, ev ? tev + 6 : nullptr));
        if (!grp[0].active.empty()) ECHK(launch_bicg_B<double>(batch<double>(0, q), stream, ev ? tev + 2 : nullptr));
        if (!grp[1].active.empty()) ECHK(launch_bicg_B<float>(batch<float>(1, q), stream, ev ? tev + 8 : nullptr));
        if (!grp[0].active.empty()) ECHK(launch_bicg_C<double>(batch<double>(0, q), stream, ev ? tev + 4 : nullptr));
        if (!grp[1].active.empty()) ECHK(launch_bicg_C<float>(batch<float>(1, q), stream, ev ? tev + 10 : nullptr));
      }
      n_launch += (double)chunk * 3 * ((grp[0].active.empty() ? 0 : 1) + (grp[1].active.empty() ? 0 : 1));
      n_chunks += 1;
      const int words = nact * (int)(sizeof(BState) / 8);
      hipLaunchKernelGGL(gather_states_kernel, dim3((words + 255) / 256), dim3(256), 0, stream, d_stptr, d_states,
                         nact);
      ECHK(hipGetLastError());
      ECHK(hipMemcpyAsync(h_states, d_states, nact * sizeof(BState), hipMemcpyDeviceToHost, stream));
      ECHK(hipStreamSynchronize(stream));
      // sampled kernel times; bytes counted for every active solve (a solve
      // that stopped inside the chunk makes the sample slightly optimistic)
      for (int g = 0; g < 2; ++g) {
        if (grp[g].active.empty()) continue;
        double b[3] = {0, 0, 0};
        for (EngineRequest* r : grp[g].active)
          for (int k = 0; k < 3; ++k) b[k] += kernel_bytes_eng(r->n, r->nnz, k, g == 0 ? 8 : 4);
        for (int k = 0; k < 3; ++k) {
          float e = 0.f;
          ECHK(hipEventElapsedTime(&e, tev[6 * g + 2 * k], tev[6 * g + 2 * k + 1]));
          ms[g][k] += e;
          launches[g][k] += 1;
          bytes[g][k] += b[k];
        }
      }
      // retire
      bool any_done = false;
      {
        std::lock_guard<std::mutex> lk(mu);
        for (int i = 0; i < nact; ++i) {
          EngineRequest* r = order[i];
          const BState& st = h_states[i];
          r->it_c = st.it_c;
          if (st.status == BS_RUNNING) continue;
          if (st.status == BS_RESTART) {
            restart[r->prec] = true;
            continue;
          }
          r->result = st;
          r->done = true;
          rows_x_iters += (double)r->n * (double)st.it_c;
          n_done += 1;
          any_done = true;
          auto& act = grp[r->prec].active;
          act.erase(std::find(act.begin(), act.end(), r));
          grp[r->prec].dirty = true;
        }
      }
      if (any_done) cv_done.notify_all();
    }
  } catch (const std::exception& e) {
    fail(e.what());
  }
}
