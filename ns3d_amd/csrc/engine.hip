// engine.hip -- the batched solve engine (see engine.h).
//
// The dispatcher loop, per round:
//   1. admit the pending requests that fit (tiles and solve slots) to the
//      lane of their precision with fewer rows: wait for each request's
//      ready event, launch init over a table holding only the new solves,
//      and add them to the active set of their group;
//   2. rebuild the descriptor table and tiles of a group whose membership
//      changed; upload every active group's tile table (it carries the
//      iteration bases);
//   3. queue `chunk` iterations, each A, B, C for every active group (one
//      launch per kernel per group, over all its solves, on its lane's
//      stream);
//   4. join the lanes, gather every active solve's BState into one array
//      (one kernel, one copy) and record the round's completion.
// Rounds are pipelined: round k+1 is queued before round k's states are
// read back and processed (retire stopped solves and wake their callers,
// flag restarts).  Kernels of solves that stopped exit at once, so a chunk
// never computes past a solve's end.
#include "engine.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

using namespace ns3d;

namespace {

constexpr int kEngMaxSolves = 4096;
constexpr int kEngMaxTiles = 1 << 17;
constexpr int kTableHeader = 256;  // bytes before a tile table's entries (its tile count)

// Grid of a graph launch over ntiles tiles: rounded up to one of four steps
// per octave, so graphs are reused while the tile count drifts.
int graph_grid(int ntiles) {
  int g = 64;
  while (g < ntiles) g *= 2;
  if (g > 64) {
    const int step = g / 8;  // four steps in (g/2, g]
    g = (ntiles + step - 1) / step * step;
  }
  return std::min(tile_grid(g), tile_grid(kEngMaxTiles));
}

#define ECHK(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// The states of a group's n solves (in its descriptor table's order) into
// dst, pinned host memory the host reads once the lane's event has fired.
template <typename T>
__global__ void gather_states_kernel(const SolveDesc<T>* __restrict__ desc, BState* __restrict__ dst, int n) {
  constexpr int kWords = (int)(sizeof(BState) / 8);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * kWords) return;
  const int s = i / kWords, w = i % kWords;
  reinterpret_cast<uint64_t*>(dst + s)[w] = reinterpret_cast<const uint64_t*>(desc[s].w.st)[w];
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

#ifndef NS3D_CTX_STREAM
#define NS3D_CTX_STREAM 1
#endif

ns3d_engine::ns3d_engine(int dev, int ch) : device(dev), chunk(ch > 0 ? ch : 8) {
  ECHK(hipSetDevice(device));
  if (const char* e = std::getenv("NS3D_ENGINE_EVENTS")) events = std::atoi(e) != 0;
  if (NS3D_CTX_STREAM) ECHK(hipStreamCreateWithFlags(&ctx_stream, hipStreamNonBlocking));
  // streams of the lanes in use only (each takes one of the process's HIP
  // hardware queues); more are created if the schedule asks for them
  for (int l = 0; l < nlanes; ++l) ECHK(hipStreamCreateWithFlags(&lane_stream[l], hipStreamNonBlocking));
  stream = lane_stream[0];
  auto dalloc = [&](size_t bytes) {
    void* p = nullptr;
    ECHK(hipMalloc(&p, bytes));
    allocs.push_back(p);
    return p;
  };
  const size_t dsz = std::max(sizeof(SolveDesc<double>), sizeof(SolveDesc<float>));
  const size_t tsz = std::max(sizeof(TileDesc<double>), sizeof(TileDesc<float>));
  for (Group& g : grp) {
    g.d_desc = dalloc(dsz * kEngMaxSolves);
    g.d_idesc = dalloc(dsz * kEngMaxSolves);
    char* tb = (char*)dalloc(kTableHeader + tsz * kEngMaxTiles);
    g.d_nt = (int*)tb;
    g.d_tiles = tb + kTableHeader;
    g.d_itiles = (char*)dalloc(kTableHeader + tsz * kEngMaxTiles) + kTableHeader;
    for (int b = 0; b < 2; ++b) {
      g.h_desc[b].reserve(dsz * 64);
      g.h_idesc[b].reserve(dsz * 64);
    }
  }
  for (int b = 0; b < 2; ++b) {
    // coherent: the gather kernels' stores reach host memory directly, with
    // no dependence on a cache write-back at the end of the round
    ECHK(hipHostMalloc((void**)&h_states[b], sizeof(BState) * kGroups * kEngMaxSolves, hipHostMallocCoherent));
    for (auto& e : lane_done[b]) ECHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
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
  for (auto& st : lane_stream)
    if (st) (void)hipStreamSynchronize(st);
  for (int b = 0; b < 2; ++b) {
    if (h_states[b]) (void)hipHostFree(h_states[b]);
    for (auto& e : lane_done[b])
      if (e) (void)hipEventDestroy(e);
    for (auto& e : tev[b])
      if (e) (void)hipEventDestroy(e);
  }
  for (auto& c : gcache)
    for (auto& kv : c) (void)hipGraphExecDestroy(kv.second);
  for (void* p : allocs) (void)hipFree(p);
  for (auto& st : lane_stream)
    if (st) (void)hipStreamDestroy(st);
  if (ctx_stream) {
    (void)hipStreamSynchronize(ctx_stream);
    (void)hipStreamDestroy(ctx_stream);
  }
}

void HostBuf::reserve(size_t n) {
  if (n <= cap) return;
  release();
  const size_t c = std::max(n, cap + cap / 2);
  ECHK(hipHostMalloc((void**)&p, c, hipHostMallocDefault));
  cap = c;
}

void HostBuf::release() {
  if (p) (void)hipHostFree(p);
  p = nullptr;
  cap = size = 0;
}

void ns3d_engine::set_graphs(bool on) { graphs = on; }

bool ns3d_engine::set_schedule(int ch, int lanes, bool use_graphs) {
  if (ch < 1 || ch > 64 || lanes < 1 || lanes > kMaxLanes) throw std::runtime_error("set_schedule: bad arguments");
  std::lock_guard<std::mutex> lk(mu);
  if (!pending.empty() || any_active() || rounds_done != rounds_queued) return false;
  ECHK(hipSetDevice(device));
  for (int l = nlanes; l < lanes; ++l)
    if (!lane_stream[l]) ECHK(hipStreamCreateWithFlags(&lane_stream[l], hipStreamNonBlocking));
  if (ch != chunk) {
    for (int l = 0; l < kMaxLanes; ++l) {
      if (lane_stream[l]) ECHK(hipStreamSynchronize(lane_stream[l]));
      for (auto& kv : gcache[l]) ECHK(hipGraphExecDestroy(kv.second));
      gcache[l].clear();
    }
  }
  chunk = ch;
  nlanes = lanes;
  graphs = use_graphs;
  return true;
}

void ns3d_engine::solve(EngineRequest& req) {
  EngineRequest* r = &req;
  solve_many(&r, 1);
}

void ns3d_engine::solve_many(EngineRequest* const* reqs, int n) {
  std::unique_lock<std::mutex> lk(mu);
  if (!fatal.empty()) throw std::runtime_error("solve engine failed: " + fatal);
  if (stopping) throw std::runtime_error("solve engine is shutting down");
  std::condition_variable cv;
  for (int i = 0; i < n; ++i) {
    reqs[i]->done = false;
    reqs[i]->waiter = &cv;
    pending.push_back(reqs[i]);
  }
  waiters.insert(&cv);
  cv_work.notify_one();
  auto all_done = [&] {
    for (int i = 0; i < n; ++i)
      if (!reqs[i]->done) return false;
    return true;
  };
  cv.wait(lk, [&] { return all_done() || !fatal.empty(); });
  waiters.erase(&cv);
  for (int i = 0; i < n; ++i) reqs[i]->waiter = nullptr;
  if (!all_done()) throw std::runtime_error("solve engine failed: " + fatal);
}

void ns3d_engine::fence() {
  std::unique_lock<std::mutex> lk(mu);
  const long long target = rounds_queued;
  fence_waiters.fetch_add(1);
  cv_done.wait(lk, [&] { return rounds_done >= target || !fatal.empty(); });
  fence_waiters.fetch_sub(1);
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
  v[23] = ch_ms;
  v[24] = ch_n;
  v[25] = ch_bytes;
  v[26] = host_queue_ms;
  v[27] = host_process_ms;
  v[28] = host_wait_ms;
  v[29] = tiles_live;
  v[30] = tiles_all;
  for (int q = 0; q < count && q < 32; ++q) out[q] = v[q];
}

void ns3d_engine::wait_round(int buf) {
  for (int l = 0; l < kMaxLanes; ++l)
    if (lane_in[buf][l]) ECHK(hipEventSynchronize(lane_done[buf][l]));
}

void ns3d_engine::fail(const std::string& msg) {
  std::lock_guard<std::mutex> lk(mu);
  fatal = msg;
  for (std::condition_variable* w : waiters) w->notify_one();
  cv_done.notify_all();
}

int64_t ns3d_engine::group_rows(int g) const {
  int64_t rows = 0;
  for (const EngineRequest* r : grp[g].active) rows += r->n;
  return rows;
}

template <typename T>
Batch<T> ns3d_engine::batch(int g, int q) const {
  const Group& G = grp[g];
  Batch<T> B{};
  B.d = (const SolveDesc<T>*)G.d_desc;
  B.tiles = (const TileDesc<T>*)G.d_tiles;
  B.nsolves = (int)G.active.size();
  B.ntiles = G.ntiles;
  B.q = q;
  return B;
}

// Host images of the tables (index buf) are rewritten only for a round queued
// after the one that last used them has been processed, so no copy still in
// flight reads a buffer that changed.
template <typename T>
static void tile_table(HostBuf& img, const std::vector<Tile>& tiles, const SolveDesc<T>* desc,
                       const std::vector<EngineRequest*>* reqs) {
  img.reserve(kTableHeader + tiles.size() * sizeof(TileDesc<T>));
  img.size = kTableHeader + tiles.size() * sizeof(TileDesc<T>);
  std::memset(img.p, 0, kTableHeader);
  *(int*)img.p = (int)tiles.size();
  TileDesc<T>* e = (TileDesc<T>*)(img.p + kTableHeader);
  for (size_t i = 0; i < tiles.size(); ++i) {
    const int s = tiles[i].solve;
    e[i].d = desc[s];
    e[i].begin = tiles[i].begin;
    e[i].end = tiles[i].end;
    e[i].it0 = reqs ? (*reqs)[s]->it_queued : 0;
    e[i].solve = s;
    e[i].step = tiles[i].step;
  }
}

void ns3d_engine::admit(std::vector<EngineRequest*>& reqs, int buf, int inflight) {
  // lane: the one of the request's precision with fewer rows so far
  int64_t rows[kGroups];
  for (int g = 0; g < kGroups; ++g) rows[g] = group_rows(g);
  for (EngineRequest* r : reqs) {
    const int g0 = kMaxLanes * r->prec;
    r->group = g0;
    for (int l = 1; l < nlanes; ++l)
      if (rows[g0 + l] < rows[r->group]) r->group = g0 + l;
    rows[r->group] += r->n;
  }
  // A solve retired by the last processed round may still have kernels in
  // the round in flight, on its old lane.  When a new solve reuses its state
  // (the same context's next solve) on another lane, that lane first waits
  // for the old lane's round: the stale kernels then see the old solve's
  // final status, never the new solve's.  (On the same lane, stream order
  // gives this already.)
  auto stale_lane = [&](const EngineRequest* r) {
    if (inflight < 0) return -1;
    const void* st = r->prec == 0 ? (const void*)r->d64.w.st : (const void*)r->d32.w.st;
    for (const RoundEntry& o : order[inflight])
      if (o.st == st && o.group % kMaxLanes != r->group % kMaxLanes) return o.group % kMaxLanes;
    return -1;
  };
  for (int g = 0; g < kGroups; ++g) {
    Group& G = grp[g];
    hipStream_t gs = lane_stream[g % kMaxLanes];
    std::vector<Tile> tiles;
    int ns = 0;
    for (EngineRequest* r : reqs) {
      if (r->group != g) continue;
      if (r->ready) ECHK(hipStreamWaitEvent(gs, r->ready, 0));
      const int old = stale_lane(r);
      if (old >= 0 && lane_in[inflight][old]) ECHK(hipStreamWaitEvent(gs, lane_done[inflight][old], 0));
    }
    auto init = [&](auto tag) {
      using T = decltype(tag);
      int nnew = 0;
      for (EngineRequest* r : reqs) nnew += r->group == g;
      G.h_idesc[buf].reserve(std::max(nnew, 1) * sizeof(SolveDesc<T>));
      SolveDesc<T>* hd = (SolveDesc<T>*)G.h_idesc[buf].p;
      for (EngineRequest* r : reqs) {
        if (r->group != g) continue;
        SolveDesc<T> d = desc_of<T>(r);
        make_tiles(ns, d.S.n, r->band_off, tiles, d.t0, d.t1);
        hd[ns++] = d;
      }
      if (ns == 0) return;
      tile_table(G.h_itiles[buf], tiles, hd, nullptr);
      ECHK(hipMemcpyAsync(G.d_idesc, hd, ns * sizeof(SolveDesc<T>), hipMemcpyHostToDevice, gs));
      ECHK(hipMemcpyAsync((char*)G.d_itiles - kTableHeader, G.h_itiles[buf].p, G.h_itiles[buf].size,
                          hipMemcpyHostToDevice, gs));
      Batch<T> B{};
      B.d = (const SolveDesc<T>*)G.d_idesc;
      B.tiles = (const TileDesc<T>*)G.d_itiles;
      B.nsolves = ns;
      B.ntiles = (int)tiles.size();
      ECHK(launch_bicg_init<T>(B, gs));
    };
    if (g / kMaxLanes == 0) init(double{});
    else init(float{});
    for (EngineRequest* r : reqs)
      if (r->group == g) {
        r->serial = next_serial++;
        r->it_next = 0;
        r->it_done = 0;
        r->it_queued = -1;
        r->restart_queued = false;
        live[r->serial] = r;
        G.active.push_back(r);
      }
    if (ns) G.dirty = true;
  }
}

// The tile table of the round being queued, with each active solve's
// iteration base (launch q of the chunk is iteration it0 + q); each solve is
// then expected to be a chunk further on.
void ns3d_engine::upload_tiles(int g, int buf, bool restarted) {
  Group& G = grp[g];
  if (G.active.empty()) return;
  for (EngineRequest* r : G.active) {
    r->it_queued = r->it_next;
    r->restart_queued = restarted;
    r->it_next += chunk;
  }
  auto up = [&](auto tag) {
    using T = decltype(tag);
    tile_table(G.h_tiles[buf], G.tiles, (const SolveDesc<T>*)G.descs.data(), &G.active);
  };
  if (g / kMaxLanes == 0) up(double{});
  else up(float{});
  ECHK(hipMemcpyAsync(G.d_nt, G.h_tiles[buf].p, G.h_tiles[buf].size, hipMemcpyHostToDevice,
                      lane_stream[g % kMaxLanes]));
}

// Tiles and descriptors of the group's active solves (after it changed).
void ns3d_engine::rebuild(int g, int buf) {
  Group& G = grp[g];
  G.tiles.clear();
  auto build = [&](auto tag) {
    using T = decltype(tag);
    G.descs.resize(std::max<size_t>(G.active.size(), 1) * sizeof(SolveDesc<T>));
    SolveDesc<T>* cur = (SolveDesc<T>*)G.descs.data();
    for (size_t i = 0; i < G.active.size(); ++i) {
      SolveDesc<T> d = desc_of<T>(G.active[i]);
      make_tiles((int)i, d.S.n, G.active[i]->band_off, G.tiles, d.t0, d.t1);
      G.active[i]->tiles = d.t1 - d.t0;
      cur[i] = d;
    }
    G.ntiles = (int)G.tiles.size();
    if (G.active.empty()) return;
    G.h_desc[buf].reserve(G.active.size() * sizeof(SolveDesc<T>));
    SolveDesc<T>* hd = (SolveDesc<T>*)G.h_desc[buf].p;
    std::copy(cur, cur + G.active.size(), hd);
    ECHK(hipMemcpyAsync(G.d_desc, hd, G.active.size() * sizeof(SolveDesc<T>), hipMemcpyHostToDevice,
                        lane_stream[g % kMaxLanes]));
  };
  if (g / kMaxLanes == 0) build(double{});
  else build(float{});
  G.dirty = false;
}

// Tables, restarts, one chunk on each lane, and on each lane the gather of
// the states of its solves, into buffer buf.  A lane reads back only its own
// solves' states, in its own stream order, so the lanes need not wait for
// each other: each starts its next round as soon as its own gather is done.
void ns3d_engine::queue_round(int buf, bool restart[kGroups]) {
  for (int g = 0; g < kGroups; ++g) {
    if (grp[g].dirty) rebuild(g, buf);
    // the restart kernel acts only on solves whose status asks for it
    const bool rs = restart[g] && !grp[g].active.empty();
    upload_tiles(g, buf, rs);
    if (rs) {
      hipStream_t gs = lane_stream[g % kMaxLanes];
      if (g / kMaxLanes == 0) ECHK(launch_bicg_restart<double>(batch<double>(g, 0), gs));
      else ECHK(launch_bicg_restart<float>(batch<float>(g, 0), gs));
    }
    restart[g] = false;
  }
  std::vector<RoundEntry>& ord = order[buf];
  ord.clear();
  int nact = 0;
  int off[kGroups + 1];
  for (int g = 0; g < kGroups; ++g) {
    off[g] = nact;
    for (EngineRequest* r : grp[g].active) {
      const void* st = g / kMaxLanes == 0 ? (const void*)r->d64.w.st : (const void*)r->d32.w.st;
      ord.push_back(RoundEntry{r->serial, r->it_queued, g, r->n, r->nnz, r->tiles, st});
      ++nact;
    }
  }
  off[kGroups] = nact;
  max_active = std::max(max_active, (double)nact);
  bool* on = round_on[buf];
  int ngroups = 0;
  for (int g = 0; g < kGroups; ++g) ngroups += (on[g] = !grp[g].active.empty());
  const bool sampled = !graphs || rounds_started % kSampleEvery == 0;
  round_sampled[buf] = sampled && events;
  ++rounds_started;
  if (sampled) {
    launch_direct(buf, on);
  } else {
    for (int l = 0; l < nlanes; ++l) {
      int grid[2] = {0, 0};
      for (int p = 0; p < 2; ++p)
        if (on[p * kMaxLanes + l]) grid[p] = graph_grid(grp[p * kMaxLanes + l].ntiles);
      if (grid[0] || grid[1]) ECHK(hipGraphLaunch(chunk_graph(l, grid), lane_stream[l]));
    }
  }
  n_launch += (double)chunk * 3 * ngroups;
  n_chunks += 1;
  // per lane: gather its groups' states, copy them back, mark the lane done
  for (int l = 0; l < kMaxLanes; ++l) {
    lane_in[buf][l] = false;
    for (int p = 0; p < 2; ++p) {
      const int g = p * kMaxLanes + l;
      const int ns = off[g + 1] - off[g];
      if (!on[g] || ns == 0) continue;
      // the group's descriptor table (uploaded by rebuild on this stream)
      // is in active order, as the round's entries are
      const int words = ns * (int)(sizeof(BState) / 8);
      if (p == 0)
        hipLaunchKernelGGL(gather_states_kernel<double>, dim3((words + 255) / 256), dim3(256), 0, lane_stream[l],
                           (const SolveDesc<double>*)grp[g].d_desc, h_states[buf] + off[g], ns);
      else
        hipLaunchKernelGGL(gather_states_kernel<float>, dim3((words + 255) / 256), dim3(256), 0, lane_stream[l],
                           (const SolveDesc<float>*)grp[g].d_desc, h_states[buf] + off[g], ns);
      ECHK(hipGetLastError());
      lane_in[buf][l] = true;
    }
    if (lane_in[buf][l]) ECHK(hipEventRecord(lane_done[buf][l], lane_stream[l]));
  }
  std::lock_guard<std::mutex> lk(mu);
  rounds_queued += 1;
}

// One chunk launched kernel by kernel: the middle iteration's launches carry
// dispatch events, and (chunks of 3 or more) the chunk's first A and last C.
void ns3d_engine::launch_direct(int buf, const bool* on) {
  const int qt = chunk / 2;
  const bool span = chunk >= 3;
  for (int q = 0; q < chunk; ++q)
    for (int k = 0; k < 3; ++k)
      for (int g = 0; g < kGroups; ++g) {
        if (!on[g]) continue;
        hipStream_t gs = lane_stream[g % kMaxLanes];
        hipEvent_t* ev = nullptr;
        if (!events) ev = nullptr;
        else if (q == qt) ev = tev[buf] + kEv * g + 2 * k;
        else if (span && q == 0 && k == 0) ev = tev[buf] + kEv * g + 6;
        else if (span && q == chunk - 1 && k == 2) ev = tev[buf] + kEv * g + 8;
        if (g / kMaxLanes == 0) {
          const Batch<double> B = batch<double>(g, q);
          ECHK(k == 0 ? launch_bicg_A<double>(B, gs, ev)
                      : k == 1 ? launch_bicg_B<double>(B, gs, ev) : launch_bicg_C<double>(B, gs, ev));
        } else {
          const Batch<float> B = batch<float>(g, q);
          ECHK(k == 0 ? launch_bicg_A<float>(B, gs, ev)
                      : k == 1 ? launch_bicg_B<float>(B, gs, ev) : launch_bicg_C<float>(B, gs, ev));
        }
      }
}

// A chain of chunk x (A, B, C) kernel nodes per active precision of the
// lane, in the order of the direct launches; the kernels take their tables
// from fixed device buffers and the tile counts from the table headers, so
// the graph serves every round with the same grids.
hipGraphExec_t ns3d_engine::chunk_graph(int lane, const int grid[2]) {
  const uint64_t key = ((uint64_t)chunk << 48) | ((uint64_t)grid[0] << 24) | (uint64_t)grid[1];
  auto& cache = gcache[lane];
  auto f = cache.find(key);
  if (f != cache.end()) return f->second;
  if (cache.size() >= 64) {
    ECHK(hipStreamSynchronize(lane_stream[lane]));
    for (auto& kv : cache) ECHK(hipGraphExecDestroy(kv.second));
    cache.clear();
  }
  hipGraph_t gr = nullptr;
  ECHK(hipGraphCreate(&gr, 0));
  hipGraphNode_t prev = nullptr;
  for (int q = 0; q < chunk; ++q)
    for (int k = 0; k < 3; ++k)
      for (int p = 0; p < 2; ++p) {
        if (!grid[p]) continue;
        const int g = p * kMaxLanes + lane;
        hipGraphNode_t nd = nullptr;
        if (p == 0) {
          Batch<double> B = batch<double>(g, q);
          B.ntiles_dev = grp[g].d_nt;
          ECHK(graph_add_bicg<double>(gr, k, B, grid[p], prev, &nd));
        } else {
          Batch<float> B = batch<float>(g, q);
          B.ntiles_dev = grp[g].d_nt;
          ECHK(graph_add_bicg<float>(gr, k, B, grid[p], prev, &nd));
        }
        prev = nd;
      }
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  ECHK(e);
  cache[key] = ex;
  return ex;
}

// After round buf has completed: sampled kernel times, then every solve of the
// round that is still live.  The round queued since (with base it_queued)
// advances a running solve by a chunk if that base equals its count, and
// does nothing to it otherwise (its kernels exit at their iteration check);
// a solve that asked for a restart is restarted by that round only if the
// round restarts its group, else it is flagged for the next one.  Stopped
// solves are retired and their callers woken.  Returns whether any was.
bool ns3d_engine::process_round(int buf, bool restart[kGroups]) {
  wait_round(buf);
  bool any_done = false;
  std::lock_guard<std::mutex> lk(mu);
  const std::vector<RoundEntry>& ord = order[buf];
  // algorithmic bytes of the work the round did: a solve iterated from its
  // count before the round (the last one read back) or its base, whichever
  // is later, up to its count after it; the sampled launches count the
  // solves whose iteration it0 + qt was among those
  const int qt = chunk / 2;
  double sb[kGroups][3] = {};
  double cb = 0;
  for (size_t i = 0; i < ord.size(); ++i) {
    const RoundEntry& o = ord[i];
    auto lv = live.find(o.serial);
    if (lv == live.end()) continue;  // retired by the previous round's results
    EngineRequest* r = lv->second;
    const long long lo = std::max(r->it_done, o.it0);
    const long long hi = std::min<long long>(h_states[buf][i].it_c, o.it0 + chunk);
    r->it_done = h_states[buf][i].it_c;
    if (hi <= lo) continue;
    const int vb = o.group / kMaxLanes ? 4 : 8;
    double per = 0;
    const bool at_qt = lo <= o.it0 + qt && o.it0 + qt < hi;
    for (int k = 0; k < 3; ++k) {
      const double kb = kernel_bytes_eng(o.n, o.nnz, k, vb);
      per += kb;
      if (at_qt) sb[o.group][k] += kb;
    }
    if (at_qt && round_sampled[buf]) tiles_live += o.tiles;
    cb += per * (double)(hi - lo);
  }
  const bool* on = round_on[buf];
  if (round_sampled[buf])
    for (const RoundEntry& o : ord) tiles_all += o.tiles;
  hipEvent_t ref = nullptr;
  double lo_t = 0, hi_t = 0;
  bool span_ok = chunk >= 3;
  for (int g = 0; g < kGroups && round_sampled[buf]; ++g) {
    if (!on[g]) continue;
    hipEvent_t* e = tev[buf] + kEv * g;
    for (int k = 0; k < 3; ++k) {
      float d = 0.f;
      if (sb[g][k] <= 0 || hipEventElapsedTime(&d, e[2 * k], e[2 * k + 1]) != hipSuccess) continue;
      ms[g / kMaxLanes][k] += d;
      launches[g / kMaxLanes][k] += 1;
      bytes[g / kMaxLanes][k] += sb[g][k];
    }
    if (!span_ok) continue;
    float t0 = 0.f, t1 = 0.f;
    if (!ref) {
      ref = e[6];
      if (hipEventElapsedTime(&t1, ref, e[9]) != hipSuccess) span_ok = false;
    } else if (hipEventElapsedTime(&t0, ref, e[6]) != hipSuccess || hipEventElapsedTime(&t1, ref, e[9]) != hipSuccess) {
      span_ok = false;
    }
    lo_t = std::min(lo_t, (double)t0);
    hi_t = std::max(hi_t, (double)t1);
  }
  if (round_sampled[buf] && span_ok && ref && hi_t > lo_t && cb > 0) {
    ch_ms += hi_t - lo_t;
    ch_n += 1;
    ch_bytes += cb;
  }
  for (size_t i = 0; i < ord.size(); ++i) {
    auto lv = live.find(ord[i].serial);
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
        restart[r->group] = true;
        r->it_next = st.it_c;
      }
      continue;
    }
    live.erase(lv);
    r->result = st;
    r->done = true;
    if (r->waiter) r->waiter->notify_one();  // under mu: the waiting call (and its cv) cannot have returned
    rows_x_iters += (double)r->n * (double)st.it_c;
    n_done += 1;
    any_done = true;
    auto& act = grp[r->group].active;
    auto it = std::find(act.begin(), act.end(), r);
    if (it != act.end()) {
      act.erase(it);
      grp[r->group].dirty = true;
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
    bool restart[kGroups] = {};
    int qbuf = 0;       // buffer of the next round to queue
    int inflight = -1;  // buffer of the queued round not yet processed
    for (;;) {
      std::vector<EngineRequest*> admitted;
      {
        std::unique_lock<std::mutex> lk(mu);
        if (inflight < 0) cv_work.wait(lk, [&] { return stopping || !pending.empty() || any_active(); });
        if (stopping && pending.empty() && !any_active() && inflight < 0) break;
        // room per precision (both lanes together; each lane gets at most that)
        int tiles[2] = {0, 0}, solves[2] = {0, 0};
        for (int g = 0; g < kGroups; ++g) {
          tiles[g / kMaxLanes] += grp[g].ntiles;
          solves[g / kMaxLanes] += (int)grp[g].active.size();
        }
        while (!pending.empty()) {
          EngineRequest* r = pending.front();
          const int64_t n = r->prec == 0 ? r->d64.S.n : r->d32.S.n;
          const int nt = tiles_for(n, r->band_off != nullptr);
          if (solves[r->prec] > 0 && (tiles[r->prec] + nt > kEngMaxTiles || solves[r->prec] + 1 > kEngMaxSolves))
            break;  // waits for room (first-come first-served)
          tiles[r->prec] += nt;
          solves[r->prec] += 1;
          admitted.push_back(r);
          pending.pop_front();
        }
      }
      bool queued = false;
      const auto h0 = std::chrono::steady_clock::now();
      if (!admitted.empty()) admit(admitted, qbuf, inflight);
      if (any_active()) {
        queue_round(qbuf, restart);
        queued = true;
      }
      const auto h1 = std::chrono::steady_clock::now();
      bool woke = false;
      if (inflight >= 0) {
        wait_round(inflight);
        const auto h2 = std::chrono::steady_clock::now();
        woke = process_round(inflight, restart);
        const auto h3 = std::chrono::steady_clock::now();
        host_wait_ms += std::chrono::duration<double, std::milli>(h2 - h1).count();
        host_process_ms += std::chrono::duration<double, std::milli>(h3 - h2).count();
      }
      host_queue_ms += std::chrono::duration<double, std::milli>(h1 - h0).count();
      (void)woke;  // process_round woke the callers of the solves it retired
      if (inflight >= 0 && fence_waiters.load() > 0) cv_done.notify_all();
      inflight = queued ? qbuf : -1;
      if (queued) qbuf ^= 1;
    }
  } catch (const std::exception& e) {
    fail(e.what());
  }
}
