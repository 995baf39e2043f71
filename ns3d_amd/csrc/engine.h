// engine.h -- the batched solve engine: one dispatcher thread per engine
// runs the BiCGSTAB solves submitted by any number of contexts (each driven
// by its own host thread) together, one launch per kernel phase over all of
// them (continuous batching: solves join and leave at chunk boundaries).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "kernels.h"

namespace ns3d {

// One solve handed to the engine.  The caller fills prec, the descriptor of
// that precision (its partials buffer of kPartialSlots x kMaxTilesPerSolve
// doubles included; t0/t1 are set by the engine), n, nnz, kind (0 momentum,
// 1 pressure, for the statistics) and ready (an event recorded on the
// caller's stream after the inputs were written; may be null), then blocks
// in ns3d_engine::solve until done.
struct EngineRequest {
  int prec = 0;  // 0 FP64, 1 FP32
  SolveDesc<double> d64{};
  SolveDesc<float> d32{};
  int64_t n = 0, nnz = 0;
  int kind = 0;
  // the system's plane-band offsets (host; null: contiguous tiles), valid
  // while the solve runs
  const int* band_off = nullptr;
  hipEvent_t ready = nullptr;
  BState result{};
  bool done = false;
  // engine-private: admission serial; executed iterations expected when the
  // next round is queued; iteration base of the last queued round and
  // whether that round restarted this solve's group first
  long long serial = 0;
  int group = 0;  // kMaxLanes * prec + lane
  int tiles = 0;  // tiles in the current table
  long long it_next = 0;
  long long it_done = 0;  // executed iterations as last read back
  long long it_queued = -1;
  bool restart_queued = false;
  // the condition variable of the solve_many call waiting for this solve
  std::condition_variable* waiter = nullptr;
};

// Pinned host memory (hipHostMalloc) for the tables the dispatcher uploads:
// a copy from it is a plain asynchronous DMA, where a copy from pageable
// memory is staged and returns only once the stream has reached it.  Grows
// on demand (dropping its contents), which callers do only when no copy from
// it can still be in flight.
struct HostBuf {
  char* p = nullptr;
  size_t cap = 0, size = 0;
  void reserve(size_t n);
  void release();
  ~HostBuf() { release(); }
};

}  // namespace ns3d

struct ns3d_engine {
  explicit ns3d_engine(int device, int chunk);
  ~ns3d_engine();
  // Blocks until the solve has stopped (result holds the final state);
  // throws std::runtime_error if the engine failed.
  void solve(ns3d::EngineRequest& req);
  // The same for n requests at once (admitted together, so they run in the
  // same rounds); returns when all have stopped.
  void solve_many(ns3d::EngineRequest* const* reqs, int n);
  void stats(double* out, int count);
  // Chunks are submitted as one graph per lane (graph kernels read each
  // table's size from the device), except every kSampleEvery-th round, whose
  // launches go one by one with dispatch events for the statistics; off:
  // every round is launched one by one.
  void set_graphs(bool on);
  // Chunk length, lane count (1..kMaxLanes) and graph use; only while no
  // solve is pending or running.  Returns false (nothing changed) otherwise.
  // None of these changes any result.
  bool set_schedule(int chunk, int lanes, bool graphs);
  // Returns once every round queued before the call has been processed, so
  // no queued kernel refers to a solve that has already returned.  A context
  // calls it before freeing memory its solves used.
  void fence();

  static constexpr int kMaxLanes = 4;
  int device = 0;
  int chunk = 8;
  int nlanes = 3;  // streams in use (1..kMaxLanes)
  std::atomic<bool> graphs{true};
  // contexts attached (ns3d_attach_engine / ns3d_destroy keep the count):
  // with only one, its solves run on its own stream instead (nothing to
  // batch with, and the single-solve launches are shorter)
  std::atomic<int> attached{0};
  // One stream for the work of every attached context (assembly, residuals,
  // corrections), created before the lanes' streams so that it has a HIP
  // hardware queue of its own: a context's stream that shared a queue with
  // a lane would wait behind that lane's queued chunks at every
  // synchronisation.  Null with NS3D_CTX_STREAM=0 (each context keeps its own).
  hipStream_t ctx_stream = nullptr;

 private:
  struct Group {
    std::vector<ns3d::EngineRequest*> active;
    void* d_desc = nullptr;
    void* d_tiles = nullptr;   // TileDesc<T> entries, after a kTableHeader-byte header
    int* d_nt = nullptr;       // the header: the table's tile count (read by graph launches)
    void* d_idesc = nullptr;
    void* d_itiles = nullptr;
    // host images of the tables, rewritten only for a round queued after
    // the one that last used them has completed, so no copy in flight ever
    // reads a changed buffer (two of each: a round is queued while the
    // previous one is in flight).  The tile table carries each solve's
    // iteration base and is uploaded every round.
    ns3d::HostBuf h_desc[2], h_idesc[2];
    ns3d::HostBuf h_tiles[2], h_itiles[2];
    std::vector<ns3d::Tile> tiles;  // current tiles (rebuilt when the active set changes)
    std::vector<char> descs;        // current descriptors, one per active solve
    int ntiles = 0;
    bool dirty = true;
  };
  // Solves are split into groups: group g holds precision g / kMaxLanes
  // (0 FP64, 1 FP32) on lane g % kMaxLanes, for the first `nlanes` lanes.
  // Each lane has its own stream, so the launches of the lanes of a
  // precision run concurrently and one's ramp-up and drain overlap the
  // others' work.
  static constexpr int kGroups = 2 * kMaxLanes;
  void run();
  // inflight: buffer of the round still in flight (-1: none)
  void admit(std::vector<ns3d::EngineRequest*>& reqs, int buf, int inflight);
  void rebuild(int g, int buf);
  void upload_tiles(int g, int buf, bool restarted);
  // The instantiated graph of one chunk on lane `lane` for the given grids
  // (per precision; 0: no launches), built on first use.
  hipGraphExec_t chunk_graph(int lane, const int grid[2]);
  void launch_direct(int buf, const bool* on);
  void queue_round(int buf, bool restart[kGroups]);
  // blocks until every lane of round buf has its states back on the host
  void wait_round(int buf);
  bool process_round(int buf, bool restart[kGroups]);
  template <typename T>
  ns3d::Batch<T> batch(int g, int q) const;
  void fail(const std::string& msg);
  bool any_active() const {
    for (const Group& G : grp)
      if (!G.active.empty()) return true;
    return false;
  }
  int64_t group_rows(int g) const;

  static constexpr int kSampleEvery = 8;
  // dispatch events on the sampled rounds (for stats); NS3D_ENGINE_EVENTS=0
  // turns them off (no kernel times in the statistics), e.g. under a
  // profiler that records every dispatch itself
  bool events = true;
  long long rounds_started = 0;
  bool round_sampled[2] = {};
  std::unordered_map<uint64_t, hipGraphExec_t> gcache[kMaxLanes];
  hipStream_t lane_stream[kMaxLanes] = {};
  hipStream_t stream = nullptr;  // lane 0's: gathers and completion events
  std::thread thr;
  std::mutex mu;
  // cv_work: the dispatcher waits for work; cv_done: fence() callers wait
  // for rounds (each solve_many call waits on a condition variable of its
  // own, woken only when one of its solves stops or the engine fails)
  std::condition_variable cv_work, cv_done;
  std::atomic<int> fence_waiters{0};
  std::unordered_set<std::condition_variable*> waiters;
  std::deque<ns3d::EngineRequest*> pending;
  bool stopping = false;
  std::string fatal;
  Group grp[kGroups];
  // per round buffer: gathered states, the solves in gather order, the
  // round's completion and sampled events
  ns3d::BState* h_states[2] = {};  // pinned; the lanes' gather kernels write it directly
  // (serials of the round's solves: a solve retired by the previous round's
  // results may still appear in this one, and its request is gone by then)
  struct RoundEntry {
    long long serial;
    long long it0;  // the solve's iteration base in the round
    int group;
    int64_t n, nnz;
    int tiles;
    const void* st;  // the solve's device BState (a later solve may reuse it)
  };
  std::vector<RoundEntry> order[2];
  std::unordered_map<long long, ns3d::EngineRequest*> live;
  long long next_serial = 1;
  bool round_on[2][kGroups] = {};      // groups with launches in the round
  // lane l's gather and copy of round buf done; lanes with solves in it
  hipEvent_t lane_done[2][kMaxLanes] = {};
  bool lane_in[2][kMaxLanes] = {};
  // dispatch events per group: [A, B, C of the sampled iteration][start,
  // stop], then the chunk's first A and last C [start, stop]
  static constexpr int kEv = 10;
  hipEvent_t tev[2][kGroups * kEv] = {};
  long long rounds_queued = 0, rounds_done = 0;
  std::vector<void*> allocs;
  // statistics
  double n_chunks = 0, n_launch = 0, n_done = 0, max_active = 0;
  // per precision and kernel: sampled launch milliseconds, launches, and
  // the algorithmic bytes of the solves that iterated in them; whole chunks
  // (all groups, from the first A start to the last C stop): milliseconds,
  // count, algorithmic bytes of the iterations executed in them
  double ms[2][3] = {}, launches[2][3] = {}, bytes[2][3] = {};
  double ch_ms = 0, ch_n = 0, ch_bytes = 0;
  // dispatcher thread: milliseconds admitting and queueing, waiting for the
  // round in flight, processing its states
  double host_queue_ms = 0, host_process_ms = 0, host_wait_ms = 0;
  // sampled rounds: tiles of the solves that iterated at the sampled
  // iteration, and all tiles launched in it
  double tiles_live = 0, tiles_all = 0;
  double rows_x_iters = 0;
};
