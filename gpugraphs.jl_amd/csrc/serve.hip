// serve.hip -- continuous-batching BFS / SSSP query server over HBM-resident
// graphs (gg_server_*).
//
// Each registered graph owns two engines.  Engines of large graphs (at least
// kOwnStreamNnz entries) run on HIP streams of their own; those of small
// graphs share one stream, so their short steps do not queue behind the long
// steps of a large graph on a shared hardware queue (the device has few):
//  - BFS: 64 slots of a bitmask traversal (the or-and SpMM step of
//    gg_msbfs_distances).  A query joins a free slot between two levels; its
//    source bit starts at the current level and its distances are counted from
//    there, so queries admitted at different times share every level kernel.
//    A slot is finished when its bit is absent from the whole next frontier.
//  - SSSP: 64 columns of a row-major min-plus working set (the in-place step
//    of gg_shortest_path_multi).  A query resets its column between two steps
//    and marks its source as changed; it is finished after a step that
//    lowered nothing in its column.  A step relaxes only the groups of 16
//    bytes of columns that hold a busy slot (a CSR engine: the columns up to
//    the highest busy slot; slots are taken lowest first).
// Every query is exactly the single-source BFS / Bellman-Ford from its source:
// slots never interact (bit-parallel or column-parallel), and each slot's
// traversal is the level-synchronous one.
//
// Admission: a query needs a free slot in its engine and an n-element result
// block from a fixed HBM arena (the memory budget given at creation); queries
// wait until both are free, and one that can never fit is failed with
// GG_ERR_BUDGET.  A scheduler thread polls the engines' step events,
// retires finished slots (copying their column into the result block),
// admits waiting queries and launches the next step, so the engines of
// different graphs run concurrently on the device.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batched.hpp"

namespace gg {

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct SeedList {
  int n;
  int32_t slot[64];
  int64_t src[64];  // 0-based
};

__global__ void serve_bfs_clear_kernel(uint64_t* seen, uint64_t* front, int64_t m, uint64_t mask) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < m) {
    seen[v] &= ~mask;
    front[v] &= ~mask;
  }
}

__global__ void serve_bfs_fill_kernel(int32_t* dist, int64_t m, SeedList sl) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m * sl.n) return;
  const int64_t q = i / m, v = i - q * m;
  dist[v + (int64_t)sl.slot[q] * m] = v == sl.src[q] ? 0 : 2147483647;
}

// Masks at state indices (vinv maps an original vertex to its internal id + 1
// for a relabelled SELL; NULL otherwise).
__global__ void serve_bfs_seed_kernel(SeedList sl, const int32_t* vinv, uint64_t* front,
                                      uint64_t* seen, uint32_t* fany, int32_t* start,
                                      int32_t level) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int q = 0; q < sl.n; ++q) {
    const int64_t v = vinv ? (int64_t)vinv[sl.src[q]] - 1 : sl.src[q];
    const uint64_t b = 1ull << sl.slot[q];
    const uint32_t fb = 1u << (v & 31);
    // a row outside the frontier may hold a stale word (a top-down step writes
    // only the rows it touched, so an untouched row keeps what it held two
    // steps before, bits of finished or reused slots included): start it
    // from the seed alone
    if (!(fany[v >> 5] & fb)) front[v] = 0ull;
    front[v] |= b;
    seen[v] |= b;
    fany[v >> 5] |= fb;
    start[sl.slot[q]] = level;
  }
}

__global__ void serve_zero_kernel(unsigned int* found, unsigned long long* mask) {
  *found = 0u;
  *mask = 0ull;
}

// Working set rows are state indices (ns of them).  One thread per row
// resets every seeded column of its row (a row's columns share its line, so
// the line is written once whatever the number of seeds) and marks the
// seeds' rows changed.
template <class TD>
__global__ void serve_sssp_seed_kernel(TD* d, int64_t ns, int k, SeedList sl,
                                       const int32_t* vinv, uint32_t* chp) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ns) return;
  TD* row = d + r * k;
  for (int q = 0; q < sl.n; ++q) {
    const int64_t rs = vinv ? (int64_t)vinv[sl.src[q]] - 1 : sl.src[q];
    row[sl.slot[q]] = r == rs ? TD(0) : Lim<TD>::max();
    if (r == rs) atomicOr(&chp[r >> 5], 1u << (r & 31));
  }
}

// Columns of the working set that retire together, with their result blocks.
struct ExtractList {
  int n;
  int32_t slot[64];
  void* dst[64];
};

// One thread per original vertex copies its row's retiring columns into
// their result blocks and resets them to the unreached value for the slots'
// next queries (the row's line is read and written once for all of them).
// The state rows no original vertex maps to (empty slots of a relabelled
// SELL) are never relaxed and keep the unreached value they were filled with.
template <class TD>
__global__ void serve_extract_kernel(TD* d, int64_t m, int k, ExtractList xl,
                                     const int32_t* vinv) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= m) return;
  TD* row = d + (vinv ? (int64_t)vinv[v] - 1 : v) * k;
  for (int j = 0; j < xl.n; ++j) {
    ((TD*)xl.dst[j])[v] = row[xl.slot[j]];
    row[xl.slot[j]] = Lim<TD>::max();
  }
}

// Seeds of clean slots (their columns hold the unreached value everywhere):
// the source's entry and its changed bit (one thread).
template <class TD>
__global__ void serve_sssp_seed_src_kernel(TD* d, int k, SeedList sl, const int32_t* vinv,
                                           uint32_t* chp) {
  for (int q = 0; q < sl.n; ++q) {
    const int64_t rs = vinv ? (int64_t)vinv[sl.src[q]] - 1 : sl.src[q];
    d[rs * k + sl.slot[q]] = TD(0);
    chp[rs >> 5] |= 1u << (rs & 31);
  }
}

template <class TD> __global__ void serve_fill_kernel(TD* p, int64_t n, TD v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// First-fit allocator over the result arena (256-byte granules).
class Arena {
 public:
  void init(size_t bytes) {
    total_ = bytes & ~size_t(255);
    free_.clear();
    if (total_) free_[0] = total_;
    used_ = 0;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = (bytes + 255) & ~size_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t roff = it->first + bytes;
        free_.erase(it);
        if (rest) free_[roff] = rest;
        used_ += bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t total() const { return total_; }
  size_t used() const { return used_; }

 private:
  std::map<size_t, size_t> free_;
  size_t total_ = 0, used_ = 0;
};

enum QState { Q_PENDING = 0, Q_RUNNING = 1, Q_DONE = 2, Q_FAILED = 3 };

struct Query {
  int64_t id;
  int graph, kind;
  int engine = 0;  // index into gg_server::engines
  int64_t source;
  double t_arrive = 0, t_admit = 0, t_done = 0;
  int state = Q_PENDING;
  int status = GG_OK;
  size_t off = 0, bytes = 0;
  bool has_block = false;
  hipEvent_t ready = nullptr;
  bool polled = false;
  std::string msg;  // why a failed query failed
};

// Graphs with at least this many entries get streams of their own.
constexpr int64_t kOwnStreamNnz = int64_t(1) << 24;

struct Engine {
  int kind;  // GG_QUERY_BFS / GG_QUERY_SSSP
  int slots;
  gg_matrix A;
  int64_t m;                    // vertices (result length)
  int64_t ns;                   // rows of the traversal state
  const int32_t* vinv = nullptr;  // original -> internal id + 1 (relabelled SELL)
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipEvent_t step_start = nullptr, step_done = nullptr;  // around each step, on the stream
  bool inflight = false;
  uint64_t active = 0, stepped = 0;
  // SSSP: slots whose column may hold values of a failed query (a retiring
  // column is reset to the unreached value by its extraction, so a clean
  // slot's column is all unreached and a seed writes its source alone)
  uint64_t dirty = 0;
  std::vector<int64_t> slot_q;          // query id per slot
  std::vector<int64_t> slot_steps;      // steps each SSSP slot has run (Bellman-Ford cap)
  std::vector<std::pair<int, int64_t>> seeds;  // (slot, 0-based source) waiting for the next step
  std::deque<int64_t> waiting;          // query ids
  unsigned long long* mask_h = nullptr;  // pinned: alive (BFS) / changed (SSSP) mask
  unsigned long long* mask_d = nullptr;
  unsigned int* cnt_d = nullptr;
  // BFS
  uint64_t *front = nullptr, *next = nullptr, *seen = nullptr;
  uint32_t *fany = nullptr, *nany = nullptr;
  int32_t *dist = nullptr, *start = nullptr;
  void* heavy_acc = nullptr;  // segment accumulators of heavy rows (kept zero)
  // top-down steps (symmetric unit pattern): a step whose union frontier has
  // few out-edges pushes them instead of pulling every row
  bool has_push = false;
  MsPush push{};

  int32_t level = 0;
  // SSSP
  int dist_dtype = GG_INT32;
  size_t esz = 4;
  void* buf = nullptr;                      // ns x slots row-major distances
  uint32_t *chp = nullptr, *chn = nullptr;  // rows that fell in the last / this step
  bool sparse = false;  // sparse steps (symmetric relabelled SELL): only rows next to fallen rows
  BfSparse sp{};
  std::vector<void*> allocs;
  int64_t steps = 0;
  // pacing: when the step was launched and a running mean of the steps' own
  // time on the device (start to done event on the engine's stream, so the
  // steps of other engines queued before it on a shared stream, and the
  // scheduler's polling, are not counted)
  double t_launch = 0.0, step_s = 0.0;
  // gg_server_engine_stats: device seconds of all steps, busy slots summed
  // over the steps, queries finished
  double device_s = 0.0;
  int64_t busy_slot_steps = 0, finished = 0;
};

}  // namespace gg

struct gg_server {
  int device = 0;
  char* arena_base = nullptr;
  gg::Arena arena;
  std::mutex mu;
  std::condition_variable cv_in, cv_done;
  std::deque<int64_t> inbox;
  std::unordered_map<int64_t, gg::Query> queries;
  std::vector<std::unique_ptr<gg::Engine>> engines;  // 2 per graph: BFS, SSSP
  std::vector<hipEvent_t> event_pool;
  std::thread worker;
  std::atomic<bool> running{false}, stop{false};
  int64_t next_id = 1;
  int64_t completed = 0, failed = 0;
  hipStream_t fetch_stream = nullptr;
  hipStream_t small_stream = nullptr;  // shared by the engines of small graphs
  int num_cus = 256;
  std::string worker_error;
  double pacing = 24.0;  // gg_server_set_pacing
};

namespace gg {

static int engine_alloc(gg_server* s, Engine& e, size_t bytes, void** p) {
  if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess)
    return set_error(GG_ERR_NOMEM, "server engine allocation of %zu bytes failed", bytes);
  e.allocs.push_back(*p);
  (void)s;
  return GG_OK;
}

static int make_engine(gg_server* s, const gg_matrix* A, int kind, std::unique_ptr<Engine>& out) {
  auto e = std::make_unique<Engine>();
  e->kind = kind;
  e->A = *A;
  e->m = A->m;
  e->ns = state_size(A);
  e->vinv = (A->format == GG_SELL && (A->flags & GG_FLAG_INTERNAL_IDS)) ? A->vinv : nullptr;
  // 64 bitmask slots per BFS engine, 64 columns per Bellman-Ford engine
  e->slots = 64;
  e->slot_q.assign(e->slots, -1);
  e->slot_steps.assign(e->slots, 0);
  const int64_t m = A->m, ns = e->ns;
  if (A->nnz >= kOwnStreamNnz) {
    GG_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    e->own_stream = true;
  } else {
    e->stream = s->small_stream;
  }
  GG_HIP(hipEventCreate(&e->step_start));
  GG_HIP(hipEventCreate(&e->step_done));
  GG_HIP(hipHostMalloc((void**)&e->mask_h, 64, hipHostMallocDefault));
  void* p;
  GG_TRY(engine_alloc(s, *e, 64, &p)); e->mask_d = (unsigned long long*)p;
  GG_TRY(engine_alloc(s, *e, 64, &p)); e->cnt_d = (unsigned int*)p;
  const unsigned nb = blocks_for(ns, 256);
  if (kind == GG_QUERY_BFS) {
    const int64_t words = (ns + 31) / 32;
    GG_TRY(engine_alloc(s, *e, 8 * ns, &p)); e->front = (uint64_t*)p;
    GG_TRY(engine_alloc(s, *e, 8 * ns, &p)); e->next = (uint64_t*)p;
    GG_TRY(engine_alloc(s, *e, 8 * ns, &p)); e->seen = (uint64_t*)p;
    GG_TRY(engine_alloc(s, *e, 4 * words, &p)); e->fany = (uint32_t*)p;
    GG_TRY(engine_alloc(s, *e, 4 * words, &p)); e->nany = (uint32_t*)p;
    GG_TRY(engine_alloc(s, *e, 4 * m * 64, &p)); e->dist = (int32_t*)p;
    GG_TRY(engine_alloc(s, *e, 4 * 64, &p)); e->start = (int32_t*)p;
    const size_t hb = ms_heavy_bytes(A, 1);
    GG_TRY(engine_alloc(s, *e, hb, &p)); e->heavy_acc = p;
    hipLaunchKernelGGL(serve_fill_kernel<uint32_t>, dim3(blocks_for(hb / 4, 256)), dim3(256), 0,
                       e->stream, (uint32_t*)p, (int64_t)(hb / 4), 0u);
    if (ms_push_ok(A)) {
      const size_t pb = ms_push_bytes(A, 1);
      GG_TRY(engine_alloc(s, *e, pb, &p));
      hipLaunchKernelGGL(serve_fill_kernel<uint32_t>, dim3(blocks_for((int64_t)(pb / 4), 256)),
                         dim3(256), 0, e->stream, (uint32_t*)p, (int64_t)(pb / 4), 0u);
      e->push = ms_push_layout(p, A, 1);
      e->push.alpha = 16.0;
      e->has_push = true;
    }

    hipLaunchKernelGGL(serve_fill_kernel<uint64_t>, dim3(nb), dim3(256), 0, e->stream, e->front, ns, 0ull);
    hipLaunchKernelGGL(serve_fill_kernel<uint64_t>, dim3(nb), dim3(256), 0, e->stream, e->seen, ns, 0ull);
    hipLaunchKernelGGL(serve_fill_kernel<uint32_t>, dim3(blocks_for(words, 256)), dim3(256), 0,
                       e->stream, e->fany, words, 0u);
  } else {
    e->dist_dtype = A->val_dtype == GG_BOOL ? GG_INT32 : A->val_dtype;
    e->esz = dtype_size(e->dist_dtype);
    if (e->dist_dtype != GG_INT32 && e->dist_dtype != GG_FLOAT32 && e->dist_dtype != GG_INT64 &&
        e->dist_dtype != GG_FLOAT64)
      return set_error(GG_ERR_ARGUMENT, "unsupported SSSP value type");
    GG_TRY(engine_alloc(s, *e, e->esz * ns * e->slots, &p));
    e->buf = p;
    GG_TRY(with_type(e->dist_dtype, [&](auto t) -> int {
      typedef typename decltype(t)::type TD;
      hipLaunchKernelGGL(serve_fill_kernel<TD>, dim3(blocks_for(ns * e->slots, 256)), dim3(256), 0,
                         e->stream, (TD*)p, ns * e->slots, Lim<TD>::max());
      return GG_OK;
    }));
    const int64_t words = (ns + 31) / 32;
    GG_TRY(engine_alloc(s, *e, 4 * words, &p)); e->chp = (uint32_t*)p;
    GG_TRY(engine_alloc(s, *e, 4 * words, &p)); e->chn = (uint32_t*)p;
    hipLaunchKernelGGL(serve_fill_kernel<uint32_t>, dim3(blocks_for(words, 256)), dim3(256), 0,
                       e->stream, e->chp, words, 0u);
    if (bf_sparse_ok(A)) {
      const size_t sb = bf_sparse_bytes(A);
      GG_TRY(engine_alloc(s, *e, sb, &p));
      hipLaunchKernelGGL(serve_fill_kernel<uint32_t>, dim3(blocks_for((int64_t)(sb / 4), 256)),
                         dim3(256), 0, e->stream, (uint32_t*)p, (int64_t)(sb / 4), 0u);
      e->sp = bf_sparse_layout(p, A, 16.0);
      e->sparse = true;
    }
  }
  GG_HIP(hipGetLastError());
  GG_HIP(hipStreamSynchronize(e->stream));
  out = std::move(e);
  return GG_OK;
}

static void free_engine(Engine& e) {
  if (e.stream) (void)hipStreamSynchronize(e.stream);
  for (void* p : e.allocs) (void)hipFree(p);
  e.allocs.clear();
  if (e.mask_h) (void)hipHostFree(e.mask_h);
  if (e.step_start) (void)hipEventDestroy(e.step_start);
  if (e.step_done) (void)hipEventDestroy(e.step_done);
  if (e.stream && e.own_stream) (void)hipStreamDestroy(e.stream);
}

static hipEvent_t take_event(gg_server* s) {
  if (!s->event_pool.empty()) {
    hipEvent_t ev = s->event_pool.back();
    s->event_pool.pop_back();
    return ev;
  }
  hipEvent_t ev = nullptr;
  (void)hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  return ev;
}

// Fail the query of a slot: its result block goes back to the arena and the
// slot is free again (called with s->mu held; no step of the engine is in
// flight, so nothing writes the block any more).
static void fail_slot(gg_server* s, Engine& e, int sl, int status, const char* msg) {
  Query& q = s->queries[e.slot_q[sl]];
  if (q.has_block) {
    s->arena.release(q.off, q.bytes);
    q.has_block = false;
  }
  q.state = Q_FAILED;
  q.status = status;
  q.msg = msg;
  q.t_done = now_s();
  ++s->failed;
  e.dirty |= 1ull << sl;
  e.active &= ~(1ull << sl);
  e.slot_q[sl] = -1;
}

// Retire the slots whose traversal ended in the step that just completed
// (called with s->mu held).
static int retire(gg_server* s, Engine& e) {
  const uint64_t mask = *e.mask_h;  // BFS: slots still alive; SSSP: columns that changed
  const uint64_t done = e.stepped & ~mask;
  if (!done) return GG_OK;
  const double t = now_s();
  ExtractList xl{};
  for (int sl = 0; sl < e.slots; ++sl) {
    if (!((done >> sl) & 1ull)) continue;
    Query& q = s->queries[e.slot_q[sl]];
    void* dst = s->arena_base + q.off;
    if (e.kind == GG_QUERY_BFS) {
      GG_HIP(hipMemcpyAsync(dst, e.dist + (int64_t)sl * e.m, 4 * e.m, hipMemcpyDeviceToDevice,
                            e.stream));
    } else {
      xl.slot[xl.n] = sl;
      xl.dst[xl.n] = dst;
      ++xl.n;
    }
  }
  if (xl.n) {
    GG_TRY(with_type(e.dist_dtype, [&](auto tt) -> int {
      typedef typename decltype(tt)::type TD;
      hipLaunchKernelGGL(serve_extract_kernel<TD>, dim3(blocks_for(e.m, 256)), dim3(256), 0,
                         e.stream, (TD*)e.buf, e.m, e.slots, xl, e.vinv);
      return GG_OK;
    }));
  }
  // an event per query, recorded after the copies, marks its result ready
  for (int sl = 0; sl < e.slots; ++sl) {
    if (!((done >> sl) & 1ull)) continue;
    Query& q = s->queries[e.slot_q[sl]];
    q.ready = take_event(s);
    GG_HIP(hipEventRecord(q.ready, e.stream));
    q.t_done = t;
    q.state = Q_DONE;
    ++s->completed;
    ++e.finished;
    e.active &= ~(1ull << sl);
    e.slot_q[sl] = -1;
  }
  s->cv_done.notify_all();
  return GG_OK;
}

// A Bellman-Ford column still falling after m + 16 steps has a negative
// cycle in reach (a shortest path has at most m - 1 edges): its query fails
// with GG_ERR_NOT_CONVERGED, as gg_shortest_path does (s->mu held).
static void cap_sssp_slots(gg_server* s, Engine& e) {
  if (e.kind != GG_QUERY_SSSP) return;
  bool any = false;
  for (int sl = 0; sl < e.slots; ++sl) {
    if (!((e.active >> sl) & 1ull) || e.slot_steps[sl] <= e.m + 16) continue;
    fail_slot(s, e, sl, GG_ERR_NOT_CONVERGED,
              "Bellman-Ford did not converge (a negative cycle is reachable from the source)");
    any = true;
  }
  if (any) s->cv_done.notify_all();
}

// Move waiting queries into free slots while the arena has room (s->mu held).
static void admit(gg_server* s, Engine& e) {
  const double t = now_s();
  while (!e.waiting.empty()) {
    int free_slot = -1;
    for (int sl = 0; sl < e.slots; ++sl)
      if (!((e.active >> sl) & 1ull)) { free_slot = sl; break; }
    if (free_slot < 0) return;
    Query& q = s->queries[e.waiting.front()];
    size_t off;
    if (!s->arena.alloc(q.bytes, &off)) return;  // wait for memory
    e.waiting.pop_front();
    q.off = off;
    q.has_block = true;
    q.t_admit = t;
    q.state = Q_RUNNING;
    e.active |= 1ull << free_slot;
    e.slot_q[free_slot] = q.id;
    e.slot_steps[free_slot] = 0;
    e.seeds.emplace_back(free_slot, q.source - 1);
  }
}

static int launch_step(gg_server* s, Engine& e) {
  const int64_t m = e.m, ns = e.ns;
  SeedList sl{};
  sl.n = (int)e.seeds.size();
  uint64_t seedmask = 0;
  for (int i = 0; i < sl.n; ++i) {
    sl.slot[i] = e.seeds[i].first;
    sl.src[i] = e.seeds[i].second;
    seedmask |= 1ull << sl.slot[i];
  }
  e.seeds.clear();
  GG_HIP(hipEventRecord(e.step_start, e.stream));
  hipLaunchKernelGGL(serve_zero_kernel, dim3(1), dim3(1), 0, e.stream, e.cnt_d, e.mask_d);
  if (e.kind == GG_QUERY_BFS) {
    if (sl.n) {
      hipLaunchKernelGGL(serve_bfs_clear_kernel, dim3(blocks_for(ns, 256)), dim3(256), 0,
                         e.stream, e.seen, e.front, ns, seedmask);
      hipLaunchKernelGGL(serve_bfs_fill_kernel, dim3(blocks_for(m * sl.n, 256)), dim3(256), 0,
                         e.stream, e.dist, m, sl);
      hipLaunchKernelGGL(serve_bfs_seed_kernel, dim3(1), dim3(1), 0, e.stream, sl, e.vinv,
                         e.front, e.seen, e.fany, e.start, e.level);
    }
    MsArgs a{};  // msbfs_level clears next_any
    a.m = m;
    a.k = 64;
    a.layout = GG_COL_MAJOR;
    a.front = e.front;
    a.front_any = e.fany;
    a.next = e.next;
    a.next_any = e.nany;
    a.seen = e.seen;
    a.dist = e.dist;
    a.level = e.level + 1;
    a.found = e.cnt_d;
    a.live[0] = e.active;
    a.start = e.start;
    a.alive = e.mask_d;
    a.hacc = (unsigned long long*)e.heavy_acc;
    a.hcnt = (unsigned*)(a.hacc + (e.A.format == GG_SELL ? e.A.n_heavy : 0));
    // the slots of a step sit at different levels; their union frontier
    // decides the direction, as in a batched traversal
    GG_TRY(msbfs_level(e.stream, &e.A, a, 1, e.has_push ? &e.push : nullptr, s->num_cus));
    e.level += 1;
    std::swap(e.front, e.next);
    std::swap(e.fany, e.nany);
  } else {
    if (sl.n) {
      // a dirty slot among the seeds: every seeded column is reset in one
      // pass over the rows; otherwise the sources alone are written
      const bool reset = (seedmask & e.dirty) != 0;
      e.dirty &= ~seedmask;
      GG_TRY(with_type(e.dist_dtype, [&](auto tt) -> int {
        typedef typename decltype(tt)::type TD;
        if (reset)
          hipLaunchKernelGGL(serve_sssp_seed_kernel<TD>, dim3(blocks_for(ns, 256)), dim3(256),
                             0, e.stream, (TD*)e.buf, ns, e.slots, sl, e.vinv, e.chp);
        else
          hipLaunchKernelGGL(serve_sssp_seed_src_kernel<TD>, dim3(1), dim3(1), 0, e.stream,
                             (TD*)e.buf, e.slots, sl, e.vinv, e.chp);
        return GG_OK;
      }));
    }
    const int64_t words = (ns + 31) / 32;
    hipLaunchKernelGGL(serve_fill_kernel<uint32_t>, dim3(blocks_for(words, 256)), dim3(256), 0,
                       e.stream, e.chn, words, 0u);
    // a SELL engine relaxes the groups of V = 16 / (distance bytes) columns
    // that hold a busy slot (queries finish in any order, so the busy slots
    // scatter over the 64); otherwise the columns up to the highest busy
    // slot (slots are taken lowest first), in groups of 4 so that the
    // 16-byte column kernels apply
    ColGroups cg{};
    if (e.A.format == GG_SELL) {
      const int V = 16 / (int)e.esz;
      for (int gi = 0; gi < e.slots / V; ++gi)
        if ((e.active >> (gi * V)) & ((1ull << V) - 1ull)) cg.g[cg.n++] = (uint8_t)gi;
    }
    if (cg.n > 0) {
      GG_TRY(bf_multi_step_sp(e.stream, &e.A, e.dist_dtype, e.slots, e.buf, e.chp, e.chn,
                              e.cnt_d, e.mask_d, nullptr, e.sparse ? &e.sp : nullptr, s->num_cus,
                              e.slots, &cg));
    } else {
      const int hi = 64 - __builtin_clzll(e.active | 1ull);
      const int64_t kc = std::min<int64_t>(e.slots, (hi + 3) & ~3);
      GG_TRY(bf_multi_step_sp(e.stream, &e.A, e.dist_dtype, kc, e.buf, e.chp, e.chn, e.cnt_d,
                              e.mask_d, nullptr, e.sparse ? &e.sp : nullptr, s->num_cus, e.slots));
    }
    std::swap(e.chp, e.chn);
  }
  GG_HIP(hipGetLastError());
  GG_HIP(hipMemcpyAsync(e.mask_h, e.mask_d, 8, hipMemcpyDeviceToHost, e.stream));
  GG_HIP(hipEventRecord(e.step_done, e.stream));
  e.stepped = e.active;
  e.inflight = true;
  e.t_launch = now_s();
  e.steps += 1;
  e.busy_slot_steps += __builtin_popcountll(e.active);
  for (int i = 0; i < e.slots; ++i)
    if ((e.active >> i) & 1ull) e.slot_steps[i] += 1;
  return GG_OK;
}

// Step pacing.  A step costs about the same whatever the number of its busy
// slots (it reads the whole graph, or the whole frontier of every slot), so
// an engine that steps whenever one slot is busy spends a full step on one
// query.  While other engines have steps in flight, an engine whose slots are
// not full waits pacing x (its mean step time on the device) x (its free
// fraction) x (its step time / the longest engine step time) after its last
// launch, so that arrivals join its next step instead of starting their own.
// The last factor keeps the wait for the engines whose steps are expensive
// (what pacing saves is their device time) and drops it for cheap ones, whose
// steps cost little whoever shares them; a full engine, or one alone on the
// device, steps at once.
static bool paced(gg_server* s, const Engine& e, int inflight, double max_step_s) {
  if (s->pacing <= 0.0 || inflight == 0 || max_step_s <= 0.0) return false;
  const int busy = __builtin_popcountll(e.active);
  if (busy >= e.slots) return false;
  const double wait =
      s->pacing * e.step_s * (1.0 - (double)busy / e.slots) * (e.step_s / max_step_s);
  return now_s() - e.t_launch < wait;
}

static void worker_main(gg_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    // route new queries
    while (!s->inbox.empty()) {
      const int64_t id = s->inbox.front();
      s->inbox.pop_front();
      Query& q = s->queries[id];
      Engine& e = *s->engines[q.engine];
      if (q.bytes > s->arena.total()) {
        q.state = Q_FAILED;
        q.status = GG_ERR_BUDGET;
        q.t_done = now_s();
        ++s->failed;
        s->cv_done.notify_all();
        continue;
      }
      e.waiting.push_back(id);
    }
    bool busy = false, progress = false;
    int inflight = 0;
    double max_step_s = 0.0;
    for (auto& ep : s->engines) {
      inflight += ep->inflight ? 1 : 0;
      max_step_s = std::max(max_step_s, ep->step_s);
    }
    for (auto& ep : s->engines) {
      Engine& e = *ep;
      if (e.inflight) {
        if (hipEventQuery(e.step_done) != hipSuccess) {
          busy = true;
          continue;
        }
        e.inflight = false;
        {
          float ms = 0.f;
          if (hipEventElapsedTime(&ms, e.step_start, e.step_done) == hipSuccess) {
            const double d = 1e-3 * ms;
            e.step_s = e.step_s > 0.0 ? 0.75 * e.step_s + 0.25 * d : d;
            e.device_s += d;
          }
        }
        int st = retire(s, e);
        if (st != GG_OK) s->worker_error = gg_last_error();
        cap_sssp_slots(s, e);
        progress = true;
      }
      admit(s, e);
      if (e.active && !paced(s, e, inflight, max_step_s)) {
        int st = launch_step(s, e);
        if (st != GG_OK) {
          // the step did not run: every query of the engine's slots fails
          // with the launch error and gives its block back
          s->worker_error = gg_last_error();
          (void)hipStreamSynchronize(e.stream);
          for (int sl = 0; sl < e.slots; ++sl)
            if ((e.active >> sl) & 1ull) fail_slot(s, e, sl, st, s->worker_error.c_str());
          e.seeds.clear();
          s->cv_done.notify_all();
        } else {
          busy = true;
          progress = true;
          ++inflight;
        }
      } else if (e.active) {
        busy = true;  // paced: the engine steps again shortly
      }
    }
    // on stop: finish the steps in flight; queries still waiting stay pending
    if (s->stop && !busy) break;
    if (!progress) {
      if (busy) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
      } else {
        s->cv_in.wait_for(lk, std::chrono::milliseconds(2));
      }
    }
  }
}

}  // namespace gg

using namespace gg;

extern "C" {

int gg_server_create(int device, size_t budget_bytes, gg_server** out) {
  if (!out) return set_error(GG_ERR_ARGUMENT, "out is NULL");
  clear_error();
  GG_HIP(hipSetDevice(device));
  gg_server* s = new gg_server();
  s->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) s->num_cus = prop.multiProcessorCount;
  if (budget_bytes) {
    if (hipMalloc((void**)&s->arena_base, budget_bytes) != hipSuccess) {
      delete s;
      return set_error(GG_ERR_NOMEM, "result arena of %zu bytes", budget_bytes);
    }
  }
  s->arena.init(budget_bytes);
  if (hipStreamCreateWithFlags(&s->fetch_stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&s->small_stream, hipStreamNonBlocking) != hipSuccess) {
    if (s->fetch_stream) (void)hipStreamDestroy(s->fetch_stream);
    if (s->arena_base) (void)hipFree(s->arena_base);
    delete s;
    return set_error(GG_ERR_HIP, "hipStreamCreate failed");
  }
  *out = s;
  return GG_OK;
}

int gg_server_add_graph(gg_server* s, const gg_matrix* A_T, int* graph_id) {
  if (!s || !graph_id) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  GG_TRY(check_matrix(A_T));
  if (A_T->format == GG_CSC || A_T->m != A_T->n)
    return set_error(GG_ERR_ARGUMENT, "a served graph is a square CSR or SELL A_T");
  if (s->running) return set_error(GG_ERR_ARGUMENT, "add graphs before gg_server_start");
  GG_HIP(hipSetDevice(s->device));
  std::unique_ptr<Engine> eb, es;
  GG_TRY(make_engine(s, A_T, GG_QUERY_BFS, eb));
  GG_TRY(make_engine(s, A_T, GG_QUERY_SSSP, es));
  *graph_id = (int)(s->engines.size() / 2);
  s->engines.push_back(std::move(eb));
  s->engines.push_back(std::move(es));
  return GG_OK;
}

int gg_server_set_pacing(gg_server* s, double factor) {
  if (!s) return set_error(GG_ERR_ARGUMENT, "NULL server");
  if (!(factor >= 0.0)) return set_error(GG_ERR_ARGUMENT, "pacing factor must be >= 0");
  std::lock_guard<std::mutex> g(s->mu);
  s->pacing = factor;
  return GG_OK;
}

int gg_server_start(gg_server* s) {
  if (!s) return set_error(GG_ERR_ARGUMENT, "NULL server");
  if (s->running) return GG_OK;
  s->stop = false;
  s->running = true;
  s->worker = std::thread(worker_main, s);
  return GG_OK;
}

int gg_server_submit(gg_server* s, int graph_id, int kind, int64_t source, int64_t* request_id) {
  if (!s || !request_id) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  if (graph_id < 0 || 2 * (size_t)graph_id >= s->engines.size())
    return set_error(GG_ERR_ARGUMENT, "unknown graph %d", graph_id);
  if (kind != GG_QUERY_BFS && kind != GG_QUERY_SSSP)
    return set_error(GG_ERR_ARGUMENT, "unknown query kind %d", kind);
  // On a Bool unit pattern every edge weighs one (safe_add(true, d) = d + 1),
  // so the Bellman-Ford distances are the BFS levels (Int32, typemax when
  // unreached, src/algorithms/shortest_path.jl:18): such an SSSP query runs in
  // the graph's BFS engine, where it shares the 64-slot level steps.
  Engine& es = *s->engines[2 * graph_id + 1];
  const bool hops = kind == GG_QUERY_SSSP && es.A.val_dtype == GG_BOOL &&
                    (es.A.flags & GG_FLAG_UNIT_PATTERN) != 0;
  const int engine = 2 * graph_id + (kind == GG_QUERY_BFS || hops ? 0 : 1);
  Engine& e = *s->engines[engine];
  if (kind == GG_QUERY_BFS && (e.A.val_dtype == GG_FLOAT32 || e.A.val_dtype == GG_FLOAT64) &&
      !(e.A.flags & GG_FLAG_UNIT_PATTERN))
    return set_error(GG_ERR_ARGUMENT, "or-and BFS needs Bool or integer matrix values");
  if (source < 1 || source > e.m)
    return set_error(GG_ERR_BOUNDS, "source %lld outside 1:%lld", (long long)source, (long long)e.m);
  {
    std::lock_guard<std::mutex> g(s->mu);
    Query q;
    q.id = s->next_id++;
    q.graph = graph_id;
    q.kind = kind;
    q.engine = engine;
    q.source = source;
    q.t_arrive = now_s();
    q.bytes = (size_t)e.m * (e.kind == GG_QUERY_BFS ? 4 : e.esz);
    s->queries.emplace(q.id, q);
    s->inbox.push_back(q.id);
    *request_id = q.id;
  }
  s->cv_in.notify_one();
  return GG_OK;
}

int gg_server_poll(gg_server* s, gg_query_info* out, int64_t max, int64_t* n_out, double wait_s) {
  if (!s || !n_out || (max > 0 && !out)) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto collect = [&]() -> int64_t {
    int64_t n = 0;
    for (auto& kv : s->queries) {
      Query& q = kv.second;
      if (n >= max) break;
      if ((q.state == Q_DONE || q.state == Q_FAILED) && !q.polled) {
        q.polled = true;
        gg_query_info& r = out[n++];
        r.request_id = q.id;
        r.graph = q.graph;
        r.kind = q.kind;
        r.source = q.source;
        r.status = q.status;
        r.arrival_s = q.t_arrive;
        r.admit_s = q.t_admit;
        r.complete_s = q.t_done;
      }
    }
    return n;
  };
  int64_t n = collect();
  if (n == 0 && wait_s > 0) {
    s->cv_done.wait_for(lk, std::chrono::duration<double>(wait_s));
    n = collect();
  }
  *n_out = n;
  return GG_OK;
}

static int fetch_impl(gg_server* s, int64_t request_id, void* dst, size_t bytes, hipMemcpyKind kind,
                      double wait_s) {
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->queries.find(request_id);
  if (it == s->queries.end()) return set_error(GG_ERR_ARGUMENT, "unknown request %lld", (long long)request_id);
  const double deadline = now_s() + wait_s;
  while (it->second.state == Q_PENDING || it->second.state == Q_RUNNING) {
    if (now_s() >= deadline) return set_error(GG_ERR_ARGUMENT, "request %lld not finished", (long long)request_id);
    s->cv_done.wait_for(lk, std::chrono::milliseconds(5));
    it = s->queries.find(request_id);
  }
  Query q = it->second;
  if (q.state == Q_FAILED) {
    s->queries.erase(it);
    s->cv_in.notify_one();
    if (q.status == GG_ERR_BUDGET)
      return set_error(q.status, "request %lld failed: result of %zu bytes exceeds the %zu-byte budget",
                       (long long)request_id, q.bytes, s->arena.total());
    return set_error(q.status, "request %lld failed: %s", (long long)request_id, q.msg.c_str());
  }
  lk.unlock();
  int st = GG_OK;
  if (dst) {
    if (bytes < q.bytes) st = set_error(GG_ERR_ARGUMENT, "destination holds %zu bytes, result needs %zu", bytes, q.bytes);
    else if (hipEventSynchronize(q.ready) != hipSuccess ||
             hipMemcpyAsync(dst, s->arena_base + q.off, q.bytes, kind, s->fetch_stream) != hipSuccess ||
             hipStreamSynchronize(s->fetch_stream) != hipSuccess)
      st = set_error(GG_ERR_HIP, "result copy failed");
  } else {
    (void)hipEventSynchronize(q.ready);
  }
  lk.lock();
  if (q.has_block) s->arena.release(q.off, q.bytes);
  if (q.ready) s->event_pool.push_back(q.ready);
  s->queries.erase(request_id);
  lk.unlock();
  s->cv_in.notify_one();  // freed memory may admit a waiting query
  return st;
}

int gg_server_fetch(gg_server* s, int64_t request_id, void* host_dst, size_t bytes, double wait_s) {
  if (!s) return set_error(GG_ERR_ARGUMENT, "NULL server");
  return fetch_impl(s, request_id, host_dst, bytes, hipMemcpyDeviceToHost, wait_s);
}

int gg_server_fetch_device(gg_server* s, int64_t request_id, void* dev_dst, size_t bytes,
                           double wait_s) {
  if (!s) return set_error(GG_ERR_ARGUMENT, "NULL server");
  return fetch_impl(s, request_id, dev_dst, bytes, hipMemcpyDeviceToDevice, wait_s);
}

int gg_server_release(gg_server* s, int64_t request_id, double wait_s) {
  if (!s) return set_error(GG_ERR_ARGUMENT, "NULL server");
  return fetch_impl(s, request_id, nullptr, 0, hipMemcpyDeviceToHost, wait_s);
}

int gg_server_stats(gg_server* s, gg_server_counters* out) {
  if (!s || !out) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> g(s->mu);
  out->completed = s->completed;
  out->failed = s->failed;
  out->arena_bytes = s->arena.total();
  out->arena_used = s->arena.used();
  int64_t w = 0, r = 0, steps = 0;
  for (auto& e : s->engines) {
    w += (int64_t)e->waiting.size();
    r += __builtin_popcountll(e->active);
    steps += e->steps;
  }
  out->waiting = w + (int64_t)s->inbox.size();
  out->running = r;
  out->steps = steps;
  return s->worker_error.empty() ? GG_OK : set_error(GG_ERR_HIP, "%s", s->worker_error.c_str());
}

int gg_server_engine_stats(gg_server* s, int graph_id, int kind, gg_engine_counters* out) {
  if (!s || !out) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  if (kind != GG_QUERY_BFS && kind != GG_QUERY_SSSP)
    return set_error(GG_ERR_ARGUMENT, "unknown engine kind %d", kind);
  std::lock_guard<std::mutex> g(s->mu);
  if (graph_id < 0 || 2 * (size_t)graph_id >= s->engines.size())
    return set_error(GG_ERR_ARGUMENT, "unknown graph %d", graph_id);
  const Engine& e = *s->engines[2 * graph_id + (kind == GG_QUERY_BFS ? 0 : 1)];
  out->steps = e.steps;
  out->busy_slot_steps = e.busy_slot_steps;
  out->finished = e.finished;
  out->slots = e.slots;
  out->device_s = e.device_s;
  return GG_OK;
}

int gg_server_stop(gg_server* s) {
  if (!s) return set_error(GG_ERR_ARGUMENT, "NULL server");
  if (!s->running) return GG_OK;
  s->stop = true;
  s->cv_in.notify_all();
  if (s->worker.joinable()) s->worker.join();
  s->running = false;
  for (auto& e : s->engines) (void)hipStreamSynchronize(e->stream);
  return GG_OK;
}

int gg_server_destroy(gg_server* s) {
  if (!s) return GG_OK;
  gg_server_stop(s);
  for (auto& e : s->engines) free_engine(*e);
  for (auto& kv : s->queries)
    if (kv.second.ready) (void)hipEventDestroy(kv.second.ready);
  for (hipEvent_t ev : s->event_pool) (void)hipEventDestroy(ev);
  if (s->fetch_stream) (void)hipStreamDestroy(s->fetch_stream);
  if (s->small_stream) (void)hipStreamDestroy(s->small_stream);
  if (s->arena_base) (void)hipFree(s->arena_base);
  delete s;
  return GG_OK;
}

}  // extern "C"
