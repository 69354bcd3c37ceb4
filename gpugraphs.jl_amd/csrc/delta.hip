// delta.hip -- shortest paths by delta-stepping over the out-edge CSR of A,
// from one source (gg_shortest_path_delta) or from k sources at once
// (gg_shortest_path_delta_multi).
//
// src/algorithms/shortest_path.jl:34-83 runs Jacobi Bellman-Ford over A_T:
// every iteration relaxes every row, and the `updated` vector it allocates
// (:49-53) is never read.  Its fixpoint is the least solution of
// d[v] = min(d[v], min_u safe_add(w(u->v), d[u])) with d[source] = 0, which
// every relaxation order reaches as long as each vertex whose distance fell
// is expanded again.  With non-negative weights, delta-stepping (Meyer and
// Sanders, J. Algorithms 49, 2003) orders that work by distance: the vertices
// below a threshold T are expanded -- their out-edges relaxed push-style with
// an atomic min on the target -- until none is left below T, then T moves to
// the end of the next bucket of width delta that holds a pending vertex.
// Most vertices are expanded once or twice, where Bellman-Ford re-reads every
// entry every iteration.
//
// k sources are k independent runs that share the phases: a queue entry is
// j * n + u (query j, vertex u), each query keeps its own distance column,
// threshold, pending bitmap and stamps, and a phase expands the entries of
// every query together, so the fixed cost of a phase (its launches and the
// latency of a short queue) is paid once for all k.
//
// State (workspace): two queues (the phase's input and output), a pending
// bitmap per query (vertices whose distance fell to the query's threshold or
// above: they wait for a later bucket), and a stamp per (query, vertex) (the
// phase that last queued it, so it enters a phase's output once).  One phase
// is four launches:
//   ds_expand_kernel  tiles of up to 256 queue entries (fewer when the queue
//                     is short, to spread it over the grid), their out-degrees
//                     scanned in LDS and their edges walked 256 at a time (each lane finds
//                     its entry by binary search in the scan); an entry with
//                     more than kDsBig out-edges is cut into pieces; output
//                     entries are staged in LDS and published with one
//                     device-wide atomic per flush;
//   ds_big_kernel     the pieces, one block each (only when some vertex has
//                     more than kDsBig out-edges);
//   ds_bucket_kernel  for each query whose expansion queued nothing: its
//                     pending vertices below T + delta move to the output
//                     queue, and the smallest distance of those that stay is
//                     recorded;
//   ds_advance_kernel one block: moves the thresholds, decides done.
// Phases run U at a time in a captured graph, relaunched until done.
#include "frontier.hpp"

namespace gg {

constexpr int kDs = 256;        // threads per block = queue entries per tile
constexpr int kDsBig = 1024;    // out-degree above which an entry is cut into pieces
constexpr int kDsMaxK = 1024;   // sources per call
constexpr unsigned long long kKeyInf = ~0ull;
constexpr double kDeltaScale = 4.0;  // default delta = kDeltaScale x the mean weight

struct DsCtl {                  // shared by the k queries
  unsigned qlen[2];             // queue lengths by parity
  unsigned nbig;                // hub pieces of the running phase
  int32_t phase;                // phases run (the running phase stamps phase + 1)
  int32_t done;                 // 1: every query finished; 2: phase cap
  int32_t neg;                  // a negative (or NaN) weight was met
  int32_t active;               // queries not finished after the last phase
  int32_t pad;
  unsigned long long expanded;  // queue entries expanded
};

struct DsQ {                    // one entry per query
  double* thr;                  // distances below thr are expanded
  double* thr_next;             // a bucket step gathers distances below thr_next
  unsigned long long* minkey;   // smallest key of the pending vertices left by a bucket step
  int32_t* pushed;              // the running phase queued an entry of the query
  int32_t* gathered;            // its bucket step moved pending vertices in
  int32_t* bucket_ran;          // it took a bucket step this phase
  int32_t* qdone;
  int32_t* buckets;             // bucket steps taken
};

// Order-preserving keys of non-negative distances (the bit patterns of
// non-negative IEEE values order like unsigned integers).
template <class TD> __device__ __forceinline__ unsigned long long ds_key(TD d) {
  if constexpr (std::is_same<TD, float>::value) return (unsigned long long)__float_as_uint(d);
  else if constexpr (std::is_same<TD, double>::value)
    return (unsigned long long)__double_as_longlong(d);
  else return (unsigned long long)d;
}
template <class TD> __device__ __forceinline__ double ds_val(unsigned long long k) {
  if constexpr (std::is_same<TD, float>::value) return (double)__uint_as_float((unsigned)k);
  else if constexpr (std::is_same<TD, double>::value) return __longlong_as_double((long long)k);
  else return (double)(long long)k;
}

// Atomic min of a non-negative distance; returns the previous value.
template <class TD> __device__ __forceinline__ TD ds_atomic_min(TD* p, TD v) {
  if constexpr (std::is_same<TD, float>::value)
    return __uint_as_float(atomicMin((unsigned*)p, __float_as_uint(v)));
  else if constexpr (std::is_same<TD, double>::value)
    return __longlong_as_double((long long)atomicMin((unsigned long long*)p,
                                                     (unsigned long long)__double_as_longlong(v)));
  else if constexpr (std::is_same<TD, int64_t>::value)
    return (int64_t)atomicMin((unsigned long long*)p, (unsigned long long)v);
  else return atomicMin(p, v);
}

template <class TD> __device__ __forceinline__ TD ds_sadd(TD w, TD d) {
  const TD sat = Lim<TD>::max();
  return jmax(w, d) == sat ? sat : jadd(w, d);  // GPUGraphs_safe_add, src/algebra.jl:47-48
}

template <class TD> struct DsArgs {
  const int32_t* rowptr;   // out-edge CSR of A, 1-based
  const int32_t* colval;
  const void* val;         // weights; NULL for a unit pattern (every weight one)
  TD* dist;                // n x k, column-major
  uint32_t* queue[2];      // entries j * n + u
  uint32_t* pend;          // k bitmaps of `words` words
  int32_t* stamp;          // n x k
  uint2* big;              // pieces (entry, first edge) of the queued hubs
  DsCtl* ctl;
  DsQ q;
  int64_t n, words;
  int32_t k;
};

// Output-queue entries staged per block in LDS: one device-wide atomic per
// flush instead of one per wavefront (a counter that every wavefront of the
// 8 XCDs bumps serialises the whole phase).
constexpr int kOutBuf = 2048;
struct DsOut {
  uint32_t buf[kOutBuf];
  unsigned cnt;
  unsigned base;
};

// Stages e when push; all 64 lanes of the wavefront call it.
__device__ __forceinline__ void ds_stage(DsOut& o, bool push, uint32_t e) {
  const unsigned long long b = __ballot(push);
  if (b) {
    const int wl = threadIdx.x & 63;
    const int leader = __ffsll(b) - 1;
    unsigned base = 0;
    if (wl == leader) base = atomicAdd(&o.cnt, (unsigned)__popcll(b));
    base = __shfl(base, leader, 64);
    if (push) o.buf[base + __popcll(b & ((1ull << wl) - 1))] = e;
  }
}

// Every thread of the block calls it between rounds (a round stages at most
// kDs entries): flushes when another round might not fit, or always when
// `last`.  Ends with the staging area reusable.
__device__ void ds_flush(DsOut& o, uint32_t* out, unsigned* out_len, bool last) {
  __syncthreads();
  const unsigned n = o.cnt;
  const bool go = n > 0 && (last || n > (unsigned)(kOutBuf - kDs));
  if (go) {
    if (threadIdx.x == 0) o.base = atomicAdd(out_len, n);
    __syncthreads();
    const unsigned base = o.base;
    for (unsigned i = threadIdx.x; i < n; i += kDs) out[base + i] = o.buf[i];
    __syncthreads();
    if (threadIdx.x == 0) o.cnt = 0;
    __syncthreads();
  }
}

// Query and vertex of an entry.
template <bool MULTI>
__device__ __forceinline__ void ds_split(uint32_t e, int64_t n, int32_t& j, int64_t& u) {
  if (MULTI) {
    j = (int32_t)(e / (uint32_t)n);
    u = (int64_t)e - (int64_t)j * n;
  } else {
    j = 0;
    u = e;
  }
}

// Relaxes edge i (u -> v) of query j from distance du.  A lowered v is
// staged for the output queue when it falls below the query's threshold
// (once per phase, by its stamp), else joins the query's pending set.  All
// 64 lanes must call it.
template <class TA, class TD, bool UNIT>
__device__ __forceinline__ void ds_relax(const DsArgs<TD>& a, bool valid, int32_t j, TD du,
                                         int64_t i, double thr, int32_t stampv, DsOut& o) {
  bool push = false;
  uint32_t e = 0;
  if (valid) {
    const int64_t v = a.colval[i] - 1;
    const TD w = UNIT ? TD(1) : (TD)((const TA*)a.val)[i];
    if (!(w >= TD(0))) {
      a.ctl->neg = 1;  // benign race: every writer stores 1
    } else {
      const TD nd = ds_sadd(w, du);
      const int64_t jv = (int64_t)j * a.n + v;
      if (nd < a.dist[jv] && nd < ds_atomic_min(&a.dist[jv], nd)) {
        if ((double)nd < thr) {
          push = atomicExch(&a.stamp[jv], stampv) != stampv;
          e = (uint32_t)jv;
          if (push && __atomic_load_n(&a.q.pushed[j], __ATOMIC_RELAXED) == 0)
            __atomic_store_n(&a.q.pushed[j], 1, __ATOMIC_RELAXED);
        } else {
          uint32_t* pb = a.pend + (int64_t)j * a.words;
          if (!bit_get(pb, v)) atomicOr(&pb[v >> 5], 1u << (v & 31));
        }
      }
    }
  }
  ds_stage(o, push, e);
}

template <class TA, class TD, bool UNIT, bool MULTI>
__global__ __launch_bounds__(kDs) void ds_expand_kernel(DsArgs<TD> a, int p) {
  __shared__ int32_t s_beg[kDs];
  __shared__ int32_t s_scan[kDs + 1];
  __shared__ int32_t s_wsum[kDs / 64];
  __shared__ int32_t s_j[kDs];
  __shared__ TD s_du[kDs];
  __shared__ double s_thr[MULTI ? kDsMaxK : 1];
  __shared__ DsOut o;
  DsCtl* c = a.ctl;
  if (c->done) return;
  const int64_t qlen = c->qlen[p];
  // a short queue gets small tiles, so that its entries spread over the grid
  const int tile = (int)min<int64_t>(kDs, max<int64_t>(1, (qlen + gridDim.x - 1) / gridDim.x));
  if ((int64_t)blockIdx.x * tile >= qlen) return;
  const uint32_t* q = a.queue[p];
  uint32_t* out = a.queue[p ^ 1];
  unsigned* out_len = &c->qlen[p ^ 1];
  for (int j = threadIdx.x; j < (MULTI ? a.k : 1); j += kDs) s_thr[j] = a.q.thr[j];
  if (threadIdx.x == 0) o.cnt = 0;
  const int32_t stampv = c->phase + 1;
  const int wl = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int64_t t0 = (int64_t)blockIdx.x * tile; t0 < qlen; t0 += (int64_t)gridDim.x * tile) {
    const int64_t qi = t0 + threadIdx.x;
    int32_t beg = 0, deg = 0, j = 0;
    TD du = TD(0);
    if ((int)threadIdx.x < tile && qi < qlen) {
      const uint32_t e = q[qi];
      int64_t u;
      ds_split<MULTI>(e, a.n, j, u);
      // u is expanded now: a pending mark from an earlier fall is served
      uint32_t* pb = a.pend + (int64_t)j * a.words;
      if (bit_get(pb, u)) atomicAnd(&pb[u >> 5], ~(1u << (u & 31)));
      du = a.dist[e];
      beg = a.rowptr[u] - 1;
      deg = a.rowptr[u + 1] - 1 - beg;
      if (deg > kDsBig) {
        const int np = (deg + kDsBig - 1) / kDsBig;
        const unsigned p0 = atomicAdd(&c->nbig, (unsigned)np);
        for (int k = 0; k < np; ++k) a.big[p0 + k] = make_uint2(e, (unsigned)(beg + k * kDsBig));
        deg = 0;
      }
    }
    int32_t incl = deg;  // exclusive scan of the tile's degrees
    for (int o2 = 1; o2 < 64; o2 <<= 1) {
      const int32_t y = __shfl_up(incl, o2, 64);
      if (wl >= o2) incl += y;
    }
    if (wl == 63) s_wsum[wv] = incl;
    __syncthreads();
    int32_t off = 0;
    for (int k = 0; k < wv; ++k) off += s_wsum[k];
    s_beg[threadIdx.x] = beg;
    s_du[threadIdx.x] = du;
    s_j[threadIdx.x] = j;
    s_scan[threadIdx.x] = off + incl - deg;
    if (threadIdx.x == kDs - 1) s_scan[kDs] = off + incl;
    __syncthreads();
    const int32_t total = s_scan[kDs];
    for (int32_t e0 = 0; e0 < total; e0 += kDs) {
      const int32_t ei = e0 + threadIdx.x;
      const bool valid = ei < total;
      int lo = 0;
      if (valid) {
        int hi = kDs;  // s_scan[lo] <= ei < s_scan[hi]; lo is never a 0-degree entry
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_scan[mid] <= ei) lo = mid; else hi = mid;
        }
      }
      const int32_t jj = s_j[lo];
      ds_relax<TA, TD, UNIT>(a, valid, jj, s_du[lo],
                             valid ? (int64_t)s_beg[lo] + (ei - s_scan[lo]) : 0,
                             s_thr[MULTI ? jj : 0], stampv, o);
      ds_flush(o, out, out_len, false);
    }
    __syncthreads();  // the next tile reuses the LDS arrays
  }
  ds_flush(o, out, out_len, true);
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&c->expanded, (unsigned long long)qlen);
}

template <class TA, class TD, bool UNIT, bool MULTI>
__global__ __launch_bounds__(kDs) void ds_big_kernel(DsArgs<TD> a, int p) {
  __shared__ DsOut o;
  DsCtl* c = a.ctl;
  if (c->done) return;
  const int64_t nbig = c->nbig;  // written by ds_expand_kernel
  if (blockIdx.x >= nbig) return;
  const int32_t stampv = c->phase + 1;
  if (threadIdx.x == 0) o.cnt = 0;
  for (int64_t k = blockIdx.x; k < nbig; k += gridDim.x) {
    const uint2 pc = a.big[k];
    int32_t j;
    int64_t u;
    ds_split<MULTI>(pc.x, a.n, j, u);
    const TD du = a.dist[pc.x];
    const double thr = a.q.thr[j];
    const int32_t stop = min((int32_t)pc.y + kDsBig, a.rowptr[u + 1] - 1);
    for (int32_t i0 = (int32_t)pc.y; i0 < stop; i0 += kDs) {
      const int32_t i = i0 + (int32_t)threadIdx.x;
      ds_relax<TA, TD, UNIT>(a, i < stop, j, du, i, thr, stampv, o);
      ds_flush(o, a.queue[p ^ 1], &c->qlen[p ^ 1], false);
    }
  }
  ds_flush(o, a.queue[p ^ 1], &c->qlen[p ^ 1], true);
}

// The bucket step of each query whose expansion queued nothing (blockIdx.y
// = query): its pending vertices below thr_next go to the output queue
// (their bits cleared; only this step touches the query's bitmap while it
// runs); the smallest key among the rest is kept.
template <class TD>
__global__ __launch_bounds__(kDs) void ds_bucket_kernel(DsArgs<TD> a, int p) {
  __shared__ unsigned s_wsum[kDs / 64];
  __shared__ unsigned s_base;
  __shared__ unsigned long long s_min[kDs / 64];
  DsCtl* c = a.ctl;
  const int32_t j = blockIdx.y;
  if (c->done || a.q.qdone[j] || a.q.pushed[j]) return;
  const double tn = a.q.thr_next[j];
  uint32_t* out = a.queue[p ^ 1];
  uint32_t* pb = a.pend + (int64_t)j * a.words;
  const TD* dj = a.dist + (int64_t)j * a.n;
  const int64_t words = a.words;
  const int wl = threadIdx.x & 63, wv = threadIdx.x >> 6;
  unsigned long long mn = kKeyInf;
  unsigned any = 0;
  const int64_t stride = (int64_t)gridDim.x * kDs;
  const int64_t wend = (words + stride - 1) / stride * stride;  // whole blocks iterate
  for (int64_t w = (int64_t)blockIdx.x * kDs + threadIdx.x; w < wend; w += stride) {
    uint32_t bits = w < words ? pb[w] : 0u;
    uint32_t sel = 0;
    for (uint32_t r = bits; r; r &= r - 1) {
      const int bi = __ffs(r) - 1;
      const TD d = dj[w * 32 + bi];
      if ((double)d < tn) sel |= 1u << bi;
      else mn = min(mn, ds_key(d));
    }
    if (sel) pb[w] = bits & ~sel;
    // block-wide exclusive scan of the selected counts, one atomic per block
    const int cnt = __popc(sel);
    int incl = cnt;
    for (int o2 = 1; o2 < 64; o2 <<= 1) {
      const int y = __shfl_up(incl, o2, 64);
      if (wl >= o2) incl += y;
    }
    if (wl == 63) s_wsum[wv] = (unsigned)incl;
    __syncthreads();
    unsigned off = 0, tot = 0;
    for (int k = 0; k < kDs / 64; ++k) {
      if (k < wv) off += s_wsum[k];
      tot += s_wsum[k];
    }
    if (threadIdx.x == 0 && tot) s_base = atomicAdd(&c->qlen[p ^ 1], tot);
    __syncthreads();
    if (cnt) {
      unsigned pos = s_base + off + incl - cnt;
      const uint32_t e0 = (uint32_t)((int64_t)j * a.n + w * 32);
      for (uint32_t r = sel; r; r &= r - 1) out[pos++] = e0 + (uint32_t)(__ffs(r) - 1);
    }
    any |= tot;
    __syncthreads();  // s_wsum / s_base are rewritten next iteration
  }
  for (int o2 = 32; o2 > 0; o2 >>= 1) mn = min(mn, (unsigned long long)__shfl_xor(mn, o2, 64));
  if (wl == 0) s_min[wv] = mn;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kDs / 64; ++k) mn = min(mn, s_min[k]);
    if (mn != kKeyInf) atomicMin(&a.q.minkey[j], mn);
    if (any) a.q.gathered[j] = 1;  // benign race: every writer stores 1
    if (blockIdx.x == 0) a.q.bucket_ran[j] = 1;
  }
}

// One block: per query, move the threshold after a bucket step (to the next
// bucket when it gathered, else to the bucket of the smallest pending
// distance, or finish the query when nothing is pending); then the shared
// counters.
template <class TD>
__global__ __launch_bounds__(kDs) void ds_advance_kernel(DsArgs<TD> a, int p, double delta,
                                                         int64_t max_phases) {
  __shared__ int s_act[kDs / 64];
  DsCtl* c = a.ctl;
  if (c->done) return;
  int act = 0;
  for (int j = threadIdx.x; j < a.k; j += kDs) {
    if (a.q.qdone[j]) continue;
    if (a.q.bucket_ran[j]) {
      a.q.bucket_ran[j] = 0;
      a.q.buckets[j] += 1;
      if (a.q.gathered[j]) {
        a.q.thr[j] = a.q.thr_next[j];
      } else if (a.q.minkey[j] == kKeyInf) {
        a.q.qdone[j] = 1;  // nothing queued, nothing pending
      } else {
        // no pending vertex below thr_next: the next step takes the bucket
        // of the smallest pending distance
        const double mv = ds_val<TD>(a.q.minkey[j]);
        double kb = floor(mv / delta);
        if ((kb + 1.0) * delta <= mv) kb += 1.0;
        a.q.thr[j] = kb * delta;
      }
      a.q.thr_next[j] = a.q.thr[j] + delta;
      a.q.minkey[j] = kKeyInf;
      a.q.gathered[j] = 0;
    }
    a.q.pushed[j] = 0;
    act += a.q.qdone[j] ? 0 : 1;
  }
  for (int o2 = 32; o2 > 0; o2 >>= 1) act += __shfl_xor(act, o2, 64);
  if ((threadIdx.x & 63) == 0) s_act[threadIdx.x >> 6] = act;
  __syncthreads();
  if (threadIdx.x == 0) {
    act = 0;
    for (int k = 0; k < kDs / 64; ++k) act += s_act[k];
    c->active = act;
    c->phase += 1;
    c->qlen[p] = 0;  // the consumed input becomes the next phase's output
    c->nbig = 0;
    if (c->neg) c->done = 1;
    else if (act == 0) c->done = 1;
    else if (c->phase >= max_phases) c->done = 2;
  }
}

// Query j starts at sources[j] (0-based) with distance 0.
template <class TD>
__global__ void ds_seed_kernel(DsArgs<TD> a, const int32_t* __restrict__ sources, double delta) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0) {
    DsCtl* c = a.ctl;
    c->qlen[0] = (unsigned)a.k;
    c->qlen[1] = 0;
    c->nbig = 0;
    c->phase = 0;
    c->done = 0;
    c->neg = 0;
    c->active = a.k;
    c->expanded = 0;
  }
  if (j >= a.k) return;
  const int64_t e = (int64_t)j * a.n + sources[j];
  a.dist[e] = TD(0);
  a.queue[0][j] = (uint32_t)e;
  a.q.thr[j] = delta;
  a.q.thr_next[j] = 2.0 * delta;
  a.q.minkey[j] = kKeyInf;
  a.q.pushed[j] = 0;
  a.q.gathered[j] = 0;
  a.q.bucket_ran[j] = 0;
  a.q.qdone[j] = 0;
  a.q.buckets[j] = 0;
}

template <class T> __global__ void ds_fill_kernel(T* p, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Sum of the weights (for the default delta), one pass.
template <class TA>
__global__ __launch_bounds__(kDs) void ds_weight_sum_kernel(const TA* __restrict__ val, int64_t nnz,
                                                            double* sum) {
  __shared__ double part[kDs / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kDs + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * kDs)
    s += (double)val[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kDs / 64; ++k) s += part[k];
    atomicAdd(sum, s);
  }
}

// Largest out-degree (whether hub pieces can occur), one pass over rowptr.
__global__ __launch_bounds__(kDs) void ds_maxdeg_kernel(const int32_t* __restrict__ rowptr, int64_t n,
                                                        unsigned* out) {
  unsigned mx = 0;
  for (int64_t i = (int64_t)blockIdx.x * kDs + threadIdx.x; i < n; i += (int64_t)gridDim.x * kDs)
    mx = max(mx, (unsigned)(rowptr[i + 1] - rowptr[i]));
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, (unsigned)__shfl_xor(mx, o, 64));
  if ((threadIdx.x & 63) == 0 && mx) atomicMax(out, mx);
}

// Dist type of shortest_path: Int32 for Bool values, else the value type
// (src/algorithms/shortest_path.jl:18).
static int ds_dist_type(int val_dtype) { return val_dtype == GG_BOOL ? GG_INT32 : val_dtype; }

// k runs from sources (0-based, host) into dist (n x k column-major).
static int delta_run(gg_ctx* ctx, const gg_matrix* A, const int32_t* sources, int64_t k,
                     void* dist, int dist_dtype, double delta, int64_t* phases) {
  GG_TRY(check_matrix(A));
  if (A->format != GG_CSR)
    return set_error(GG_ERR_ARGUMENT, "delta-stepping runs on the CSR of A (out-neighbour rows)");
  if (A->m != A->n) return set_error(GG_ERR_DIMENSION, "A must be square");
  if (k < 1 || k > kDsMaxK)
    return set_error(GG_ERR_ARGUMENT, "delta-stepping takes 1 to %d sources per call", kDsMaxK);
  const int64_t n = A->m, words = (n + 31) / 32;
  if ((double)k * (double)n > 4294967295.0)
    return set_error(GG_ERR_ARGUMENT, "sources x vertices (%lld x %lld) must stay below 2^32",
                     (long long)k, (long long)n);
  for (int64_t j = 0; j < k; ++j)
    if (sources[j] < 0 || sources[j] >= n)
      return set_error(GG_ERR_BOUNDS, "source %lld outside 1:%lld", (long long)sources[j] + 1,
                       (long long)n);
  if (dist_dtype != ds_dist_type(A->val_dtype))
    return set_error(GG_ERR_ARGUMENT, "dist element type must be %s of the matrix values",
                     A->val_dtype == GG_BOOL ? "Int32 for Bool" : "the type");
  if (!(delta >= 0.0)) return set_error(GG_ERR_ARGUMENT, "delta must be >= 0 (0: from the weights)");
  // every weight is one only for an all-true Bool matrix (the unit-pattern
  // flag also admits odd integers, which are weights here)
  const bool unit = (A->flags & GG_FLAG_UNIT_PATTERN) && A->val_dtype == GG_BOOL;
  const bool multi = k > 1;
  return with_type(A->val_dtype, [&](auto ta) -> int {
    typedef typename decltype(ta)::type TA;
    typedef typename std::conditional<std::is_same<TA, jbool>::value, int32_t, TA>::type TD;
    void *pq, *pst, *pbm, *pbig, *pc, *pqs;
    const size_t kn = (size_t)k * (size_t)n;
    GG_TRY(ws_get(ctx, WS_SSSP0, sizeof(uint32_t) * 2 * kn, &pq));
    GG_TRY(ws_get(ctx, WS_SSSP1, sizeof(int32_t) * kn, &pst));
    GG_TRY(ws_get(ctx, WS_SSSP2, sizeof(uint32_t) * (size_t)k * (size_t)words, &pbm));
    // an entry enters a queue at most once, so a phase cuts at most
    // k * nnz / kDsBig pieces plus one per entry of more than kDsBig out-edges
    GG_TRY(ws_get(ctx, WS_BFS5, sizeof(uint2) * (size_t)(2 * k * (A->nnz / kDsBig) + 16), &pbig));
    GG_TRY(ws_get(ctx, WS_FLAG, sizeof(DsCtl) + 2 * sizeof(double), &pc));
    // per query: thr, thr_next, minkey (8 bytes each), six int32 flags, the sources
    GG_TRY(ws_get(ctx, WS_BFS4, (size_t)k * (3 * 8 + 6 * 4 + 4) + 64, &pqs));
    DsArgs<TD> a{};
    a.rowptr = A->ptr;
    a.colval = A->idx;
    a.val = unit ? nullptr : A->val;
    a.dist = (TD*)dist;
    a.queue[0] = (uint32_t*)pq;
    a.queue[1] = a.queue[0] + kn;
    a.pend = (uint32_t*)pbm;
    a.stamp = (int32_t*)pst;
    a.big = (uint2*)pbig;
    a.ctl = (DsCtl*)pc;
    a.n = n;
    a.words = words;
    a.k = (int32_t)k;
    a.q.thr = (double*)pqs;
    a.q.thr_next = a.q.thr + k;
    a.q.minkey = (unsigned long long*)(a.q.thr_next + k);
    a.q.pushed = (int32_t*)(a.q.minkey + k);
    a.q.gathered = a.q.pushed + k;
    a.q.bucket_ran = a.q.gathered + k;
    a.q.qdone = a.q.bucket_ran + k;
    a.q.buckets = a.q.qdone + k;
    int32_t* d_src = a.q.buckets + k;
    // sources through the context's pinned page when they fit (the copy
    // stays asynchronous), else a synchronous copy
    if ((size_t)k * sizeof(int32_t) <= 1024) {
      memcpy(ctx->host_pinned, sources, (size_t)k * sizeof(int32_t));
      GG_HIP(hipMemcpyAsync(d_src, ctx->host_pinned, (size_t)k * sizeof(int32_t),
                            hipMemcpyHostToDevice, ctx->stream));
    } else {
      GG_HIP(hipMemcpyAsync(d_src, sources, (size_t)k * sizeof(int32_t), hipMemcpyHostToDevice,
                            ctx->stream));
    }
    // pre-pass: the largest out-degree (are hub pieces possible?) and, for
    // the default delta, the mean weight; one readback, into the pinned page
    // after the copy above has read it (stream order)
    double* sum = (double*)((char*)pc + sizeof(DsCtl));
    unsigned* maxdeg = (unsigned*)(sum + 1);
    const bool need_mean = delta == 0.0 && !unit && A->nnz > 0;
    hipLaunchKernelGGL(ds_fill_kernel<double>, dim3(1), dim3(64), 0, ctx->stream, sum, (int64_t)2,
                       0.0);
    hipLaunchKernelGGL(ds_maxdeg_kernel, dim3((unsigned)ctx->num_cus * 4), dim3(kDs), 0,
                       ctx->stream, A->ptr, n, maxdeg);
    if (need_mean)
      hipLaunchKernelGGL(ds_weight_sum_kernel<TA>, dim3((unsigned)ctx->num_cus * 4), dim3(kDs), 0,
                         ctx->stream, (const TA*)A->val, A->nnz, sum);
    GG_HIP(hipGetLastError());
    GG_HIP(hipMemcpyAsync(ctx->host_pinned, sum, 2 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    GG_HIP(hipStreamSynchronize(ctx->stream));
    const double wsum = ((const double*)ctx->host_pinned)[0];
    const bool hubs = *(const unsigned*)((const double*)ctx->host_pinned + 1) > (unsigned)kDsBig;
    double d = delta;
    if (d == 0.0) {  // buckets a few edges deep
      const double mean = need_mean ? wsum / (double)A->nnz : 1.0;
      d = kDeltaScale * (mean > 0.0 ? mean : 1.0);
    }
    // phases stamp Int32s: the cap stays far below 2^31
    const int64_t max_phases = std::min<int64_t>(8 * n + 1024, int64_t(1) << 30);
    hipLaunchKernelGGL(ds_fill_kernel<TD>, dim3(blocks_for((int64_t)kn, kDs)), dim3(kDs), 0,
                       ctx->stream, (TD*)dist, (int64_t)kn, Lim<TD>::max());
    hipLaunchKernelGGL(ds_fill_kernel<int32_t>, dim3(blocks_for((int64_t)kn, kDs)), dim3(kDs), 0,
                       ctx->stream, a.stamp, (int64_t)kn, 0);
    hipLaunchKernelGGL(ds_fill_kernel<uint32_t>, dim3(blocks_for(k * words, kDs)), dim3(kDs), 0,
                       ctx->stream, a.pend, k * words, 0u);
    hipLaunchKernelGGL(ds_seed_kernel<TD>, dim3(blocks_for(k, kDs)), dim3(kDs), 0, ctx->stream, a,
                       (const int32_t*)d_src, d);
    GG_HIP(hipGetLastError());
    const unsigned nb_exp = (unsigned)ctx->num_cus * 8;
    const unsigned nb_big = (unsigned)ctx->num_cus * 2;
    const unsigned bx = std::max<unsigned>(
        1u, std::min<unsigned>(blocks_for(words, kDs),
                               std::max<unsigned>(1u, (unsigned)(ctx->num_cus * 4 / k))));
    auto phase = [&](hipStream_t s, int p) {
      auto go = [&](auto unit_tag, auto multi_tag) {
        constexpr bool U = decltype(unit_tag)::value;
        constexpr bool M = decltype(multi_tag)::value;
        hipLaunchKernelGGL((ds_expand_kernel<TA, TD, U, M>), dim3(nb_exp), dim3(kDs), 0, s, a, p);
        if (hubs)
          hipLaunchKernelGGL((ds_big_kernel<TA, TD, U, M>), dim3(nb_big), dim3(kDs), 0, s, a, p);
      };
      if (unit) {
        if (multi) go(std::true_type{}, std::true_type{}); else go(std::true_type{}, std::false_type{});
      } else {
        if (multi) go(std::false_type{}, std::true_type{}); else go(std::false_type{}, std::false_type{});
      }
      hipLaunchKernelGGL(ds_bucket_kernel<TD>, dim3(bx, (unsigned)k), dim3(kDs), 0, s, a, p);
      hipLaunchKernelGGL(ds_advance_kernel<TD>, dim3(1), dim3(kDs), 0, s, a, p, d, max_phases);
    };
    const int U = ctx->graph_steps > 0 ? (ctx->graph_steps + 1) & ~1 : 8;
    if (ctx->graph_steps > 0) {
      std::vector<uint64_t> key{0x64656c7461ull, (uint64_t)U, (uint64_t)unit, (uint64_t)hubs};
      sig_add(key, *A);
      sig_add(key, a);
      sig_add(key, d);
      auto steps = [&](hipStream_t s) -> int {
        for (int i = 0; i < U; ++i) phase(s, i & 1);
        GG_HIP(hipGetLastError());
        return GG_OK;
      };
      GG_TRY(graph_loop(ctx, key, &a.ctl->done, max_phases / U + 2, steps, nullptr));
    } else {
      // host-stepped phases: U phases per check of the done flag
      for (int64_t l = 0;; ++l) {
        for (int i = 0; i < U; ++i) phase(ctx->stream, i & 1);
        GG_HIP(hipGetLastError());
        GG_HIP(hipMemcpyAsync(ctx->host_pinned, &a.ctl->done, sizeof(int32_t),
                              hipMemcpyDeviceToHost, ctx->stream));
        GG_HIP(hipStreamSynchronize(ctx->stream));
        if (*(const int32_t*)ctx->host_pinned) break;
        if (l > max_phases / U + 2)
          return set_error(GG_ERR_NOT_CONVERGED, "delta-stepping did not finish");
      }
    }
    GG_HIP(hipMemcpyAsync(ctx->host_pinned, a.ctl, sizeof(DsCtl), hipMemcpyDeviceToHost,
                          ctx->stream));
    GG_HIP(hipStreamSynchronize(ctx->stream));
    const DsCtl hc = *(const DsCtl*)ctx->host_pinned;
    if (hc.neg)
      return set_error(GG_ERR_ARGUMENT,
                       "delta-stepping needs non-negative weights on the edges it reaches");
    if (hc.done == 2)
      return set_error(GG_ERR_NOT_CONVERGED, "delta-stepping did not finish in %lld phases",
                       (long long)max_phases);
    if (phases) *phases = hc.phase;
    return GG_OK;
  });
}

}  // namespace gg

using namespace gg;

extern "C" {

int gg_shortest_path_delta(gg_ctx* ctx, const gg_matrix* A, int64_t source, void* dist,
                           int dist_dtype, double delta, int64_t* phases) {
  if (!ctx || !dist || !A) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  if (source < 1 || source > A->m)
    return set_error(GG_ERR_BOUNDS, "source %lld outside 1:%lld", (long long)source,
                     (long long)A->m);
  const int32_t s0 = (int32_t)(source - 1);
  timer_begin(ctx);
  const int st = delta_run(ctx, A, &s0, 1, dist, dist_dtype, delta, phases);
  timer_end(ctx);
  return st;
}

int gg_shortest_path_delta_multi(gg_ctx* ctx, const gg_matrix* A, const int32_t* sources,
                                 int64_t k, void* dist, int dist_dtype, double delta,
                                 int64_t* phases) {
  if (!ctx || !dist || !A || !sources) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  if (k < 1 || k > kDsMaxK)
    return set_error(GG_ERR_ARGUMENT, "delta-stepping takes 1 to %d sources per call", kDsMaxK);
  std::vector<int32_t> s0((size_t)k);
  for (int64_t j = 0; j < k; ++j) s0[j] = sources[j] - 1;  // range checked by delta_run
  timer_begin(ctx);
  const int st = delta_run(ctx, A, s0.data(), k, dist, dist_dtype, delta, phases);
  timer_end(ctx);
  return st;
}

}  // extern "C"
