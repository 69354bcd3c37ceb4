// delta.hip -- single-source shortest paths by delta-stepping over the
// out-edge CSR of A (gg_shortest_path_delta).
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
// State (workspace): two queues (the phase's input and output), a pending
// bitmap (vertices whose distance fell to T or above: they wait for a later
// bucket), and a per-vertex stamp (the phase that last queued the vertex, so
// it enters a phase's output once).  One phase is four launches:
//   ds_expand_kernel  tiles of 256 queue entries, their out-degrees scanned in
//                     LDS and their edges walked 256 at a time (each lane finds
//                     its vertex by binary search in the scan); a vertex with
//                     more than kDsBig out-edges is cut into pieces;
//   ds_big_kernel     the pieces, one block each;
//   ds_bucket_kernel  only when the phase queued nothing: the pending vertices
//                     below T + delta move to the output queue, and the
//                     smallest distance of those that stay is recorded;
//   ds_advance_kernel one thread: counts, moves T, decides done.
// Phases run U at a time in a captured graph, relaunched until done.
#include "frontier.hpp"

namespace gg {

constexpr int kDs = 256;        // threads per block = queue entries per tile
constexpr int kDsBig = 1024;    // out-degree above which a vertex is cut into pieces
constexpr unsigned long long kKeyInf = ~0ull;
constexpr double kDeltaScale = 4.0;  // default delta = kDeltaScale x the mean weight

struct DsCtl {
  unsigned qlen[2];             // queue lengths by parity
  unsigned nbig;                // hub pieces of the running phase
  int32_t phase;                // phases run (the running phase stamps phase + 1)
  unsigned long long minkey;    // smallest key of the pending vertices left by a bucket step
  double thr;                   // distances below thr are expanded
  double thr_next;              // a bucket step gathers distances below thr_next
  int32_t done;
  int32_t buckets;              // bucket steps taken
  int32_t neg;                  // a negative (or NaN) weight was met
  int32_t bucket_ran;           // the running phase took a bucket step
  int64_t expanded;             // queue entries expanded
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
  TD* dist;
  int32_t* queue[2];
  uint32_t* pend;
  int32_t* stamp;
  int2* big;               // pieces (vertex, first entry) of the queued hubs
  DsCtl* ctl;
};

// Output-queue entries staged per block in LDS: one device-wide atomic per
// flush instead of one per wavefront (a counter that every wavefront of the
// 8 XCDs bumps serialises the whole phase).
constexpr int kOutBuf = 2048;
struct DsOut {
  int32_t buf[kOutBuf];
  unsigned cnt;
  unsigned base;
};

// Stages v when push; all 64 lanes of the wavefront call it.
__device__ __forceinline__ void ds_stage(DsOut& o, bool push, int32_t v) {
  const unsigned long long b = __ballot(push);
  if (b) {
    const int wl = threadIdx.x & 63;
    const int leader = __ffsll(b) - 1;
    unsigned base = 0;
    if (wl == leader) base = atomicAdd(&o.cnt, (unsigned)__popcll(b));
    base = __shfl(base, leader, 64);
    if (push) o.buf[base + __popcll(b & ((1ull << wl) - 1))] = v;
  }
}

// Every thread of the block calls it between rounds (a round stages at most
// kDs entries): flushes when another round might not fit, or always when
// `last`.  Ends with the staging area reusable.
__device__ void ds_flush(DsOut& o, int32_t* out, unsigned* out_len, bool last) {
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

// Relaxes entry i (u -> v) from distance du.  A lowered v is staged for the
// output queue when it falls below thr (once per phase, by its stamp), else
// joins the pending set.  All 64 lanes must call it.
template <class TA, class TD, bool UNIT>
__device__ __forceinline__ void ds_relax(const DsArgs<TD>& a, bool valid, TD du, int64_t i,
                                         double thr, int32_t stampv, DsOut& o) {
  bool push = false;
  int32_t v = 0;
  if (valid) {
    v = a.colval[i] - 1;
    const TD w = UNIT ? TD(1) : (TD)((const TA*)a.val)[i];
    if (!(w >= TD(0))) {
      a.ctl->neg = 1;  // benign race: every writer stores 1
    } else {
      const TD nd = ds_sadd(w, du);
      if (nd < a.dist[v] && nd < ds_atomic_min(&a.dist[v], nd)) {
        if ((double)nd < thr) {
          push = atomicExch(&a.stamp[v], stampv) != stampv;
        } else if (!bit_get(a.pend, v)) {
          atomicOr(&a.pend[v >> 5], 1u << (v & 31));
        }
      }
    }
  }
  ds_stage(o, push, v);
}

template <class TA, class TD, bool UNIT>
__global__ __launch_bounds__(kDs) void ds_expand_kernel(DsArgs<TD> a, int p) {
  __shared__ int32_t s_beg[kDs];
  __shared__ int32_t s_scan[kDs + 1];
  __shared__ int32_t s_wsum[kDs / 64];
  __shared__ TD s_du[kDs];
  __shared__ DsOut o;
  DsCtl* c = a.ctl;
  if (c->done) return;
  const int64_t qlen = c->qlen[p];
  if ((int64_t)blockIdx.x * kDs >= qlen) return;
  const int32_t* q = a.queue[p];
  int32_t* out = a.queue[p ^ 1];
  unsigned* out_len = &c->qlen[p ^ 1];
  if (threadIdx.x == 0) o.cnt = 0;
  const double thr = c->thr;
  const int32_t stampv = c->phase + 1;
  const int wl = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int64_t t0 = (int64_t)blockIdx.x * kDs; t0 < qlen; t0 += (int64_t)gridDim.x * kDs) {
    const int64_t qi = t0 + threadIdx.x;
    int32_t beg = 0, deg = 0;
    TD du = TD(0);
    if (qi < qlen) {
      const int32_t u = q[qi];
      // u is expanded now: a pending mark from an earlier fall is served
      if (bit_get(a.pend, u)) atomicAnd(&a.pend[u >> 5], ~(1u << (u & 31)));
      du = a.dist[u];
      beg = a.rowptr[u] - 1;
      deg = a.rowptr[u + 1] - 1 - beg;
      if (deg > kDsBig) {
        const int np = (deg + kDsBig - 1) / kDsBig;
        const unsigned p0 = atomicAdd(&c->nbig, (unsigned)np);
        for (int k = 0; k < np; ++k) a.big[p0 + k] = make_int2(u, beg + k * kDsBig);
        deg = 0;
      }
    }
    int32_t incl = deg;  // exclusive scan of the tile's degrees
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(incl, o, 64);
      if (wl >= o) incl += y;
    }
    if (wl == 63) s_wsum[wv] = incl;
    __syncthreads();
    int32_t off = 0;
    for (int k = 0; k < wv; ++k) off += s_wsum[k];
    s_beg[threadIdx.x] = beg;
    s_du[threadIdx.x] = du;
    s_scan[threadIdx.x] = off + incl - deg;
    if (threadIdx.x == kDs - 1) s_scan[kDs] = off + incl;
    __syncthreads();
    const int32_t total = s_scan[kDs];
    for (int32_t e0 = 0; e0 < total; e0 += kDs) {
      const int32_t e = e0 + threadIdx.x;
      const bool valid = e < total;
      int lo = 0;
      if (valid) {
        int hi = kDs;  // s_scan[lo] <= e < s_scan[hi]; lo is never a 0-degree entry
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_scan[mid] <= e) lo = mid; else hi = mid;
        }
      }
      ds_relax<TA, TD, UNIT>(a, valid, s_du[lo], valid ? (int64_t)s_beg[lo] + (e - s_scan[lo]) : 0,
                             thr, stampv, o);
      ds_flush(o, out, out_len, false);
    }
    __syncthreads();  // the next tile reuses the LDS arrays
  }
  ds_flush(o, out, out_len, true);
  if (blockIdx.x == 0 && threadIdx.x == 0 && qlen) atomicAdd((unsigned long long*)&c->expanded,
                                                            (unsigned long long)qlen);
}

template <class TA, class TD, bool UNIT>
__global__ __launch_bounds__(kDs) void ds_big_kernel(DsArgs<TD> a, int p) {
  __shared__ DsOut o;
  DsCtl* c = a.ctl;
  if (c->done) return;
  const int64_t nbig = c->nbig;  // written by ds_expand_kernel
  if (blockIdx.x >= nbig) return;
  const double thr = c->thr;
  const int32_t stampv = c->phase + 1;
  if (threadIdx.x == 0) o.cnt = 0;
  for (int64_t k = blockIdx.x; k < nbig; k += gridDim.x) {
    const int2 pc = a.big[k];
    const TD du = a.dist[pc.x];
    const int32_t stop = min(pc.y + kDsBig, a.rowptr[pc.x + 1] - 1);
    for (int32_t i0 = pc.y; i0 < stop; i0 += kDs) {
      const int32_t i = i0 + (int32_t)threadIdx.x;
      ds_relax<TA, TD, UNIT>(a, i < stop, du, i, thr, stampv, o);
      ds_flush(o, a.queue[p ^ 1], &c->qlen[p ^ 1], false);
    }
  }
  ds_flush(o, a.queue[p ^ 1], &c->qlen[p ^ 1], true);
}

// The bucket step of a phase that queued nothing: pending vertices below
// thr_next go to the output queue (their bits cleared; this kernel is the
// bitmap's only user while it runs); the smallest key among the rest is kept.
template <class TD>
__global__ __launch_bounds__(kDs) void ds_bucket_kernel(DsArgs<TD> a, int p, int64_t n) {
  __shared__ unsigned s_wsum[kDs / 64];
  __shared__ unsigned s_base;
  DsCtl* c = a.ctl;
  if (c->done || c->qlen[p ^ 1] != 0) return;
  const double tn = c->thr_next;
  int32_t* out = a.queue[p ^ 1];
  const int64_t words = (n + 31) >> 5;
  const int wl = threadIdx.x & 63, wv = threadIdx.x >> 6;
  unsigned long long mn = kKeyInf;
  const int64_t stride = (int64_t)gridDim.x * kDs;
  const int64_t wend = (words + stride - 1) / stride * stride;  // whole wavefronts iterate
  for (int64_t w = (int64_t)blockIdx.x * kDs + threadIdx.x; w < wend; w += stride) {
    uint32_t bits = w < words ? a.pend[w] : 0u;
    uint32_t sel = 0;
    for (uint32_t r = bits; r; r &= r - 1) {
      const int bi = __ffs(r) - 1;
      const TD d = a.dist[w * 32 + bi];
      if ((double)d < tn) sel |= 1u << bi;
      else mn = min(mn, ds_key(d));
    }
    if (sel) a.pend[w] = bits & ~sel;
    // block-wide exclusive scan of the selected counts, one atomic per block
    const int cnt = __popc(sel);
    int incl = cnt;
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if (wl >= o) incl += y;
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
      for (uint32_t r = sel; r; r &= r - 1) out[pos++] = (int32_t)(w * 32 + __ffs(r) - 1);
    }
    __syncthreads();  // s_wsum / s_base are rewritten next iteration
  }
  for (int o = 32; o > 0; o >>= 1) mn = min(mn, (unsigned long long)__shfl_xor(mn, o, 64));
  if (wl == 0) {
    if (mn != kKeyInf) atomicMin(&c->minkey, mn);
    if (blockIdx.x == 0 && threadIdx.x == 0) c->bucket_ran = 1;
  }
}

template <class TD>
__global__ void ds_advance_kernel(DsCtl* c, int p, double delta, int64_t max_phases) {
  if (c->done) return;
  c->phase += 1;
  c->qlen[p] = 0;  // the consumed input becomes the next phase's output
  c->nbig = 0;
  if (c->neg || c->phase >= max_phases) {
    c->done = c->neg ? 1 : 2;
    return;
  }
  if (c->bucket_ran) {
    c->bucket_ran = 0;
    c->buckets += 1;
    if (c->qlen[p ^ 1] > 0) {
      c->thr = c->thr_next;
    } else if (c->minkey == kKeyInf) {
      c->done = 1;  // nothing queued, nothing pending
    } else {
      // no pending vertex below thr_next: the next step takes the bucket
      // of the smallest pending distance
      const double mv = ds_val<TD>(c->minkey);
      double k = floor(mv / delta);
      if ((k + 1.0) * delta <= mv) k += 1.0;
      c->thr = k * delta;
    }
    c->thr_next = c->thr + delta;
    c->minkey = kKeyInf;
  }
}

template <class TD>
__global__ void ds_seed_kernel(TD* dist, int32_t* q0, int32_t s, DsCtl* c, double delta) {
  dist[s] = TD(0);
  q0[0] = s;
  c->qlen[0] = 1;
  c->qlen[1] = 0;
  c->nbig = 0;
  c->phase = 0;
  c->minkey = kKeyInf;
  c->thr = delta;
  c->thr_next = 2.0 * delta;
  c->done = 0;
  c->buckets = 0;
  c->neg = 0;
  c->bucket_ran = 0;
  c->expanded = 0;
}

template <class T> __global__ void ds_fill_kernel(T* p, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Sum of the weights and their count (for the default delta), one pass.
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

}  // namespace gg

using namespace gg;

extern "C" {

int gg_shortest_path_delta(gg_ctx* ctx, const gg_matrix* A, int64_t source, void* dist,
                           int dist_dtype, double delta, int64_t* phases) {
  if (!ctx || !dist) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  GG_TRY(check_matrix(A));
  if (A->format != GG_CSR)
    return set_error(GG_ERR_ARGUMENT, "delta-stepping runs on the CSR of A (out-neighbour rows)");
  if (A->m != A->n) return set_error(GG_ERR_DIMENSION, "A must be square");
  if (source < 1 || source > A->m)
    return set_error(GG_ERR_BOUNDS, "source %lld outside 1:%lld", (long long)source,
                     (long long)A->m);
  if (dist_dtype != ds_dist_type(A->val_dtype))
    return set_error(GG_ERR_ARGUMENT, "dist element type must be %s of the matrix values",
                     A->val_dtype == GG_BOOL ? "Int32 for Bool" : "the type");
  if (!(delta >= 0.0)) return set_error(GG_ERR_ARGUMENT, "delta must be >= 0 (0: from the weights)");
  // every weight is one only for an all-true Bool matrix (the unit-pattern
  // flag also admits odd integers, which are weights here)
  const bool unit = (A->flags & GG_FLAG_UNIT_PATTERN) && A->val_dtype == GG_BOOL;
  const int64_t n = A->m, words = (n + 31) / 32;
  timer_begin(ctx);
  int st = with_type(A->val_dtype, [&](auto ta) -> int {
    typedef typename decltype(ta)::type TA;
    typedef typename std::conditional<std::is_same<TA, jbool>::value, int32_t, TA>::type TD;
    void *pq, *pst, *pbm, *pbig, *pc;
    GG_TRY(ws_get(ctx, WS_SSSP0, sizeof(int32_t) * 2 * (size_t)n, &pq));
    GG_TRY(ws_get(ctx, WS_SSSP1, sizeof(int32_t) * (size_t)n, &pst));
    GG_TRY(ws_get(ctx, WS_SSSP2, sizeof(uint32_t) * (size_t)words, &pbm));
    // a vertex enters a queue at most once, so a phase cuts at most
    // nnz / kDsBig pieces plus one per vertex of more than kDsBig out-edges
    GG_TRY(ws_get(ctx, WS_BFS5, sizeof(int2) * (size_t)(2 * (A->nnz / kDsBig) + 16), &pbig));
    GG_TRY(ws_get(ctx, WS_FLAG, sizeof(DsCtl) + 2 * sizeof(double), &pc));
    DsArgs<TD> a{};
    a.rowptr = A->ptr;
    a.colval = A->idx;
    a.val = unit ? nullptr : A->val;
    a.dist = (TD*)dist;
    a.queue[0] = (int32_t*)pq;
    a.queue[1] = a.queue[0] + n;
    a.pend = (uint32_t*)pbm;
    a.stamp = (int32_t*)pst;
    a.big = (int2*)pbig;
    a.ctl = (DsCtl*)pc;
    // pre-pass: the largest out-degree (are hub pieces possible?) and, for
    // the default delta, the mean weight; one readback
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
    if (d == 0.0) {
      // buckets a few edges deep
      const double mean = need_mean ? wsum / (double)A->nnz : 1.0;
      d = kDeltaScale * (mean > 0.0 ? mean : 1.0);
    }
    const int64_t max_phases = 8 * n + 1024;
    hipLaunchKernelGGL(ds_fill_kernel<TD>, dim3(blocks_for(n, kDs)), dim3(kDs), 0, ctx->stream,
                       (TD*)dist, n, Lim<TD>::max());
    hipLaunchKernelGGL(ds_fill_kernel<int32_t>, dim3(blocks_for(n, kDs)), dim3(kDs), 0,
                       ctx->stream, a.stamp, n, 0);
    hipLaunchKernelGGL(ds_fill_kernel<uint32_t>, dim3(blocks_for(words, kDs)), dim3(kDs), 0,
                       ctx->stream, a.pend, words, 0u);
    hipLaunchKernelGGL(ds_seed_kernel<TD>, dim3(1), dim3(1), 0, ctx->stream, (TD*)dist, a.queue[0],
                       (int32_t)(source - 1), a.ctl, d);
    GG_HIP(hipGetLastError());
    const unsigned nb_exp = (unsigned)ctx->num_cus * 8;
    const unsigned nb_big = (unsigned)ctx->num_cus * 2;
    const unsigned nb_bkt = std::max<unsigned>(1u, std::min<unsigned>(blocks_for(words, kDs),
                                                                      (unsigned)ctx->num_cus * 4));
    auto phase = [&](hipStream_t s, int p) {
      if (unit) {
        hipLaunchKernelGGL((ds_expand_kernel<TA, TD, true>), dim3(nb_exp), dim3(kDs), 0, s, a, p);
        if (hubs)
          hipLaunchKernelGGL((ds_big_kernel<TA, TD, true>), dim3(nb_big), dim3(kDs), 0, s, a, p);
      } else {
        hipLaunchKernelGGL((ds_expand_kernel<TA, TD, false>), dim3(nb_exp), dim3(kDs), 0, s, a, p);
        if (hubs)
          hipLaunchKernelGGL((ds_big_kernel<TA, TD, false>), dim3(nb_big), dim3(kDs), 0, s, a, p);
      }
      hipLaunchKernelGGL(ds_bucket_kernel<TD>, dim3(nb_bkt), dim3(kDs), 0, s, a, p, n);
      hipLaunchKernelGGL(ds_advance_kernel<TD>, dim3(1), dim3(1), 0, s, a.ctl, p, d, max_phases);
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
  timer_end(ctx);
  return st;
}

}  // extern "C"
