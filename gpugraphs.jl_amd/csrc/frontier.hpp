// frontier.hpp -- bitmap frontiers and row walks shared by the graph loops.
#pragma once

#include <string.h>

#include <algorithm>

#include "internal.hpp"

namespace gg {

__device__ __forceinline__ bool bit_get(const uint32_t* bm, int64_t v) {
  return (bm[v >> 5] >> (v & 31)) & 1u;
}
// Sets bit v; true when this call set it (it was clear before).
__device__ __forceinline__ bool bit_claim(uint32_t* bm, int64_t v) {
  const uint32_t b = 1u << (v & 31);
  return !(atomicOr(&bm[v >> 5], b) & b);
}

// Block-wide (256 threads, all of them calling) counterparts of a per-lane
// atomic on one counter: one atomic per block instead of one per lane or
// wavefront (a single address serialises them).
constexpr int kBlockClaimThreads = 256;
// Exclusive position of this thread's c items in a queue of length *counter.
__device__ __forceinline__ unsigned block_claim(unsigned c, unsigned* counter) {
  __shared__ unsigned wsum[kBlockClaimThreads / 64], wbase[kBlockClaimThreads / 64];
  const int wl = threadIdx.x & 63, wv = threadIdx.x >> 6;
  unsigned incl = c;
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned y = __shfl_up(incl, o, 64);
    if (wl >= o) incl += y;
  }
  if (wl == 63) wsum[wv] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = 0;
    for (int i = 0; i < kBlockClaimThreads / 64; ++i) {
      wbase[i] = t;
      t += wsum[i];
    }
    const unsigned b = t ? atomicAdd(counter, t) : 0u;
    for (int i = 0; i < kBlockClaimThreads / 64; ++i) wbase[i] += b;
  }
  __syncthreads();
  return wbase[wv] + incl - c;
}
// *target += the block's sum of v.
__device__ __forceinline__ void block_add(unsigned long long v, unsigned long long* target) {
  __shared__ unsigned long long ws[kBlockClaimThreads / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < kBlockClaimThreads / 64; ++i) t += ws[i];
    if (t) atomicAdd(target, t);
  }
}
// Words of a bitmap per thread in the frontier scans (4 per thread, a block
// covers 1024 coalesced words).
constexpr int kScanWords = 4;

// Sets bit (base + lane) for every lane of the wavefront with `on`, where base
// is the wavefront's first id, a multiple of 64 (the internal ids of one
// wavefront's 64 slots): one OR per 32-bit word instead of one atomic per
// lane on a shared word.  All 64 lanes must call it.
__device__ __forceinline__ void wave_set_bits(uint32_t* bm, int64_t base, bool on) {
  const unsigned long long b = __ballot(on);
  if ((threadIdx.x & 63) == 0 && b) {
    const uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32);
    if (lo) atomicOr(&bm[base >> 5], lo);
    if (hi) atomicOr(&bm[(base >> 5) + 1], hi);
  }
}

// *f = 1 when any lane of the wavefront has v.  The flag is read first and
// written only while it is still 0, so the wavefronts of a launch do not
// queue on one address (an atomic per wavefront on one counter costs about
// as much as the whole level).  All 64 lanes must call it.
__device__ __forceinline__ void wave_flag(unsigned int* f, bool v) {
  if (__ballot(v) && (threadIdx.x & 63) == 0 && __atomic_load_n(f, __ATOMIC_RELAXED) == 0u)
    __atomic_store_n(f, 1u, __ATOMIC_RELAXED);
}
__device__ __forceinline__ void set_flag(unsigned int* f) {
  if (__atomic_load_n(f, __ATOMIC_RELAXED) == 0u) __atomic_store_n(f, 1u, __ATOMIC_RELAXED);
}
// *m |= bits, skipping the atomic when every bit is already set.
__device__ __forceinline__ void or_new_bits(unsigned long long* m, unsigned long long bits) {
  if (bits & ~__atomic_load_n(m, __ATOMIC_RELAXED)) atomicOr(m, bits);
}

// Value of a stored entry as an or-and edge: (value & 1) != 0, i.e. Bool true
// or an odd integer; with curr in {0,1} this is exactly when
// GPUGraphs_band(value, curr) == 1 (src/algebra.jl:98).
template <class TA> __device__ __forceinline__ bool edge_on(const TA* val, int64_t i) {
  if constexpr (std::is_floating_point<TA>::value) return val[i] != TA(0);
  else return (val[i] & TA(1)) != TA(0);
}

// Location of a matrix row in a SELL hybrid: a slot, or a heavy row.
struct SellView {
  const int32_t* slice_ptr;
  const int32_t* colval;
  const void* val;
  const int32_t* perm;
  const int32_t* slot_len;
  int64_t C, logC, n_slots, m;
  int64_t n_heavy;
  const int32_t* heavy_row;
  const int64_t* heavy_ptr;
  const int32_t* heavy_seg;
  int64_t seg_len, n_seg;
  const int32_t* heavy_col;
  const void* heavy_val;
  // internal ids (GG_FLAG_INTERNAL_IDS): state arrays indexed by slot, then
  // slot_base + heavy index; otherwise by original row
  bool internal;
  int64_t slot_base;  // nslices * slice_size
  int64_t n_int;      // size of the state index space
  const int32_t* vinv;
};

// State index of a heavy row.
__device__ __forceinline__ int64_t heavy_rid(const SellView& g, int64_t h) {
  return g.internal ? g.slot_base + h : (int64_t)g.heavy_row[h] - 1;
}
// Original (0-based) row of a state index.
__device__ __forceinline__ int64_t orig_of(const SellView& g, int64_t id) {
  if (!g.internal) return id;
  return id < g.slot_base ? (int64_t)g.perm[id] - 1 : (int64_t)g.heavy_row[id - g.slot_base] - 1;
}
// State index of an original (0-based) row.
__device__ __forceinline__ int64_t rid_of(const SellView& g, int64_t orig) {
  return g.internal ? (int64_t)g.vinv[orig] - 1 : orig;
}

inline SellView sell_view(const gg_matrix* A) {
  SellView v{};
  v.slice_ptr = A->ptr;
  v.colval = A->idx;
  v.val = A->val;
  v.perm = A->perm;
  v.slot_len = A->slot_len;
  v.C = A->slice_size;
  v.logC = -1;
  if (A->slice_size > 0 && !(A->slice_size & (A->slice_size - 1))) {
    int l = 0;
    while ((1LL << l) < A->slice_size) ++l;
    v.logC = l;
  }
  v.n_slots = sell_slots(A);
  v.m = A->m;
  v.n_heavy = A->n_heavy;
  v.heavy_row = A->heavy_row;
  v.heavy_ptr = A->heavy_ptr;
  v.heavy_seg = A->heavy_seg;
  v.seg_len = A->heavy_seg_len;
  v.n_seg = A->n_heavy > 0 ? A->heavy_nseg : 0;
  v.heavy_col = A->heavy_col;
  v.heavy_val = A->heavy_val;
  v.internal = (A->flags & GG_FLAG_INTERNAL_IDS) != 0;
  v.slot_base = A->nslices * A->slice_size;
  v.n_int = v.internal ? v.slot_base + A->n_heavy : A->m;
  v.vinv = A->vinv;
  return v;
}

// Size of the index space traversal state lives in (rows of A_T for CSR).
inline int64_t state_size(const gg_matrix* A) {
  if (A->format == GG_SELL && (A->flags & GG_FLAG_INTERNAL_IDS))
    return A->nslices * A->slice_size + A->n_heavy;
  return A->m;
}

// Slot -> (row, first entry, entry count); false for an empty / out-of-range slot.
// Without slot lengths the count is the slice width and the walk stops at -1.
__device__ __forceinline__ bool sell_slot(const SellView& g, int64_t slot, int64_t& row,
                                          int64_t& base, int64_t& len, bool& has_len) {
  if (slot >= g.n_slots) return false;
  if (g.perm) {
    const int32_t p = g.perm[slot];
    if (p == 0) return false;
    row = p - 1;
  } else {
    row = slot;
    if (row >= g.m) return false;
  }
  int64_t sl, r;
  if (g.logC >= 0) { sl = slot >> g.logC; r = slot & (g.C - 1); }
  else { sl = slot / g.C; r = slot - sl * g.C; }
  base = (int64_t)g.slice_ptr[sl] - 1 + r;
  has_len = g.slot_len != nullptr;
  len = has_len ? (int64_t)g.slot_len[slot] : ((int64_t)g.slice_ptr[sl + 1] - g.slice_ptr[sl]) / g.C;
  return true;
}

// Heavy segment sg -> (heavy row h, entries [s, e)): heavy_seg[h] <= sg <
// heavy_seg[h + 1], segments of seg_len entries from heavy_ptr[h].
__device__ __forceinline__ void heavy_segment(const SellView& g, int64_t sg, int64_t& h,
                                              int64_t& s, int64_t& e) {
  int64_t lo = 0, hi = g.n_heavy;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (g.heavy_seg[mid] <= sg) lo = mid; else hi = mid;
  }
  h = lo;
  s = g.heavy_ptr[h] + (sg - g.heavy_seg[h]) * g.seg_len;
  e = min(s + g.seg_len, g.heavy_ptr[h + 1]);
}

// *p = min(*p, v) as one atomic step on the value's bits; true when v lowered
// *p (v < *p: a NaN never lowers, as in the non-atomic update).
template <class T> __device__ __forceinline__ bool atomic_min_lowered(T* p, T v) {
  typedef typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned int>::type U;
  static_assert(sizeof(T) == sizeof(U), "4- or 8-byte distances");
  U* q = (U*)p;
  U old = __atomic_load_n(q, __ATOMIC_RELAXED);
  const U vb = __builtin_bit_cast(U, v);
  while (v < __builtin_bit_cast(T, old)) {
    const U prev = atomicCAS(q, old, vb);
    if (prev == old) return true;
    old = prev;
  }
  return false;
}

// On a symmetric pattern, state row u of A_T lists u's out-neighbours.
struct StateRow {
  const int32_t* col;
  int64_t base, stride, len;
};
__device__ __forceinline__ StateRow state_row(const SellView& g, bool csr,
                                              const int32_t* __restrict__ rowptr, int64_t u) {
  StateRow r{g.colval, 0, 1, 0};
  if (csr) {
    r.base = rowptr[u] - 1;
    r.len = rowptr[u + 1] - rowptr[u];
  } else if (u < g.slot_base) {
    if (g.perm[u] == 0) return r;  // empty slot
    const int64_t sl = u / g.C;
    r.base = (int64_t)g.slice_ptr[sl] - 1 + (u - sl * g.C);
    r.stride = g.C;
    r.len = g.slot_len[u];
  } else {
    const int64_t h = u - g.slot_base;
    r.col = g.heavy_col;
    r.base = g.heavy_ptr[h];
    r.len = g.heavy_ptr[h + 1] - r.base;
  }
  return r;
}

// Out-edges of queue[0..qlen) (state ids of a symmetric pattern), `tile`
// (1..256) queue entries per block step: their degrees are scanned in LDS and
// each lane walks one edge per step, finding its row by binary search in the
// scan.  Rows with more than big_deg edges are appended to big[] as pieces
// (row, first entry) for expand_pieces instead.  on_edge(u, v, valid) is
// called by all threads of the block together (v 0-based; valid false on
// padding lanes).
template <class F>
__device__ void expand_tiles(const int32_t* __restrict__ queue, int64_t qlen, int tile,
                             const SellView& g, bool csr, const int32_t* __restrict__ rowptr,
                             int big_deg, int2* big, unsigned int* nbig, int64_t big_cap,
                             F&& on_edge) {
  __shared__ int64_t s_base[256];
  __shared__ int32_t s_stride[256];
  __shared__ int32_t s_u[256];
  __shared__ int32_t s_heavy[256];
  __shared__ int32_t s_scan[257];
  __shared__ int32_t s_wsum[4];
  const int wl = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int64_t t0 = (int64_t)blockIdx.x * tile; t0 < qlen; t0 += (int64_t)gridDim.x * tile) {
    const int64_t qi = t0 + threadIdx.x;
    int32_t deg = 0, u = 0;
    StateRow r{g.colval, 0, 1, 0};
    if ((int)threadIdx.x < tile && qi < qlen) {
      u = queue[qi];
      r = state_row(g, csr, rowptr, u);
      deg = (int32_t)r.len;
      if (deg > big_deg) {
        const int np = (deg + big_deg - 1) / big_deg;
        const unsigned p0 = atomicAdd(nbig, (unsigned)np);
        for (int j = 0; j < np; ++j)
          if (p0 + j < big_cap) big[p0 + j] = make_int2(u, j * big_deg);
        deg = 0;
      }
    }
    int32_t incl = deg;
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(incl, o, 64);
      if (wl >= o) incl += y;
    }
    if (wl == 63) s_wsum[wv] = incl;
    __syncthreads();
    int32_t off = 0;
    for (int q = 0; q < wv; ++q) off += s_wsum[q];
    s_u[threadIdx.x] = u;
    s_base[threadIdx.x] = r.base;
    s_stride[threadIdx.x] = (int32_t)r.stride;
    s_heavy[threadIdx.x] = r.col != g.colval;
    s_scan[threadIdx.x] = off + incl - deg;
    if (threadIdx.x == 255) s_scan[256] = off + incl;
    __syncthreads();
    const int32_t total = s_scan[256];
    for (int32_t e0 = 0; e0 < total; e0 += 256) {
      const int32_t e = e0 + threadIdx.x;
      const bool valid = e < total;
      int64_t uu = 0, v = 0;
      if (valid) {
        int lo = 0, hi = 256;  // s_scan[lo] <= e < s_scan[hi]; lo is never a 0-degree entry
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_scan[mid] <= e) lo = mid; else hi = mid;
        }
        const int32_t* col = s_heavy[lo] ? g.heavy_col : g.colval;
        uu = s_u[lo];
        v = col[s_base[lo] + (int64_t)(e - s_scan[lo]) * s_stride[lo]] - 1;
      }
      on_edge(uu, v, valid);
    }
    __syncthreads();  // the next tile reuses the LDS arrays
  }
}

// Queue entries per block step of expand_tiles: all 256 lanes' worth once the
// queue fills the grid, fewer for a short queue, so that a few rows of many
// edges spread over many blocks instead of one.
__device__ __forceinline__ int tile_for(int64_t qlen) {
  const int64_t t = (qlen + gridDim.x - 1) / gridDim.x;
  return (int)(t < 1 ? 1 : t > 256 ? 256 : t);
}

// The pieces expand_tiles left: block-strided, 256 edges per step.
template <class F>
__device__ void expand_pieces(const int2* __restrict__ big, int64_t nbig, const SellView& g,
                              bool csr, const int32_t* __restrict__ rowptr, int big_deg,
                              F&& on_edge) {
  for (int64_t pc = blockIdx.x; pc < nbig; pc += gridDim.x) {
    const int2 x = big[pc];
    const StateRow r = state_row(g, csr, rowptr, x.x);
    const int64_t hi = min((int64_t)x.y + big_deg, r.len);
    for (int64_t k0 = x.y; k0 < hi; k0 += 256) {
      const int64_t k = k0 + threadIdx.x;
      const bool valid = k < hi;
      on_edge((int64_t)x.x, valid ? (int64_t)r.col[r.base + k * r.stride] - 1 : 0, valid);
    }
  }
}

// Appends v to q (length *qn) for every lane with `add`: one atomic per
// wavefront.  All 64 lanes must call it.
__device__ __forceinline__ void wave_append(int32_t* q, unsigned int* qn, int64_t v, bool add) {
  const unsigned long long b = __ballot(add);
  if (!b) return;
  const int wl = threadIdx.x & 63;
  const int leader = __ffsll(b) - 1;
  unsigned base = 0;
  if (wl == leader) base = atomicAdd(qn, (unsigned)__popcll(b));
  base = __shfl(base, leader, 64);
  if (add) q[base + __popcll(b & ((1ull << wl) - 1))] = (int32_t)v;
}

// wave_append for the block-synchronous callbacks of expand_tiles /
// expand_pieces: one atomic per block.  All 256 threads must call it.
__device__ __forceinline__ void block_append(int32_t* q, unsigned int* qn, int64_t v, bool add) {
  const unsigned pos = block_claim(add ? 1u : 0u, qn);
  if (add) q[pos] = (int32_t)v;
}

// x in internal order: xp[slot] = x[perm[slot]] (0 for an empty slot),
// xp[slot_base + h] = x[heavy_row[h]]; rows of k values when k > 1 (B is
// m x k in b_layout, xp is row-major).
template <class T>
__global__ void gather_internal_kernel(const T* __restrict__ x, const int32_t* __restrict__ perm,
                                       int64_t slot_base, const int32_t* __restrict__ heavy_row,
                                       int64_t n_heavy, int64_t m, int64_t k, int b_rowmajor,
                                       T* __restrict__ xp) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n_int = slot_base + n_heavy;
  if (t >= n_int * k) return;
  const int64_t i = t / k, j = t - i * k;
  int64_t o;
  if (i < slot_base) o = perm ? (int64_t)perm[i] - 1 : -1;
  else o = (int64_t)heavy_row[i - slot_base] - 1;
  xp[t] = o < 0 ? T(0) : x[b_rowmajor ? o * k + j : o + j * m];
}

__global__ void fill_u32_kernel(uint32_t* p, int64_t n, uint32_t v);
__global__ void fill_i32v_kernel(int32_t* p, int64_t n, int32_t v);

}  // namespace gg
