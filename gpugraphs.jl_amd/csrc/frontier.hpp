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
