// batched.hpp -- launchers of the multi-source traversal steps, shared by the
// batched entry points and the serving engine.
#pragma once

#include "frontier.hpp"

namespace gg {

// The masks are indexed by state index (the internal id of a relabelled SELL,
// else the row); dist by original row.
struct MsArgs {
  int64_t m, k;
  int layout;                  // dist layout: GG_COL_MAJOR (m x k) or GG_ROW_MAJOR
  const uint64_t* front;       // [ns*W], ns = state_size(A)
  const uint32_t* front_any;   // [ceil(ns/32)]
  uint64_t* next;              // [ns*W]
  uint32_t* next_any;
  uint64_t* seen;              // [ns*W]
  int32_t* dist;               // original rows
  uint8_t* dist8;              // when set: byte levels (< 255), row-major m x k8, instead of dist
  int64_t k8;                  // row stride of dist8: k rounded up to a multiple of 8
  int32_t level;
  // captured loops: ctl[0] is the level (used instead of `level`), and a
  // nonzero ctl[1] (done) makes the kernel return at once; NULL otherwise
  const int32_t* ctl;
  unsigned int* found;         // set to 1 when some row got a new bit this level
  uint64_t live[8];            // bits of the sources being traversed, per word
  const int32_t* start;        // per source: level at which it was seeded (NULL: 0)
  unsigned long long* alive;   // OR of every next word, per word (NULL: not kept)
  // per heavy row of a SELL hybrid: OR of its segments' words [n_heavy*W] and
  // arrival count [n_heavy]; zero before a level, left zero after it
  unsigned long long* hacc;
  unsigned int* hcnt;
};

// Zero-filled per-heavy-row accumulators for msbfs_level (MsArgs::hacc /
// hcnt), W words per row, in one allocation of ms_heavy_bytes(A, W).
inline size_t ms_heavy_bytes(const gg_matrix* A, int W) {
  const int64_t nh = A->format == GG_SELL ? A->n_heavy : 0;
  return (size_t)nh * (8 * (size_t)W + 4) + 64;
}

// One or-and level over A_T with W = Wc words per vertex (1, 2, 4 or 8).
int msbfs_level(hipStream_t stream, const gg_matrix* A, const MsArgs& a, int Wc);

// One in-place min-plus step of k sources over a row-major working set d
// (state rows x k): d[v, j] = min(d[v, j], min_u safe_add(w(u->v), d[u, j]))
// over the in-neighbours u whose bit is set in chp (rows that fell in the
// previous step, or were seeded).  Rows that fall set their bit in chn and
// set *changed to 1; when changed_mask is given, bit j is set for every
// column j that fell (k <= 64).
int bf_multi_step(hipStream_t stream, const gg_matrix* A, int dist_dtype, int64_t k, void* d,
                  const uint32_t* chp, uint32_t* chn, unsigned int* changed,
                  unsigned long long* changed_mask, const int32_t* done = nullptr);

}  // namespace gg
