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
  int32_t* dist;               // original rows; NULL with dist8: no per-level distance writes
                               // (the caller keeps each level's next and records them later)
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
  // when set and nonzero, the pull kernels do nothing (a push step ran)
  const int32_t* pull_off;
  // when set and nonzero, the pull kernels read front without testing
  // front_any first: every front word was written by the previous level (a
  // pull level) and the frontier is dense, so the test filters little and
  // costs a dependent gather per entry
  const int32_t* dense;
};

// Buffers of the top-down (push) step of a multi-source BFS level: used on a
// symmetric unit-pattern A_T in state order (CSR, or a relabelled SELL).  A
// device-side switch picks push when the frontier's edges are few; every
// kernel of the other direction returns at once, so a level is the same
// sequence of launches either way (and can be captured in a graph).
struct MsPush {
  unsigned long long* acc;  // [ns*W] OR of pushed masks, zero outside a push step
  int32_t* queue;           // [ns] frontier state ids
  int32_t* tqueue;          // [ns] rows touched by the push
  uint32_t* touched;        // [words] bitmap of tqueue, zero outside a push step
  int2* big;                // hub pieces (state id, first entry)
  unsigned int* cnt;        // [0] qlen, [1] tlen, [2] nbig, [3] mode (1 push), [4] = [3],
                            // [5] dense pull (MsArgs::dense), [6] every word of this
                            // level's next will be written (a pull, or a wide push)
  unsigned long long* scout;  // a level's first kernel: (blocks << 40) + frontier out-edges; else 0
  int64_t words, nnz, big_cap;
  double alpha;             // push when frontier out-edges * alpha < nnz
};

inline bool ms_push_ok(const gg_matrix* A) {
  const int64_t f = A->flags;
  return (f & GG_FLAG_SYMMETRIC) && (f & GG_FLAG_UNIT_PATTERN) &&
         (A->format == GG_CSR || (A->format == GG_SELL && (f & GG_FLAG_INTERNAL_IDS)));
}
// Out-edges from which a frontier row of a top-down step is cut into pieces
// of this many edges (each piece one block's work).
constexpr int kPushBig = 1024;
inline size_t ms_push_bytes(const gg_matrix* A, int W) {
  const int64_t ns = state_size(A), words = (ns + 31) / 32;
  const int64_t big = 2 * (A->nnz / kPushBig) + 16;
  return 8 * (size_t)ns * W + 8 * (size_t)ns + 4 * (size_t)(words + 8) + 8 * (size_t)big + 64;
}
// Lays out an MsPush over one zero-filled allocation of ms_push_bytes(A, W).
inline MsPush ms_push_layout(void* base, const gg_matrix* A, int W) {
  MsPush p{};
  const int64_t ns = state_size(A);
  p.words = (ns + 31) / 32;
  p.nnz = A->nnz;
  p.big_cap = 2 * (A->nnz / kPushBig) + 16;
  char* c = (char*)base;
  p.acc = (unsigned long long*)c; c += 8 * (size_t)ns * W;
  p.big = (int2*)c; c += 8 * (size_t)p.big_cap;
  p.scout = (unsigned long long*)c; c += 8;
  p.cnt = (unsigned int*)c; c += 4 * 8;
  p.queue = (int32_t*)c; c += 4 * (size_t)ns;
  p.tqueue = (int32_t*)c; c += 4 * (size_t)ns;
  p.touched = (uint32_t*)c;
  p.alpha = 16.0;
  return p;
}

// Zero-filled per-heavy-row accumulators for msbfs_level (MsArgs::hacc /
// hcnt), W words per row, in one allocation of ms_heavy_bytes(A, W).
inline size_t ms_heavy_bytes(const gg_matrix* A, int W) {
  const int64_t nh = A->format == GG_SELL ? A->n_heavy : 0;
  return (size_t)nh * (8 * (size_t)W + 4) + 64;
}

// One or-and level over A_T with W = Wc words per vertex (1, 2, 4 or 8).
// The level clears a.next_any first.  With push (ms_push_ok(A)), it decides
// on the device between the pull kernel and the top-down step; the push
// buffers' counters (scout, cnt) must be zero before a traversal's first
// level (each level leaves them so); `ctx_cus` sizes the push grids.
int msbfs_level(hipStream_t stream, const gg_matrix* A, const MsArgs& a, int Wc,
                const MsPush* push = nullptr, int ctx_cus = 256);

// One in-place min-plus step of k sources over a row-major working set d
// (state rows x k): d[v, j] = min(d[v, j], min_u safe_add(w(u->v), d[u, j]))
// over the in-neighbours u whose bit is set in chp (rows that fell in the
// previous step, or were seeded).  Rows that fall set their bit in chn and
// set *changed to 1; when changed_mask is given, bit j is set for every
// column j that fell (k <= 64).
int bf_multi_step(hipStream_t stream, const gg_matrix* A, int dist_dtype, int64_t k, void* d,
                  const uint32_t* chp, uint32_t* chn, unsigned int* changed,
                  unsigned long long* changed_mask, const int32_t* done = nullptr);

// Sparse k-source iterations on the symmetric pattern of a relabelled SELL
// (bf_sparse_ok): the buffers of the device-side switch and of the rows to
// relax, in one allocation of bf_sparse_bytes(A) whose mark bitmap is zero.
struct BfSparse {
  int32_t* queue;           // [ns] rows that fell in the previous iteration
  int32_t* dq;              // [ns] light rows to relax
  uint32_t* mark;           // [words] bitmap of dq, zero outside a sparse iteration
  int2* big;                // pieces of long fallen rows
  unsigned int* cnt;        // [0] fallen, [1] marked, [2] pieces, [3] sparse, [4] = [3]
  unsigned long long* scout;  // out-edges of the fallen rows
  int64_t words, ns, big_cap, nnz;
  double alpha;
};
constexpr int kSparseBig = 4096;

inline bool bf_sparse_ok(const gg_matrix* A) {
  return A->format == GG_SELL && (A->flags & GG_FLAG_INTERNAL_IDS) &&
         (A->flags & GG_FLAG_SYMMETRIC);
}
inline size_t bf_sparse_bytes(const gg_matrix* A) {
  const int64_t ns = state_size(A), words = (ns + 31) / 32, big = 2 * (A->nnz / kSparseBig) + 16;
  return 8 * (size_t)ns + 4 * (size_t)(words + 16) + 8 * (size_t)big + 64;
}
inline BfSparse bf_sparse_layout(void* base, const gg_matrix* A, double alpha) {
  BfSparse s{};
  const int64_t ns = state_size(A);
  s.words = (ns + 31) / 32;
  s.ns = ns;
  s.big_cap = 2 * (A->nnz / kSparseBig) + 16;
  s.nnz = A->nnz;
  s.alpha = alpha;
  char* c = (char*)base;
  s.big = (int2*)c; c += 8 * (size_t)s.big_cap;
  s.scout = (unsigned long long*)c; c += 8;
  s.cnt = (unsigned int*)c; c += 4 * 8;
  s.queue = (int32_t*)c; c += 4 * (size_t)ns;
  s.dq = (int32_t*)c; c += 4 * (size_t)ns;
  s.mark = (uint32_t*)c;
  return s;
}

// Column groups a k-source step relaxes: the step's group i is columns
// g[i] * V .. g[i] * V + V - 1, V = 16 / sizeof(distance) (the 16-byte
// column kernels); n = 0: every group of the first k columns.  The serving
// engine lists the groups that hold a busy slot, so a step costs what its
// busy columns cost wherever finished queries left their slots free.
struct ColGroups {
  int32_t n;
  uint8_t g[32];
};

// bf_multi_step with sparse iterations when sp is set (cus sizes their grids).
// Columns 0..k-1 of rows ld elements apart (ld < 0: ld = k), or only the
// groups of cols when given; the other columns are left alone.
int bf_multi_step_sp(hipStream_t stream, const gg_matrix* A, int dist_dtype, int64_t k, void* d,
                     const uint32_t* chp, uint32_t* chn, unsigned int* changed,
                     unsigned long long* changed_mask, const int32_t* done, const BfSparse* sp,
                     int cus, int64_t ld = -1, const ColGroups* cols = nullptr);

}  // namespace gg
