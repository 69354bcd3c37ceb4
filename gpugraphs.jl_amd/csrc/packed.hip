// packed.hip -- many independent graphs in one launch (gg_pack_graphs,
// gg_bfs_packed).
//
// The A_T of G graphs are packed block-diagonally into one CSR (graph g owns
// rows and columns offsets[g] .. offsets[g+1]-1, each block padded with
// isolated vertices to a multiple of 256 rows).  One BFS per graph, from its
// own source, runs in ONE cooperative launch of a persistent kernel: every
// level, resident workgroups take 256-row chunks from an atomic chunk counter
// (the work queue), skip chunks of graphs that have finished, and pull the
// chunk's unvisited rows against the frontier bitmap (src/algorithms/
// bfs.jl:50-71, one level of the or-and SpMV with its epilogue fused).  After
// a grid-wide barrier one workgroup closes the level per graph: a graph whose
// level claimed nothing is done (its level count recorded); the others go on.
// The launch ends when no graph is active, so each graph's levels equal the
// single-graph BFS from its source while small graphs stop costing anything
// after their last level.
//
// A chunk's 256 rows own 8 whole words of the next-frontier and visited
// bitmaps (the blocks are 256-aligned), so those words are written with plain
// stores; rows longer than 32 entries are scanned by a whole wavefront.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <vector>

#include "frontier.hpp"

namespace cg = cooperative_groups;

namespace gg {

constexpr int kP = 256;      // rows per chunk = threads per workgroup
constexpr int kShort = 32;   // rows up to this length are scanned by one lane

struct PackedArgs {
  const int32_t* rowptr;
  const int32_t* colval;
  const void* val;            // edge values (NULL: unit pattern)
  int64_t n_chunks;
  const int32_t* chunk_graph;  // [n_chunks] graph of each chunk
  uint32_t* bm[2];            // frontier bitmaps: level L reads bm[(L-1)&1], writes bm[L&1]
  uint32_t* visited;
  int32_t* dist;
  int32_t* found;             // [G] 1 when the graph's level claimed a row
  int32_t* done;              // [G]
  int32_t* levels;            // [G] levels of each graph's BFS
  int32_t n_graphs;
  unsigned int* chunk_ctr;    // [2] work queue of the level parity
  int32_t* active;            // graphs still active after the last level
  int32_t first_level;        // levels before it ran top-down (separate launches)
  int32_t max_level;
};

template <class TA, bool UNIT>
__device__ __forceinline__ bool packed_edge(const PackedArgs& a, const uint32_t* front, int64_t i) {
  const int32_t c = a.colval[i] - 1;
  return bit_get(front, c) && (UNIT || edge_on((const TA*)a.val, i));
}

template <class TA, bool UNIT>
__device__ void packed_chunk(const PackedArgs& a, int64_t chunk, int32_t level,
                             const uint32_t* __restrict__ front, uint32_t* __restrict__ next,
                             int32_t* s_long, int* s_nlong, int* s_hit) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t r0 = chunk * kP;
  const int64_t row = r0 + tid;
  const bool vis = bit_get(a.visited, row);
  bool hit = false, longrow = false;
  int32_t s = 0, e = 0;
  if (!vis) {
    s = a.rowptr[row] - 1;
    e = a.rowptr[row + 1] - 1;
    if (e - s <= kShort) {
      // 4 entries per step: their column loads, then their frontier tests,
      // are independent, so a lane keeps 4 of each in flight
      for (int32_t i = s; i < e && !hit; i += 4) {
        int32_t c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = i + u < e ? a.colval[i + u] - 1 : -1;
        bool h[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
          h[u] = c[u] >= 0 && bit_get(front, c[u]) &&
                 (UNIT || edge_on((const TA*)a.val, (int64_t)i + u));
        hit = h[0] || h[1] || h[2] || h[3];
      }
    } else {
      longrow = true;
    }
  }
  if (tid == 0) *s_nlong = 0;
  s_hit[tid] = 0;
  __syncthreads();
  if (longrow) s_long[atomicAdd(s_nlong, 1)] = tid;
  __syncthreads();
  // long rows: one wavefront each, 64 entries per step, stop at the first hit
  const int nl = *s_nlong;
  for (int q = wv; q < nl; q += kP / 64) {
    const int t = s_long[q];
    const int32_t rs = a.rowptr[r0 + t] - 1, re = a.rowptr[r0 + t + 1] - 1;
    for (int32_t b = rs; b < re; b += 64) {
      const int32_t i = b + lane;
      if (__ballot(i < re && packed_edge<TA, UNIT>(a, front, i))) {
        if (lane == 0) s_hit[t] = 1;
        break;
      }
    }
  }
  __syncthreads();
  if (longrow) hit = s_hit[tid] != 0;
  if (hit) a.dist[row] = level;
  // the wavefront's 64 rows own two words of each bitmap
  const unsigned long long b = __ballot(hit);
  if (lane == 0) {
    const int64_t w0 = (r0 + wv * 64) >> 5;
    next[w0] = (uint32_t)b;
    next[w0 + 1] = (uint32_t)(b >> 32);
    if (b) {
      a.visited[w0] |= (uint32_t)b;
      a.visited[w0 + 1] |= (uint32_t)(b >> 32);
    }
  }
  if (b && lane == 0) set_flag((unsigned int*)&a.found[a.chunk_graph[chunk]]);
}

template <class TA, bool UNIT>
__global__ __launch_bounds__(kP) void packed_bfs_kernel(PackedArgs a) {
  __shared__ int32_t s_long[kP];
  __shared__ int s_nlong;
  __shared__ int s_hit[kP];
  __shared__ long long s_chunk;
  __shared__ int s_any;
  cg::grid_group grid = cg::this_grid();
  for (int32_t level = a.first_level; level <= a.max_level; ++level) {
    const uint32_t* front = a.bm[(level - 1) & 1];
    uint32_t* next = a.bm[level & 1];
    unsigned int* ctr = &a.chunk_ctr[level & 1];
    while (true) {
      if (threadIdx.x == 0) s_chunk = (long long)atomicAdd(ctr, 1u);
      __syncthreads();
      const int64_t c = s_chunk;
      __syncthreads();  // s_chunk is rewritten by the next take
      if (c >= a.n_chunks) break;
      if (__hip_atomic_load(&a.done[a.chunk_graph[c]], __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT))
        continue;
      packed_chunk<TA, UNIT>(a, c, level, front, next, s_long, &s_nlong, s_hit);
      __syncthreads();  // the chunk's LDS lists are reused by the next chunk
    }
    grid.sync();
    if (blockIdx.x == 0) {
      // close the level per graph; reset the queue of the level after next
      if (threadIdx.x == 0) s_any = 0;
      __syncthreads();
      int any = 0;
      for (int g = threadIdx.x; g < a.n_graphs; g += kP) {
        if (a.done[g]) continue;  // written only here
        if (__hip_atomic_load(&a.found[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          a.found[g] = 0;
          any = 1;
        } else {
          a.done[g] = 1;
          a.levels[g] = level - 1;
        }
      }
      if (any) s_any = 1;  // benign race: every writer stores 1
      __syncthreads();
      if (threadIdx.x == 0) {
        a.chunk_ctr[(level + 1) & 1] = 0u;
        *a.active = s_any;
      }
    }
    grid.sync();
    // device-scope read: every workgroup must see the value block 0 stored
    // before the barrier, so all leave the level loop together
    if (__hip_atomic_load(a.active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
  }
}

__global__ void packed_seed_kernel(const int64_t* __restrict__ offsets, const int32_t* __restrict__ src,
                                   int32_t n_graphs, uint32_t* front, uint32_t* visited,
                                   int32_t* dist, int32_t* found, int32_t* done,
                                   const int32_t* __restrict__ rowptr, int32_t* queue0,
                                   unsigned long long* cnt) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_graphs) return;
  const int64_t v = offsets[g] + src[g] - 1;
  dist[v] = 0;
  atomicOr(&front[v >> 5], 1u << (v & 31));
  atomicOr(&visited[v >> 5], 1u << (v & 31));
  found[g] = 0;
  done[g] = 0;
  if (queue0) {  // the sources are the first queue (top-down levels)
    queue0[g] = (int32_t)v;
    atomicAdd(&cnt[1], (unsigned long long)(rowptr[v + 1] - rowptr[v]));
  }
}

// ----------------------------------------------------- top-down first levels
// A symmetric unit pattern (row u of A_T lists u's out-neighbours) starts with
// top-down levels while the frontier is small: each is an ordinary launch
// that expands the queue of the rows the previous level claimed, and the host
// reads back the new queue's length and degree sum to pick the next level's
// direction (Beamer's rule).  The persistent pull kernel then takes over from
// the first bottom-up level.
struct PackedTD {
  const int32_t* rowptr;
  const int32_t* colval;
  const int32_t* chunk_graph;
  const int32_t* cur;          // queue of the previous level
  int64_t qn;
  int32_t* nxt;                // queue this level fills
  uint32_t* next;              // next-frontier bitmap (cleared before the launch)
  uint32_t* visited;
  int32_t* dist;
  int32_t* found;
  unsigned long long* cnt;     // [0] = queue length, [1] = degree sum, [3] = pieces
  int2* pieces;                // (row, first entry) of the kTdPiece-entry pieces of long rows
};

constexpr int32_t kTdPiece = 1024;  // rows longer than this are cut into pieces

// Claims of a wavefront staged in LDS (the count and degree sum in
// wave-uniform registers): one device-wide atomic per counter per flush
// instead of one per 64-entry step (a counter that every wavefront of the 8
// XCDs bumps serialises the level).
constexpr int kTdBuf = 512;
struct TdStage {
  int32_t* buf;             // this wavefront's kTdBuf entries in LDS
  unsigned cnt;
  unsigned long long deg;
};

// Flushes when another step might not fit, or always when `last`.  All 64
// lanes call it with the same stage.
__device__ void td_flush(const PackedTD& t, TdStage& st, bool last) {
  if (st.cnt == 0 || (!last && st.cnt <= (unsigned)(kTdBuf - 64))) return;
  __builtin_amdgcn_wave_barrier();
  const int lane = threadIdx.x & 63;
  unsigned long long base = 0;
  if (lane == 0) {
    base = atomicAdd(&t.cnt[0], (unsigned long long)st.cnt);
    atomicAdd(&t.cnt[1], st.deg);
  }
  base = __shfl(base, 0, 64);
  for (unsigned i = lane; i < st.cnt; i += 64) t.nxt[base + i] = st.buf[i];
  __builtin_amdgcn_wave_barrier();
  st.cnt = 0;
  st.deg = 0;
}

// Entries [b0, b1) of row u (0-based entries) expanded by one wavefront, 64
// at a time: unvisited neighbours are claimed with an atomic OR on the
// visited bitmap (a plain read first skips the ones already set), get their
// level and next bit, and are staged with their degree.
__device__ void td_expand(const PackedTD& t, int32_t u, int32_t b0, int32_t b1, int32_t level,
                          TdStage& st) {
  const int lane = threadIdx.x & 63;
  bool any = false;
  for (int32_t b = b0; b < b1; b += 64) {
    const int32_t i = b + lane;
    bool claim = false;
    int32_t v = 0;
    unsigned long long deg = 0;
    if (i < b1) {
      v = t.colval[i] - 1;
      const uint32_t bit = 1u << (v & 31);
      if (!(t.visited[v >> 5] & bit) && !(atomicOr(&t.visited[v >> 5], bit) & bit)) {
        claim = true;
        t.dist[v] = level;
        atomicOr(&t.next[v >> 5], bit);
        deg = (unsigned long long)(t.rowptr[v + 1] - t.rowptr[v]);
      }
    }
    const unsigned long long m = __ballot(claim);
    if (m) {
      any = true;
      unsigned long long dsum = deg;
      for (int o = 32; o > 0; o >>= 1) dsum += __shfl_xor(dsum, o, 64);
      if (claim) st.buf[st.cnt + __popcll(m & ((1ull << lane) - 1ull))] = v;
      st.cnt += (unsigned)__popcll(m);
      st.deg += dsum;
      td_flush(t, st, false);
    }
  }
  // every neighbour is in u's graph (the blocks are disjoint)
  if (any && lane == 0) set_flag((unsigned int*)&t.found[t.chunk_graph[u / kP]]);
}

// One wavefront per queued row; a row longer than kTdPiece entries queues its
// pieces for packed_td_big_kernel instead.
__global__ __launch_bounds__(kP) void packed_td_kernel(PackedTD t, int32_t level) {
  __shared__ int32_t s_buf[kP / 64][kTdBuf];
  const int lane = threadIdx.x & 63;
  TdStage st{s_buf[threadIdx.x >> 6], 0u, 0ull};
  const int64_t waves = (int64_t)gridDim.x * (kP / 64);
  for (int64_t q = (int64_t)blockIdx.x * (kP / 64) + (threadIdx.x >> 6); q < t.qn; q += waves) {
    const int32_t u = t.cur[q];
    const int32_t rs = t.rowptr[u] - 1, re = t.rowptr[u + 1] - 1;
    if (re - rs > kTdPiece) {
      if (lane == 0) {
        const int32_t np = (re - rs + kTdPiece - 1) / kTdPiece;
        const unsigned long long p0 = atomicAdd(&t.cnt[3], (unsigned long long)np);
        for (int32_t j = 0; j < np; ++j) t.pieces[p0 + j] = make_int2(u, rs + j * kTdPiece);
      }
      continue;
    }
    td_expand(t, u, rs, re, level, st);
  }
  td_flush(t, st, true);
}

// One wavefront per piece of a long row.
__global__ __launch_bounds__(kP) void packed_td_big_kernel(PackedTD t, int32_t level) {
  __shared__ int32_t s_buf[kP / 64][kTdBuf];
  TdStage st{s_buf[threadIdx.x >> 6], 0u, 0ull};
  const int64_t np = (int64_t)t.cnt[3];
  const int64_t waves = (int64_t)gridDim.x * (kP / 64);
  for (int64_t q = (int64_t)blockIdx.x * (kP / 64) + (threadIdx.x >> 6); q < np; q += waves) {
    const int2 pc = t.pieces[q];
    const int32_t re = t.rowptr[pc.x + 1] - 1;
    td_expand(t, pc.x, pc.y, min(pc.y + kTdPiece, re), level, st);
  }
  td_flush(t, st, true);
}

// Closes a top-down level per graph (block 0's job in the persistent
// kernel): a graph whose level claimed nothing is done; cnt[2] counts the
// graphs still active.
__global__ void packed_close_kernel(int32_t* found, int32_t* done, int32_t* levels, int n_graphs,
                                    int32_t level, unsigned long long* cnt) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_graphs || done[g]) return;
  if (found[g]) {
    found[g] = 0;
    atomicAdd(&cnt[2], 1ull);
  } else {
    done[g] = 1;
    levels[g] = level - 1;
  }
}

__global__ void packed_td_reset_kernel(unsigned long long* cnt) {
  cnt[0] = 0ull;
  cnt[1] = 0ull;
  cnt[2] = 0ull;
  cnt[3] = 0ull;
}

__global__ void chunk_graph_kernel(const int64_t* __restrict__ offsets, int32_t n_graphs,
                                   int64_t n_chunks, int32_t* __restrict__ chunk_graph) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_chunks) return;
  const int64_t row = c * kP;
  int lo = 0, hi = n_graphs;  // offsets[lo] <= row < offsets[lo + 1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offsets[mid] <= row) lo = mid; else hi = mid;
  }
  chunk_graph[c] = lo;
}

}  // namespace gg

using namespace gg;

extern "C" {

int gg_pack_graphs(int n_graphs, const gg_csr_host* const* graphs, gg_csr_host** out,
                   int64_t* offsets) {
  if (!out || !offsets || n_graphs < 1 || !graphs)
    return set_error(GG_ERR_ARGUMENT, "need at least one graph, an output and offsets");
  clear_error();
  int64_t rows = 0, nnz = 0;
  const int dt = graphs[0] ? graphs[0]->val_dtype : 0;
  for (int g = 0; g < n_graphs; ++g) {
    const gg_csr_host* h = graphs[g];
    if (!h || h->m != h->n || h->m < 0)
      return set_error(GG_ERR_ARGUMENT, "graph %d is not a square CSR", g);
    if (h->val_dtype != dt) return set_error(GG_ERR_ARGUMENT, "graphs differ in value type");
    offsets[g] = rows;
    rows += (h->m + kP - 1) / kP * kP;
    nnz += h->nnz;
  }
  offsets[n_graphs] = rows;
  if (rows > 2147483647LL || nnz + 1 > 2147483647LL)
    return set_error(GG_ERR_ARGUMENT, "packed graph exceeds Int32 indices");
  const size_t es = dtype_size(dt);
  if (!es) return set_error(GG_ERR_ARGUMENT, "unknown value type %d", dt);
  gg_csr_host* p = new gg_csr_host();
  p->m = p->n = rows;
  p->nnz = nnz;
  p->val_dtype = dt;
  p->rowptr = (int32_t*)malloc(sizeof(int32_t) * (rows + 1));
  p->colval = (int32_t*)malloc(sizeof(int32_t) * std::max<int64_t>(1, nnz));
  p->nzval = malloc(es * std::max<int64_t>(1, nnz));
  if (!p->rowptr || !p->colval || !p->nzval) {
    gg_csr_host_free(p);
    return set_error(GG_ERR_NOMEM, "host allocation failed");
  }
  int64_t e = 0;
  p->rowptr[0] = 1;
  for (int g = 0; g < n_graphs; ++g) {
    const gg_csr_host* h = graphs[g];
    const int64_t off = offsets[g];
    for (int64_t r = 0; r < h->m; ++r) p->rowptr[off + r + 1] = (int32_t)(e + h->rowptr[r + 1]);
    for (int64_t r = h->m; r < offsets[g + 1] - off; ++r)
      p->rowptr[off + r + 1] = (int32_t)(e + h->nnz + 1);
    for (int64_t i = 0; i < h->nnz; ++i) p->colval[e + i] = (int32_t)(h->colval[i] + off);
    memcpy((char*)p->nzval + e * es, h->nzval, es * h->nnz);
    e += h->nnz;
  }
  *out = p;
  return GG_OK;
}

int gg_bfs_packed(gg_ctx* ctx, const gg_matrix* A_T, const int64_t* offsets, int n_graphs,
                  const int32_t* sources, int32_t* dist, int32_t* levels) {
  if (!ctx || !A_T || !offsets || !sources || !dist || n_graphs < 1)
    return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  GG_TRY(check_matrix(A_T));
  if (A_T->format != GG_CSR || A_T->m != A_T->n)
    return set_error(GG_ERR_ARGUMENT, "a packed graph is a square block-diagonal CSR");
  if (offsets[0] != 0 || offsets[n_graphs] != A_T->m)
    return set_error(GG_ERR_DIMENSION, "offsets do not cover the %lld rows", (long long)A_T->m);
  // the chunk queue walks whole 256-row chunks: a last block that is not
  // padded to a multiple of 256 rows would leave its trailing rows unvisited
  if (A_T->m % kP)
    return set_error(GG_ERR_ARGUMENT, "packed rows (%lld) must be a multiple of %d",
                     (long long)A_T->m, kP);
  for (int g = 0; g < n_graphs; ++g) {
    if (offsets[g + 1] < offsets[g] || (offsets[g] % kP))
      return set_error(GG_ERR_ARGUMENT, "graph offsets must rise in multiples of %d", kP);
    if (sources[g] < 1 || sources[g] > offsets[g + 1] - offsets[g])
      return set_error(GG_ERR_BOUNDS, "source %d of graph %d outside its rows", sources[g], g);
  }
  const bool unit = (A_T->flags & GG_FLAG_UNIT_PATTERN) != 0;
  if (!unit && (A_T->val_dtype == GG_FLOAT32 || A_T->val_dtype == GG_FLOAT64))
    return set_error(GG_ERR_ARGUMENT, "or-and BFS needs Bool or integer matrix values");
  const int64_t m = A_T->m, words = m / 32, n_chunks = m / kP;
  timer_begin(ctx);
  void *p_off, *p_src, *p_cg, *p_bm, *p_st;
  GG_TRY(ws_get(ctx, WS_SPMSPV0, sizeof(int64_t) * (n_graphs + 1) + sizeof(int32_t) * n_graphs,
                &p_off));
  p_src = (int64_t*)p_off + n_graphs + 1;
  GG_TRY(ws_get(ctx, WS_SPMSPV1, sizeof(int32_t) * (n_chunks + 1), &p_cg));
  GG_TRY(ws_get(ctx, WS_BFS0, sizeof(uint32_t) * 3 * (words + 1), &p_bm));
  // per graph: found, done, levels; then the two queue counters and the active
  // flag; then, 8-byte aligned, the top-down counters (queue length, degree
  // sum, active graphs)
  const size_t st_i32 = (sizeof(int32_t) * (3 * (size_t)n_graphs + 8) + 7) & ~(size_t)7;
  GG_TRY(ws_get(ctx, WS_SPMSPV2, st_i32 + 4 * sizeof(uint64_t), &p_st));  // td_cnt[0..3]
  unsigned long long* td_cnt = (unsigned long long*)((char*)p_st + st_i32);
  // top-down first levels need row u of A_T to list u's out-neighbours
  const bool topdown = unit && (A_T->flags & GG_FLAG_SYMMETRIC) && ctx->push_alpha > 0.0;
  void *p_q0 = nullptr, *p_q1 = nullptr, *p_pc = nullptr;
  if (topdown) {
    GG_TRY(ws_get(ctx, WS_BFS3, sizeof(int32_t) * m, &p_q0));
    GG_TRY(ws_get(ctx, WS_BFS4, sizeof(int32_t) * m, &p_q1));
    // pieces of rows longer than kTdPiece: fewer than nnz / kTdPiece + one per such row
    GG_TRY(ws_get(ctx, WS_BFS5, sizeof(int2) * (size_t)(2 * (A_T->nnz / kTdPiece) + 16), &p_pc));
  }
  GG_HIP(hipMemcpyAsync(p_off, offsets, sizeof(int64_t) * (n_graphs + 1), hipMemcpyHostToDevice,
                        ctx->stream));
  GG_HIP(hipMemcpyAsync(p_src, sources, sizeof(int32_t) * n_graphs, hipMemcpyHostToDevice,
                        ctx->stream));
  PackedArgs a{};
  a.rowptr = A_T->ptr;
  a.colval = A_T->idx;
  a.val = unit ? nullptr : A_T->val;
  a.n_chunks = n_chunks;
  a.chunk_graph = (const int32_t*)p_cg;
  a.bm[0] = (uint32_t*)p_bm;
  a.bm[1] = a.bm[0] + words + 1;
  a.visited = a.bm[1] + words + 1;
  a.dist = dist;
  a.found = (int32_t*)p_st;
  a.done = a.found + n_graphs;
  a.levels = a.done + n_graphs;
  a.n_graphs = n_graphs;
  a.chunk_ctr = (unsigned int*)(a.levels + n_graphs);
  a.active = (int32_t*)(a.chunk_ctr + 2);
  a.first_level = 1;
  a.max_level = (int32_t)std::min<int64_t>(m + 1, 2147483646LL);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(3 * (words + 1), kP)), dim3(kP), 0,
                     ctx->stream, a.bm[0], 3 * (words + 1), 0u);
  hipLaunchKernelGGL(fill_i32v_kernel, dim3(blocks_for(m, kP)), dim3(kP), 0, ctx->stream, dist, m,
                     (int32_t)2147483647);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, ctx->stream, a.chunk_ctr, 3, 0u);
  hipLaunchKernelGGL(fill_i32v_kernel, dim3(blocks_for(n_graphs, kP)), dim3(kP), 0, ctx->stream,
                     a.levels, n_graphs, 0);
  hipLaunchKernelGGL(chunk_graph_kernel, dim3(blocks_for(n_chunks, kP)), dim3(kP), 0, ctx->stream,
                     (const int64_t*)p_off, n_graphs, n_chunks, (int32_t*)p_cg);
  hipLaunchKernelGGL(packed_td_reset_kernel, dim3(1), dim3(1), 0, ctx->stream, td_cnt);
  hipLaunchKernelGGL(packed_seed_kernel, dim3(blocks_for(n_graphs, kP)), dim3(kP), 0, ctx->stream,
                     (const int64_t*)p_off, (const int32_t*)p_src, n_graphs, a.bm[0], a.visited,
                     dist, a.found, a.done, A_T->ptr, (int32_t*)p_q0, td_cnt);
  GG_HIP(hipGetLastError());
  bool all_done = false;
  if (topdown) {
    // top-down levels while the queue's degree sum * alpha is below the
    // entries of the rows not reached yet; one host round trip per level
    GG_HIP(hipMemcpyAsync(ctx->host_pinned, td_cnt, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    GG_HIP(hipStreamSynchronize(ctx->stream));
    const unsigned long long* hc = (const unsigned long long*)ctx->host_pinned;
    int64_t qn = n_graphs;
    double qdeg = (double)hc[1];
    double left = (double)A_T->nnz - qdeg;
    int32_t level = 1;
    int32_t* q[2] = {(int32_t*)p_q0, (int32_t*)p_q1};
    while (qn > 0 && qdeg * ctx->push_alpha < left) {
      PackedTD t{};
      t.rowptr = A_T->ptr;
      t.colval = A_T->idx;
      t.chunk_graph = a.chunk_graph;
      t.cur = q[(level - 1) & 1];
      t.qn = qn;
      t.nxt = q[level & 1];
      t.next = a.bm[level & 1];
      t.visited = a.visited;
      t.dist = dist;
      t.found = a.found;
      t.cnt = td_cnt;
      t.pieces = (int2*)p_pc;
      hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(words + 1, kP)), dim3(kP), 0,
                         ctx->stream, t.next, words + 1, 0u);
      hipLaunchKernelGGL(packed_td_reset_kernel, dim3(1), dim3(1), 0, ctx->stream, td_cnt);
      const int64_t nb = std::min<int64_t>((qn + 3) / 4, (int64_t)ctx->num_cus * 16);
      hipLaunchKernelGGL(packed_td_kernel, dim3((unsigned)std::max<int64_t>(nb, 1)), dim3(kP), 0,
                         ctx->stream, t, level);
      hipLaunchKernelGGL(packed_td_big_kernel, dim3((unsigned)ctx->num_cus * 16), dim3(kP), 0,
                         ctx->stream, t, level);
      hipLaunchKernelGGL(packed_close_kernel, dim3(blocks_for(n_graphs, kP)), dim3(kP), 0,
                         ctx->stream, a.found, a.done, a.levels, n_graphs, level, td_cnt);
      GG_HIP(hipGetLastError());
      GG_HIP(hipMemcpyAsync(ctx->host_pinned, td_cnt, 3 * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, ctx->stream));
      GG_HIP(hipStreamSynchronize(ctx->stream));
      qn = (int64_t)hc[0];
      qdeg = (double)hc[1];
      left -= qdeg;
      ++level;
      if (hc[2] == 0) {  // every graph finished top-down
        all_done = true;
        break;
      }
    }
    a.first_level = level;
  }
  auto launch = [&](auto kern) -> int {
    int per_cu = 0;
    GG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kP, 0));
    if (per_cu < 1) return set_error(GG_ERR_HIP, "packed BFS kernel cannot be resident");
    int64_t grid = (int64_t)per_cu * ctx->num_cus;
    grid = std::max<int64_t>(1, std::min<int64_t>(grid, n_chunks));
    void* args[] = {&a};
    HotTimer ht(ctx);
    GG_HIP(hipLaunchCooperativeKernel((const void*)kern, dim3((unsigned)grid), dim3(kP), args, 0,
                                      ctx->stream));
    return GG_OK;
  };
  int st;
  if (all_done) {
    st = GG_OK;
  } else if (unit) {
    st = launch(packed_bfs_kernel<uint8_t, true>);
  } else {
    st = with_type(A_T->val_dtype, [&](auto t) -> int {
      typedef typename decltype(t)::type TA;
      if constexpr (std::is_floating_point<TA>::value) return GG_ERR_ARGUMENT;
      else return launch(packed_bfs_kernel<TA, false>);
    });
  }
  GG_TRY(st);
  if (levels)
    GG_HIP(hipMemcpyAsync(levels, a.levels, sizeof(int32_t) * n_graphs, hipMemcpyDeviceToHost,
                          ctx->stream));
  GG_HIP(hipStreamSynchronize(ctx->stream));
  timer_end(ctx);
  return GG_OK;
}

}  // extern "C"
