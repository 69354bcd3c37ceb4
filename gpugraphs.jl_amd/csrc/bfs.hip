// bfs.hip -- BFS levels and parents over A_T (gg_bfs_distances,
// gg_bfs_parents, gg_bfs_distances_do).
//
// One level of src/algorithms/bfs.jl:50-71 is a masked or-and SpMV
// next = A_T (bor.band) curr over the unvisited rows, followed by four
// broadcasts (dist, to_explore, curr, next) and an OR-reduce.  Here a level is
// one kernel: every unvisited row scans its in-neighbours against the frontier
// bitmap, stops at the first hit (true absorbs under bor, src/spmv/CSR.jl:49-51),
// and on a hit writes dist = level, sets its bit in the next frontier and in
// the visited set, and is counted.  The frontier (n bits, 512 KiB at 4M
// vertices) stays in L2 while the rows stream from HBM.  A BFS level is
// unique, so the levels equal the reference's whatever the traversal order.
//
// Parents follow src/algorithms/bfs.jl:135-183 with GPUGraphs_any resolved as
// the reference resolves it: CSR takes the first frontier in-neighbour in row
// order (any_dense_masked_csr_spmv_kernel!, src/spmv/CSR.jl:159-170), SELL the
// largest (GPUGraphs_any = max, src/algebra.jl:35, over the whole row).
//
// gg_bfs_distances_do switches between this pull step and a push step over
// A's out-edges (a frontier queue expanded with claim-by-atomic-OR) with the
// direction-optimising rule of Beamer et al. (SC'12): go bottom-up when the
// frontier's out-edges exceed the unexplored edges / alpha, return top-down
// when the frontier shrinks below n / beta.
#include <mutex>

#include "frontier.hpp"

namespace gg {

constexpr int kB = 256;
constexpr int kPullBatch = 16;  // entries per lane per batch of the SELL pull level

__global__ void fill_u32_kernel(uint32_t* p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_i32v_kernel(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

struct BfsCounters {
  unsigned int found;        // pull: 1 when the level claimed a vertex; push: claimed count
  unsigned int qlen;         // push: length of the next queue
  unsigned long long scout;  // push: out-degree sum of the claimed vertices
  unsigned int awake;        // set bits of a bitmap (popcount_kernel)
  unsigned int nbig;         // push: hub pieces queued for push_big_kernel
  int32_t level;             // level the next pull step assigns
  int32_t done;              // set by level_end_kernel after a level that found nothing
  int32_t* out;              // dist or parents
};

// out at the original source s; the bitmaps at its state index.
__global__ void bfs_seed_kernel(int32_t* out, int32_t out_val, uint32_t* visited,
                                uint32_t* frontier, int32_t* queue, int64_t s,
                                const int32_t* vinv) {
  const int64_t r = vinv ? (int64_t)vinv[s] - 1 : s;
  out[s] = out_val;
  visited[r >> 5] |= 1u << (r & 31);
  frontier[r >> 5] |= 1u << (r & 31);
  if (queue) queue[0] = (int32_t)s;
}

// Number of set bits of a bitmap, one atomic per block.
__global__ __launch_bounds__(kB) void popcount_kernel(const uint32_t* __restrict__ bm,
                                                      int64_t words, unsigned int* out) {
  __shared__ unsigned int part[kB / 64];
  unsigned int c = 0;
  for (int64_t w = (int64_t)blockIdx.x * kB + threadIdx.x; w < words; w += (int64_t)gridDim.x * kB)
    c += __popc(bm[w]);
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kB / 64; ++i) c += part[i];
    if (c) atomicAdd(out, c);
  }
}

// ------------------------------------------------------------ pull, SELL
// Light slots: one lane per row.  Heavy rows: one block per segment for
// levels (any hit settles the row), one block per row for parents (max over
// the row).
template <class TA, bool UNIT, bool PARENTS>
__global__ __launch_bounds__(kB) void bfs_pull_sell_kernel(
    SellView g, int64_t n_sell_blocks, const uint32_t* __restrict__ frontier,
    uint32_t* __restrict__ next, uint32_t* __restrict__ visited, BfsCounters* cnt) {
  if (cnt->done) return;  // a captured step after the last level
  const int32_t level = cnt->level;
  int32_t* __restrict__ out = cnt->out;
  if ((int64_t)blockIdx.x < n_sell_blocks) {
    bool newly = false;
    int64_t row, base, len;
    bool has_len;
    const int64_t slot = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (sell_slot(g, slot, row, base, len, has_len)) {
      // state index of the row; with internal ids the wavefront's 64 slots own
      // two whole bitmap words, set below with one OR each
      const int64_t rid = g.internal ? slot : row;
      if (!bit_get(visited, rid)) {
        const TA* val = (const TA*)g.val;
        int32_t hit = 0;
        int64_t k = 0;
        if (!PARENTS && has_len) {  // batches of kPullBatch independent loads, first hit wins
          for (; k + kPullBatch <= len && !hit; k += kPullBatch) {
            int32_t c[kPullBatch];
#pragma unroll
            for (int u = 0; u < kPullBatch; ++u) c[u] = g.colval[base + (k + u) * g.C];
#pragma unroll
            for (int u = 0; u < kPullBatch; ++u)
              if (!hit && bit_get(frontier, c[u] - 1) &&
                  (UNIT || edge_on(val, base + (k + u) * g.C)))
                hit = c[u];
          }
        }
        for (; k < len && !(hit && !PARENTS); ++k) {
          const int64_t i = base + k * g.C;
          const int32_t col = g.colval[i];
          if (!has_len && col == -1) break;
          if (bit_get(frontier, col - 1) && (PARENTS || UNIT || edge_on(val, i))) {
            // parents: GPUGraphs_any keeps the largest original column
            hit = PARENTS ? max(hit, (int32_t)orig_of(g, col - 1) + 1) : col;
          }
        }
        if (hit) {
          out[row] = PARENTS ? hit : level;
          if (g.internal) {
            newly = true;  // only this lane sets this bit
          } else {
            bit_claim(next, rid);
            newly = bit_claim(visited, rid);
          }
        }
      }
    }
    if (g.internal) {
      wave_set_bits(next, slot & ~(int64_t)63, newly);
      wave_set_bits(visited, slot & ~(int64_t)63, newly);
    }
    wave_flag(&cnt->found, newly);
    return;
  }
  __shared__ int s_skip;
  __shared__ int s_hit;
  __shared__ int s_red[kB / 64];
  const int64_t hb = (int64_t)blockIdx.x - n_sell_blocks;
  int64_t h, s, e;
  if (PARENTS) {
    h = hb;
    s = g.heavy_ptr[h];
    e = g.heavy_ptr[h + 1];
  } else {
    int64_t lo = 0, hi = g.n_heavy;  // heavy_seg[lo] <= hb < heavy_seg[lo+1]
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (g.heavy_seg[mid] <= hb) lo = mid; else hi = mid;
    }
    h = lo;
    s = g.heavy_ptr[h] + (hb - g.heavy_seg[h]) * g.seg_len;
    e = min(s + g.seg_len, g.heavy_ptr[h + 1]);
  }
  const int64_t row = g.heavy_row[h] - 1;
  const int64_t rid = heavy_rid(g, h);
  if (threadIdx.x == 0) {
    s_skip = bit_get(visited, rid);
    s_hit = 0;
  }
  __syncthreads();
  if (s_skip) return;
  const TA* hv = (const TA*)g.heavy_val;
  int32_t best = 0;
  for (int64_t i = s + threadIdx.x; i < e; i += kB) {
    if (!PARENTS && __atomic_load_n(&s_hit, __ATOMIC_RELAXED)) break;
    const int32_t col = g.heavy_col[i];
    if (bit_get(frontier, col - 1) && (PARENTS || UNIT || edge_on(hv, i))) {
      if (!PARENTS) {
        __atomic_store_n(&s_hit, 1, __ATOMIC_RELAXED);
        break;
      }
      best = max(best, (int32_t)orig_of(g, col - 1) + 1);
    }
  }
  if (PARENTS) {
    for (int o = 32; o > 0; o >>= 1) best = max(best, __shfl_xor(best, o, 64));
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = best;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t hit = 0;
    if (PARENTS) {
      for (int w = 0; w < kB / 64; ++w) hit = max(hit, s_red[w]);
    } else {
      hit = s_hit;
    }
    if (hit) {
      out[row] = PARENTS ? hit : level;
      bit_claim(next, rid);
      if (bit_claim(visited, rid)) set_flag(&cnt->found);
    }
  }
}

// ------------------------------------------------------------- pull, CSR
// Groups of L lanes per row, the whole wavefront iterating together so the
// ballots see every lane.  The group's first hit in row order is the parent.
template <class TA, bool UNIT, bool PARENTS>
__global__ __launch_bounds__(kB) void bfs_pull_csr_kernel(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colval,
    const TA* __restrict__ val, int64_t m, int logL, const uint32_t* __restrict__ frontier,
    uint32_t* __restrict__ next, uint32_t* __restrict__ visited, BfsCounters* cnt) {
  if (cnt->done) return;
  const int32_t level = cnt->level;
  int32_t* __restrict__ out = cnt->out;
  const int L = 1 << logL;
  const int wl = threadIdx.x & 63;
  const int lane = wl & (L - 1);
  const int gbase = wl & ~(L - 1);
  const unsigned long long gmask = L == 64 ? ~0ull : (((1ull << L) - 1) << gbase);
  const int64_t groups_per_wave = 64 >> logL;
  const int64_t wave = ((int64_t)blockIdx.x * kB + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kB) >> 6;
  for (int64_t w0 = wave * groups_per_wave; w0 < m; w0 += nwaves * groups_per_wave) {
    const int64_t row = w0 + (wl >> logL);
    bool newly = false;
    const bool live = row < m && !bit_get(visited, row);
    int32_t s = 0, e = 0;
    if (live) {
      s = rowptr[row] - 1;
      e = rowptr[row + 1] - 1;
    }
    int32_t hitpos = -1;
    for (int32_t b = s; b < e; b += L) {
      const int32_t i = b + lane;
      bool h = false;
      if (i < e) {
        const int32_t col = colval[i];
        h = bit_get(frontier, col - 1) && (PARENTS || UNIT || edge_on(val, i));
      }
      const unsigned long long bal = __ballot(h) & gmask;
      if (bal) {
        hitpos = b + (__ffsll(bal) - 1 - gbase);
        break;
      }
    }
    if (hitpos >= 0 && lane == 0) {
      out[row] = PARENTS ? colval[hitpos] : level;
      bit_claim(next, row);
      newly = bit_claim(visited, row);
    }
    wave_flag(&cnt->found, newly);
  }
}

// -------------------------------------------------------------- push, CSR
// Frontier queue over A_out (row u = out-neighbours of u), expanded edge by
// edge so a hub's edges spread over many lanes: a block takes `tile` (at most
// 256) queue entries, scans their out-degrees in LDS and walks the tile's
// edges 256 at a time (each lane finds its vertex by binary search in the
// scan); a short queue gets small tiles, so that its rows spread over many
// blocks.  A vertex with more than kBigDeg out-edges is cut into
// kBigDeg-edge pieces that a second kernel spreads over blocks.  Claimed
// vertices go to the next queue (one atomic per wavefront) and to the next
// bitmap.
constexpr int kBigDeg = 1024;

struct PushArgs {
  const int32_t* rowptr;
  const int32_t* colval;
  const int32_t* queue;
  int64_t qlen;
  int tile;             // queue entries per block step (1..256)
  int32_t* next_queue;
  uint32_t* next;
  uint32_t* visited;
  int32_t* dist;
  int32_t level;
  const int32_t* vinv;  // original -> internal id + 1 (bitmaps in state order), or NULL
  BfsCounters* cnt;
  int2* big;            // pieces (vertex, first edge) of the hubs in the queue
};

// Claimed vertices staged per block in LDS with their out-degree sum: one
// device-wide atomic per counter per flush instead of one per wavefront (a
// counter that every wavefront of the 8 XCDs bumps serialises the level).
constexpr int kPushBuf = 2048;
struct PushOut {
  int32_t buf[kPushBuf];
  unsigned cnt;
  unsigned base;
  unsigned long long deg;
};

// Edge u -> v (0-based v) when valid.  All 64 lanes must call it.
__device__ __forceinline__ void push_edge(const PushArgs& a, bool valid, int32_t v, PushOut& o) {
  bool newly = false;
  if (valid) {
    const int64_t r = a.vinv ? (int64_t)a.vinv[v] - 1 : v;
    if (!bit_get(a.visited, r) && bit_claim(a.visited, r)) {
      newly = true;
      a.dist[v] = a.level;
      bit_claim(a.next, r);
    }
  }
  const unsigned long long b = __ballot(newly);
  if (b) {
    const int wl = threadIdx.x & 63;
    const int leader = __ffsll(b) - 1;
    unsigned long long deg = newly ? (unsigned long long)(a.rowptr[v + 1] - a.rowptr[v]) : 0ull;
    for (int o2 = 32; o2 > 0; o2 >>= 1) deg += __shfl_xor(deg, o2, 64);
    unsigned base = 0;
    if (wl == leader) {
      base = atomicAdd(&o.cnt, (unsigned)__popcll(b));
      atomicAdd(&o.deg, deg);
    }
    base = __shfl(base, leader, 64);
    if (newly) o.buf[base + __popcll(b & ((1ull << wl) - 1))] = v;
  }
}

// Every thread of the block calls it between rounds (a round stages at most
// kB vertices): flushes when another round might not fit, or always when
// `last`.
__device__ void push_flush(const PushArgs& a, PushOut& o, bool last) {
  __syncthreads();
  const unsigned n = o.cnt;
  const bool go = n > 0 && (last || n > (unsigned)(kPushBuf - kB));
  if (go) {
    if (threadIdx.x == 0) {
      o.base = atomicAdd(&a.cnt->qlen, n);
      atomicAdd(&a.cnt->scout, o.deg);
    }
    __syncthreads();
    const unsigned base = o.base;
    for (unsigned i = threadIdx.x; i < n; i += kB) a.next_queue[base + i] = o.buf[i];
    __syncthreads();
    if (threadIdx.x == 0) {
      o.cnt = 0;
      o.deg = 0;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kB) void push_tile_kernel(PushArgs a) {
  __shared__ int32_t s_beg[kB];
  __shared__ int32_t s_scan[kB + 1];
  __shared__ int32_t s_wsum[kB / 64];
  __shared__ PushOut o;
  const int wl = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if ((int64_t)blockIdx.x * a.tile >= a.qlen) return;
  if (threadIdx.x == 0) {
    o.cnt = 0;
    o.deg = 0;
  }
  for (int64_t t0 = (int64_t)blockIdx.x * a.tile; t0 < a.qlen;
       t0 += (int64_t)gridDim.x * a.tile) {
    const int64_t qi = t0 + threadIdx.x;
    int32_t beg = 0, deg = 0;
    if ((int)threadIdx.x < a.tile && qi < a.qlen) {
      const int32_t u = a.queue[qi];
      beg = a.rowptr[u] - 1;
      deg = a.rowptr[u + 1] - 1 - beg;
      if (deg > kBigDeg) {  // hub: pieces for push_big_kernel
        const int np = (deg + kBigDeg - 1) / kBigDeg;
        const unsigned p0 = atomicAdd(&a.cnt->nbig, (unsigned)np);
        for (int p = 0; p < np; ++p) a.big[p0 + p] = make_int2(u, beg + p * kBigDeg);
        deg = 0;
      }
    }
    // exclusive scan of the tile's degrees: wavefront scans, then wavefront sums
    int32_t incl = deg;
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(incl, o, 64);
      if (wl >= o) incl += y;
    }
    if (wl == 63) s_wsum[wv] = incl;
    __syncthreads();
    int32_t off = 0;
    for (int w = 0; w < wv; ++w) off += s_wsum[w];
    s_beg[threadIdx.x] = beg;
    s_scan[threadIdx.x] = off + incl - deg;
    if (threadIdx.x == kB - 1) s_scan[kB] = off + incl;
    __syncthreads();
    const int32_t total = s_scan[kB];
    for (int32_t e0 = 0; e0 < total; e0 += kB) {
      const int32_t e = e0 + threadIdx.x;
      const bool valid = e < total;
      int32_t v = 0;
      if (valid) {
        int lo = 0, hi = kB;  // s_scan[lo] <= e < s_scan[hi]; lo is never a 0-degree entry
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_scan[mid] <= e) lo = mid; else hi = mid;
        }
        v = a.colval[s_beg[lo] + (e - s_scan[lo])] - 1;
      }
      push_edge(a, valid, v, o);
      push_flush(a, o, false);
    }
    __syncthreads();  // the next tile reuses the LDS arrays
  }
  push_flush(a, o, true);
}

__global__ __launch_bounds__(kB) void push_big_kernel(PushArgs a) {
  __shared__ PushOut o;
  const int64_t nbig = __atomic_load_n(&a.cnt->nbig, __ATOMIC_RELAXED);  // from push_tile_kernel
  if (blockIdx.x >= nbig) return;
  if (threadIdx.x == 0) {
    o.cnt = 0;
    o.deg = 0;
  }
  for (int64_t p = blockIdx.x; p < nbig; p += gridDim.x) {
    const int2 pc = a.big[p];
    const int32_t stop = min(pc.y + kBigDeg, a.rowptr[pc.x + 1] - 1);
    for (int32_t i0 = pc.y; i0 < stop; i0 += kB) {
      const int32_t i = i0 + (int32_t)threadIdx.x;
      const bool valid = i < stop;
      push_edge(a, valid, valid ? a.colval[i] - 1 : 0, o);
      push_flush(a, o, false);
    }
  }
  push_flush(a, o, true);
}

// Bitmap -> queue of the set bits' original vertices (order not required).
__global__ __launch_bounds__(kB) void bitmap_to_queue_kernel(const uint32_t* __restrict__ bm,
                                                             int64_t words, int64_t n, SellView g,
                                                             int32_t* __restrict__ queue,
                                                             BfsCounters* cnt) {
  const int64_t w = (int64_t)blockIdx.x * kB + threadIdx.x;
  uint32_t bits = w < words ? bm[w] : 0u;
  if (w == words - 1 && (n & 31)) bits &= (1u << (n & 31)) - 1u;
  unsigned pos = block_claim((unsigned)__popc(bits), &cnt->qlen);  // one atomic per block
  while (bits) {
    const int b = __ffs(bits) - 1;
    bits &= bits - 1;
    queue[pos++] = (int32_t)orig_of(g, w * 32 + b);
  }
}

__global__ void level_begin_kernel(BfsCounters* c, int32_t level, int32_t* out) {
  c->found = 0;
  c->qlen = 0;
  c->scout = 0;
  c->awake = 0;
  c->nbig = 0;
  c->level = level;
  c->done = 0;
  c->out = out;
}

// After a pull step: done when it found nothing, else advance the level.
__global__ void level_end_kernel(BfsCounters* c) {
  if (c->done) return;
  if (c->found == 0) {
    c->done = 1;
  } else {
    c->level += 1;
    c->found = 0;
  }
}

__global__ void degree_kernel(const int32_t* rowptr, int64_t v, BfsCounters* c) {
  c->scout = (unsigned long long)(rowptr[v + 1] - rowptr[v]);
}

// Many sources at level 0 (original 0-based ids, distinct): out = 0, the
// bitmaps at their state index, the first queue and its out-degree sum.
__global__ __launch_bounds__(kB) void bfs_seeds_kernel(int32_t* out, uint32_t* visited,
                                                       uint32_t* frontier, int32_t* queue,
                                                       const int32_t* __restrict__ seeds,
                                                       int64_t n_seeds, const int32_t* vinv,
                                                       const int32_t* __restrict__ rowptr_out,
                                                       BfsCounters* c) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= n_seeds) return;
  const int64_t s = seeds[i];
  const int64_t r = vinv ? (int64_t)vinv[s] - 1 : s;
  out[s] = 0;
  atomicOr(&visited[r >> 5], 1u << (r & 31));
  atomicOr(&frontier[r >> 5], 1u << (r & 31));
  queue[i] = (int32_t)s;
  atomicAdd(&c->scout, (unsigned long long)(rowptr_out[s + 1] - rowptr_out[s]));
}

// Level count of each packed graph = the largest finite level among its rows
// (offsets: n_graphs + 1 rising row offsets).  A wavefront walks kLvRows
// consecutive rows 64 at a time: lane 0 finds the graph of the first row by
// binary search, each lane steps from the chunk's first graph across the
// boundaries below its row (rarely more than one), and the wavefront keeps a
// running maximum while its rows stay in one graph, folding it into
// levels[g] with one atomic max per graph it meets.
constexpr int kLvRows = 4096;
__global__ __launch_bounds__(kB) void graph_levels_kernel(const int32_t* __restrict__ dist, int64_t m,
                                                          const int64_t* __restrict__ offsets,
                                                          int n_graphs, int32_t* levels) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * kB + threadIdx.x) >> 6;
  const int64_t r0 = wave * kLvRows;
  if (r0 >= m) return;
  const int64_t r1 = r0 + kLvRows < m ? r0 + kLvRows : m;
  int gb = 0;
  if (lane == 0) {
    int lo = 0, hi = n_graphs;  // offsets[lo] <= r0 < offsets[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (offsets[mid] <= r0) lo = mid; else hi = mid;
    }
    gb = lo;
  }
  gb = __shfl(gb, 0, 64);  // graph of the chunk's first row
  int g_run = -1;
  int32_t mx = -1;
  for (int64_t b = r0; b < r1; b += 64) {
    const int64_t r = b + lane;
    int g = gb;
    int32_t d = -1;
    if (r < r1) {
      while (offsets[g + 1] <= r) ++g;  // r < offsets[n_graphs] = m ends it
      d = dist[r];
      if (d == 2147483647) d = -1;
    }
    const bool uniform = __ballot(r < r1 && g != gb) == 0ull;
    if (uniform) {
      int32_t v = d;
      for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
      if (gb != g_run) {
        if (lane == 0 && g_run >= 0 && mx >= 0) atomicMax(&levels[g_run], mx);
        g_run = gb;
        mx = -1;
      }
      mx = max(mx, v);
    } else {  // rows of two or more graphs: each lane folds its own row
      if (r < r1 && d >= 0) atomicMax(&levels[g], d);
    }
    // the next chunk starts in the graph of this chunk's last row
    gb = __shfl(r < r1 ? g : gb, 63, 64);
  }
  if (lane == 0 && g_run >= 0 && mx >= 0) atomicMax(&levels[g_run], mx);
}

// --------------------------------------------------------------- host side
struct BfsWork {
  int64_t m = 0, words = 0;
  uint32_t* visited = nullptr;
  uint32_t* front = nullptr;
  uint32_t* next = nullptr;
  int32_t* q0 = nullptr;
  int32_t* q1 = nullptr;
  BfsCounters* cnt = nullptr;
};

static int bfs_work(gg_ctx* ctx, int64_t m, bool queues, BfsWork& w) {
  w.m = m;
  w.words = (m + 31) / 32;
  void* p;
  GG_TRY(ws_get(ctx, WS_BFS0, sizeof(uint32_t) * w.words, &p)); w.visited = (uint32_t*)p;
  GG_TRY(ws_get(ctx, WS_BFS1, sizeof(uint32_t) * w.words, &p)); w.front = (uint32_t*)p;
  GG_TRY(ws_get(ctx, WS_BFS2, sizeof(uint32_t) * w.words, &p)); w.next = (uint32_t*)p;
  GG_TRY(ws_get(ctx, WS_FLAG, sizeof(BfsCounters), &p)); w.cnt = (BfsCounters*)p;
  if (queues) {
    GG_TRY(ws_get(ctx, WS_BFS3, sizeof(int32_t) * m, &p)); w.q0 = (int32_t*)p;
    GG_TRY(ws_get(ctx, WS_BFS4, sizeof(int32_t) * m, &p)); w.q1 = (int32_t*)p;
  }
  return GG_OK;
}

static void fill_words(hipStream_t st, uint32_t* p, int64_t n, uint32_t v) {
  hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(n, kB)), dim3(kB), 0, st, p, n, v);
}
static void fill_words(gg_ctx* ctx, uint32_t* p, int64_t n, uint32_t v) {
  fill_words(ctx->stream, p, n, v);
}

static int read_counters(gg_ctx* ctx, const BfsWork& w, BfsCounters* host) {
  GG_HIP(hipMemcpyAsync(ctx->host_pinned, w.cnt, sizeof(BfsCounters), hipMemcpyDeviceToHost,
                        ctx->stream));
  GG_HIP(hipStreamSynchronize(ctx->stream));
  memcpy(host, ctx->host_pinned, sizeof(BfsCounters));
  return GG_OK;
}

// One pull level (levels or parents) over A_T, front -> next, at the level
// and into the output that cnt holds.  `st` is the context stream, or the
// capture stream of a graph (then the launch is not timed).
static int pull_step(gg_ctx* ctx, hipStream_t st, const gg_matrix* A, bool parents,
                     const uint32_t* front, uint32_t* next, uint32_t* visited, BfsCounters* cnt) {
  const bool unit = (A->flags & GG_FLAG_UNIT_PATTERN) != 0;
  HotTimer ht(ctx, st == ctx->stream);
  return with_type(A->val_dtype, [&](auto ta) -> int {
    typedef typename decltype(ta)::type TA;
    auto go = [&](auto unit_tag, auto par_tag) -> int {
      constexpr bool U = decltype(unit_tag)::value;
      constexpr bool P = decltype(par_tag)::value;
      if (A->format == GG_CSR) {
        const int lg = csr_log_lanes(ctx, A);
        unsigned nb = blocks_for(A->m << lg, kB);
        const unsigned cap = (unsigned)ctx->num_cus * 32;
        if (nb > cap) nb = cap;
        hipLaunchKernelGGL((bfs_pull_csr_kernel<TA, U, P>), dim3(nb), dim3(kB), 0, st, A->ptr,
                           A->idx, (const TA*)A->val, A->m, lg, front, next, visited, cnt);
      } else {
        SellView g = sell_view(A);
        const int64_t nsb = (g.n_slots + kB - 1) / kB;
        const int64_t nb = nsb + (P ? g.n_heavy : g.n_seg);
        if (nb > 0)
          hipLaunchKernelGGL((bfs_pull_sell_kernel<TA, U, P>), dim3((unsigned)nb), dim3(kB), 0, st,
                             g, nsb, front, next, visited, cnt);
      }
      GG_HIP(hipGetLastError());
      return GG_OK;
    };
    if (parents) return go(std::true_type{}, std::true_type{});
    if (unit) return go(std::true_type{}, std::false_type{});
    if constexpr (std::is_floating_point<TA>::value) {
      return set_error(GG_ERR_ARGUMENT, "or-and BFS needs Bool or integer matrix values "
                                        "(GPUGraphs_band is bitwise, src/algebra.jl:98)");
    } else {
      return go(std::false_type{}, std::false_type{});
    }
  });
}

static const int32_t* vinv_of(const gg_matrix* A) {
  return (A->format == GG_SELL && (A->flags & GG_FLAG_INTERNAL_IDS)) ? A->vinv : nullptr;
}

// Clears the bitmaps (fill kernels, never a memset node) and seeds the source.
static int bfs_seed(gg_ctx* ctx, const gg_matrix* A, int64_t source, int32_t* out,
                    int32_t src_val, BfsWork& w, bool queues) {
  hipLaunchKernelGGL(fill_i32v_kernel, dim3(blocks_for(A->m, kB)), dim3(kB), 0, ctx->stream, out,
                     A->m, (int32_t)2147483647);
  fill_words(ctx, w.visited, w.words, 0u);
  fill_words(ctx, w.front, w.words, 0u);
  hipLaunchKernelGGL(bfs_seed_kernel, dim3(1), dim3(1), 0, ctx->stream, out, src_val, w.visited,
                     w.front, queues ? w.q0 : nullptr, source - 1, vinv_of(A));
  GG_HIP(hipGetLastError());
  return GG_OK;
}

static int bfs_begin(gg_ctx* ctx, const gg_matrix* A, int64_t source, int32_t* out,
                     int32_t src_val, BfsWork& w, bool queues) {
  GG_TRY(bfs_work(ctx, state_size(A), queues, w));
  return bfs_seed(ctx, A, source, out, src_val, w, queues);
}

static int check_square_source(const gg_matrix* A, int64_t source) {
  GG_TRY(check_matrix(A));
  if (A->format == GG_CSC) return set_error(GG_ERR_ARGUMENT, "BFS runs on CSR or SELL A_T");
  if (A->m != A->n) return set_error(GG_ERR_DIMENSION, "A_T must be square");
  if (source < 1 || source > A->m) return set_error(GG_ERR_BOUNDS, "source %lld outside 1:%lld",
                                                    (long long)source, (long long)A->m);
  return GG_OK;
}

// Host-stepped pull levels (one round trip per level) until one finds nothing.
static int pull_levels_host(gg_ctx* ctx, const gg_matrix* A, int32_t* dist, BfsWork& w,
                            int64_t* nlevels) {
  BfsCounters hc{};
  int32_t level = 0;
  while (true) {
    ++level;
    fill_words(ctx, w.next, w.words, 0u);
    hipLaunchKernelGGL(level_begin_kernel, dim3(1), dim3(1), 0, ctx->stream, w.cnt, level, dist);
    GG_TRY(pull_step(ctx, ctx->stream, A, false, w.front, w.next, w.visited, w.cnt));
    GG_TRY(read_counters(ctx, w, &hc));
    if (hc.found == 0) break;
    std::swap(w.front, w.next);
  }
  if (nlevels) *nlevels = level - 1;
  return GG_OK;
}

// Pull levels until one finds nothing.  With graph steps, U levels are
// captured (front and next alternate, U even) and relaunched until the
// device flag says done; otherwise one level per host round trip.
int bfs_levels_pull(gg_ctx* ctx, const gg_matrix* A, int64_t source, int32_t* dist,
                    int64_t* nlevels) {
  BfsWork w;
  GG_TRY(bfs_begin(ctx, A, source, dist, 0, w, false));
  BfsCounters hc{};
  if (ctx->graph_steps > 0) {
    const int U = (ctx->graph_steps + 1) & ~1;
    hipLaunchKernelGGL(level_begin_kernel, dim3(1), dim3(1), 0, ctx->stream, w.cnt, 1, dist);
    std::vector<uint64_t> key{1, (uint64_t)U};
    sig_add(key, *A);
    sig_add(key, w);
    auto steps = [&](hipStream_t st) -> int {
      for (int i = 0; i < U; ++i) {
        uint32_t* f = (i & 1) ? w.next : w.front;
        uint32_t* nx = (i & 1) ? w.front : w.next;
        fill_words(st, nx, w.words, 0u);
        GG_TRY(pull_step(ctx, st, A, false, f, nx, w.visited, w.cnt));
        hipLaunchKernelGGL(level_end_kernel, dim3(1), dim3(1), 0, st, w.cnt);
      }
      GG_HIP(hipGetLastError());
      return GG_OK;
    };
    GG_TRY(graph_loop(ctx, key, &w.cnt->done, A->m / U + 2, steps, nullptr));
    GG_TRY(read_counters(ctx, w, &hc));
    if (nlevels) *nlevels = hc.level - 1;
    return GG_OK;
  }
  return pull_levels_host(ctx, A, dist, w, nlevels);
}

// ------------------------------------------------- allocation-free variant
// Traversal state of gg_bfs_distances_noalloc: the visited / frontier / next
// bitmaps of up to kScratchVertices state rows and the level counters, in one
// device array that lives as long as the process (each GPU holds its own
// copy).  Every call on a GPU shares that GPU's copy.
constexpr int64_t kScratchVertices = int64_t(1) << 26;
constexpr int64_t kScratchWords = kScratchVertices / 32;
static __device__ __align__(256) uint32_t scratch[3 * kScratchWords + 64];
static_assert(sizeof(BfsCounters) <= 64 * sizeof(uint32_t), "counters fit the scratch tail");

// Device address of `scratch` per device, resolved once per device with the
// device current (the host-side name of the array is not a device address).
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

static int scratch_address(int device, void** out) {
  if (device < 0 || device >= 64) return set_error(GG_ERR_ARGUMENT, "device %d out of range", device);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    GG_HIP(hipSetDevice(device));
    void* p = nullptr;
    GG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return GG_OK;
}

}  // namespace gg

using namespace gg;

extern "C" {

int gg_bfs_distances(gg_ctx* ctx, const gg_matrix* A_T, int64_t source, int32_t* dist,
                     int use_mask) {
  (void)use_mask;  // the unvisited mask is applied in-kernel either way; levels are equal
  if (!ctx || !dist) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  GG_TRY(check_square_source(A_T, source));
  timer_begin(ctx);
  int st = bfs_levels_pull(ctx, A_T, source, dist, nullptr);
  timer_end(ctx);
  return st;
}

int gg_bfs_distances_noalloc(gg_ctx* ctx, const gg_matrix* A_T, int64_t source, int32_t* dist,
                             int use_mask) {
  (void)use_mask;
  if (!ctx || !dist) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  GG_TRY(check_square_source(A_T, source));
  const int64_t ns = state_size(A_T);
  if (ns > kScratchVertices)
    return set_error(GG_ERR_ARGUMENT,
                     "gg_bfs_distances_noalloc handles at most %lld vertices (state rows); "
                     "this matrix needs %lld: use gg_bfs_distances",
                     (long long)kScratchVertices, (long long)ns);
  void* base;
  GG_TRY(scratch_address(ctx->device, &base));
  BfsWork w;
  w.m = ns;
  w.words = (ns + 31) / 32;
  w.visited = (uint32_t*)base;
  w.front = w.visited + kScratchWords;
  w.next = w.front + kScratchWords;
  w.cnt = (BfsCounters*)(w.next + kScratchWords);
  timer_begin(ctx);
  GG_TRY(bfs_seed(ctx, A_T, source, dist, 0, w, false));
  const int st = pull_levels_host(ctx, A_T, dist, w, nullptr);
  timer_end(ctx);
  return st;
}

int gg_bfs_parents(gg_ctx* ctx, const gg_matrix* A_T, int64_t source, int32_t* parents,
                   int64_t batch_size) {
  if (!ctx || !parents) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  GG_TRY(check_square_source(A_T, source));
  if (batch_size < 1) return set_error(GG_ERR_ARGUMENT, "batch_size must be >= 1");
  timer_begin(ctx);
  BfsWork w;
  GG_TRY(bfs_begin(ctx, A_T, source, parents, 0, w, false));
  BfsCounters hc{};
  int32_t level = 0;
  while (true) {
    ++level;
    fill_words(ctx, w.next, w.words, 0u);
    hipLaunchKernelGGL(level_begin_kernel, dim3(1), dim3(1), 0, ctx->stream, w.cnt, level, parents);
    GG_TRY(pull_step(ctx, ctx->stream, A_T, true, w.front, w.next, w.visited, w.cnt));
    GG_TRY(read_counters(ctx, w, &hc));
    if (hc.found == 0) break;
    std::swap(w.front, w.next);
  }
  timer_end(ctx);
  return GG_OK;
}

}  // extern "C"

namespace gg {

static int check_do_pair(const gg_matrix* A_T, const gg_matrix* A_out) {
  GG_TRY(check_matrix(A_out));
  if (A_out->format != GG_CSR || A_out->m != A_T->m || A_out->n != A_T->n)
    return set_error(GG_ERR_ARGUMENT, "A_out must be the CSR of A (out-neighbour rows), same size");
  if (!(A_T->flags & GG_FLAG_UNIT_PATTERN) || !(A_out->flags & GG_FLAG_UNIT_PATTERN))
    return set_error(GG_ERR_ARGUMENT, "direction-optimising BFS needs unit-pattern matrices");
  return GG_OK;
}

// Direction-optimising levels from the seeded state: w.q0 holds the qlen
// level-0 vertices (original ids), `scout` their out-degree sum.
static int do_levels(gg_ctx* ctx, const gg_matrix* A_T, const gg_matrix* A_out, BfsWork& w,
                     int64_t qlen, double scout, int32_t* dist, double alpha, double beta) {
  if (alpha <= 0) alpha = 15.0;
  if (beta <= 0) beta = 18.0;
  const int64_t n = A_T->m;
  const int32_t* vinv = vinv_of(A_T);
  SellView g{};  // state index -> original vertex for the queue rebuild
  if (A_T->format == GG_SELL) g = sell_view(A_T);
  BfsCounters hc{};
  double edges_to_check = (double)A_out->nnz;
  int32_t level = 0;
  // hub pieces: at most one per kBigDeg edges plus one per hub
  void* bp;
  GG_TRY(ws_get(ctx, WS_BFS5, sizeof(int2) * (size_t)(2 * (A_out->nnz / kBigDeg) + 16), &bp));
  int2* big = (int2*)bp;
  while (qlen > 0) {
    if (scout > edges_to_check / alpha) {
      int64_t awake = qlen, old;
      do {
        old = awake;
        ++level;
        fill_words(ctx, w.next, w.words, 0u);
        hipLaunchKernelGGL(level_begin_kernel, dim3(1), dim3(1), 0, ctx->stream, w.cnt, level,
                           dist);
        GG_TRY(pull_step(ctx, ctx->stream, A_T, false, w.front, w.next, w.visited, w.cnt));
        hipLaunchKernelGGL(popcount_kernel, dim3(std::min<int64_t>(blocks_for(w.words, kB), 1024)),
                           dim3(kB), 0, ctx->stream, w.next, w.words, &w.cnt->awake);
        GG_HIP(hipGetLastError());
        GG_TRY(read_counters(ctx, w, &hc));
        awake = hc.awake;
        std::swap(w.front, w.next);
      } while (awake > 0 && (awake >= old || (double)awake > (double)n / beta));
      if (awake == 0) break;
      hipLaunchKernelGGL(level_begin_kernel, dim3(1), dim3(1), 0, ctx->stream, w.cnt, level, dist);
      hipLaunchKernelGGL(bitmap_to_queue_kernel, dim3(blocks_for(w.words, kB)), dim3(kB), 0,
                         ctx->stream, w.front, w.words, w.m, g, w.q0, w.cnt);
      GG_HIP(hipGetLastError());
      GG_TRY(read_counters(ctx, w, &hc));
      qlen = hc.qlen;
      scout = 1;
    } else {
      edges_to_check -= scout;
      ++level;
      fill_words(ctx, w.next, w.words, 0u);
      hipLaunchKernelGGL(level_begin_kernel, dim3(1), dim3(1), 0, ctx->stream, w.cnt, level, dist);
      const unsigned cap = (unsigned)ctx->num_cus * 16;
      const int tile = (int)std::min<int64_t>(kB, std::max<int64_t>(1, (qlen + cap - 1) / cap));
      PushArgs pa{A_out->ptr, A_out->idx, w.q0, qlen, tile, w.q1, w.next, w.visited, dist,
                  level, vinv, w.cnt, big};
      unsigned nb = blocks_for(qlen, tile);
      if (nb > cap) nb = cap;
      {
        HotTimer ht(ctx);
        hipLaunchKernelGGL(push_tile_kernel, dim3(nb), dim3(kB), 0, ctx->stream, pa);
        hipLaunchKernelGGL(push_big_kernel, dim3((unsigned)ctx->num_cus * 4), dim3(kB), 0,
                           ctx->stream, pa);
      }
      GG_HIP(hipGetLastError());
      GG_TRY(read_counters(ctx, w, &hc));
      qlen = hc.qlen;
      scout = (double)hc.scout;
      std::swap(w.q0, w.q1);
      std::swap(w.front, w.next);
    }
  }
  return GG_OK;
}

}  // namespace gg

extern "C" {

int gg_bfs_distances_do(gg_ctx* ctx, const gg_matrix* A_T, const gg_matrix* A_out,
                        int64_t source, int32_t* dist, double alpha, double beta) {
  if (!ctx || !dist) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  GG_TRY(check_square_source(A_T, source));
  GG_TRY(check_do_pair(A_T, A_out));
  timer_begin(ctx);
  BfsWork w;
  GG_TRY(bfs_begin(ctx, A_T, source, dist, 0, w, true));
  BfsCounters hc{};
  hipLaunchKernelGGL(degree_kernel, dim3(1), dim3(1), 0, ctx->stream, A_out->ptr, source - 1, w.cnt);
  GG_TRY(read_counters(ctx, w, &hc));
  const int st = do_levels(ctx, A_T, A_out, w, 1, (double)hc.scout, dist, alpha, beta);
  timer_end(ctx);
  return st;
}

int gg_bfs_packed_do(gg_ctx* ctx, const gg_matrix* A_T, const gg_matrix* A_out,
                     const int64_t* offsets, int n_graphs, const int32_t* sources, int32_t* dist,
                     int32_t* levels, double alpha, double beta) {
  if (!ctx || !A_T || !offsets || !sources || !dist || n_graphs < 1)
    return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  GG_TRY(check_matrix(A_T));
  if (A_T->format == GG_CSC) return set_error(GG_ERR_ARGUMENT, "BFS runs on CSR or SELL A_T");
  if (A_T->m != A_T->n) return set_error(GG_ERR_DIMENSION, "A_T must be square");
  GG_TRY(check_do_pair(A_T, A_out));
  if (offsets[0] != 0 || offsets[n_graphs] != A_T->m)
    return set_error(GG_ERR_DIMENSION, "offsets do not cover the %lld rows", (long long)A_T->m);
  std::vector<int32_t> seeds(n_graphs);
  for (int gi = 0; gi < n_graphs; ++gi) {
    if (offsets[gi + 1] < offsets[gi])
      return set_error(GG_ERR_ARGUMENT, "graph offsets must not decrease");
    if (sources[gi] < 1 || sources[gi] > offsets[gi + 1] - offsets[gi])
      return set_error(GG_ERR_BOUNDS, "source %d of graph %d outside its rows", sources[gi], gi);
    seeds[gi] = (int32_t)(offsets[gi] + sources[gi] - 1);
  }
  const int64_t m = A_T->m;
  timer_begin(ctx);
  BfsWork w;
  GG_TRY(bfs_work(ctx, state_size(A_T), true, w));
  // offsets (int64, n + 1), then the seeds (int32, n) and the levels (int32, n)
  void* p;
  GG_TRY(ws_get(ctx, WS_SPMSPV0, sizeof(int64_t) * (n_graphs + 1) + 2 * sizeof(int32_t) * n_graphs,
                &p));
  int64_t* d_off = (int64_t*)p;
  int32_t* d_seeds = (int32_t*)(d_off + n_graphs + 1);
  int32_t* d_levels = d_seeds + n_graphs;
  GG_HIP(hipMemcpyAsync(d_off, offsets, sizeof(int64_t) * (n_graphs + 1), hipMemcpyHostToDevice,
                        ctx->stream));
  GG_HIP(hipMemcpyAsync(d_seeds, seeds.data(), sizeof(int32_t) * n_graphs, hipMemcpyHostToDevice,
                        ctx->stream));
  hipLaunchKernelGGL(fill_i32v_kernel, dim3(blocks_for(m, kB)), dim3(kB), 0, ctx->stream, dist, m,
                     (int32_t)2147483647);
  fill_words(ctx, w.visited, w.words, 0u);
  fill_words(ctx, w.front, w.words, 0u);
  hipLaunchKernelGGL(level_begin_kernel, dim3(1), dim3(1), 0, ctx->stream, w.cnt, 0, dist);
  hipLaunchKernelGGL(bfs_seeds_kernel, dim3(blocks_for(n_graphs, kB)), dim3(kB), 0, ctx->stream,
                     dist, w.visited, w.front, w.q0, d_seeds, (int64_t)n_graphs, vinv_of(A_T),
                     A_out->ptr, w.cnt);
  GG_HIP(hipGetLastError());
  BfsCounters hc{};
  GG_TRY(read_counters(ctx, w, &hc));
  GG_TRY(do_levels(ctx, A_T, A_out, w, n_graphs, (double)hc.scout, dist, alpha, beta));
  if (levels) {
    hipLaunchKernelGGL(fill_i32v_kernel, dim3(blocks_for(n_graphs, kB)), dim3(kB), 0, ctx->stream,
                       d_levels, (int64_t)n_graphs, 0);
    const int64_t waves = (m + kLvRows - 1) / kLvRows;
    hipLaunchKernelGGL(graph_levels_kernel, dim3(blocks_for(waves * 64, kB)), dim3(kB), 0,
                       ctx->stream, dist, m, d_off, n_graphs, d_levels);
    GG_HIP(hipGetLastError());
    GG_HIP(hipMemcpyAsync(levels, d_levels, sizeof(int32_t) * n_graphs, hipMemcpyDeviceToHost,
                          ctx->stream));
    GG_HIP(hipStreamSynchronize(ctx->stream));
  }
  timer_end(ctx);
  return GG_OK;
}

}  // extern "C"
