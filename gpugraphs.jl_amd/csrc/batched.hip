// batched.hip -- multi-source BFS as or-and SpMM over source bitmasks
// (gg_msbfs_distances).
//
// k BFS runs from k sources share one traversal: vertex v holds W = ceil(k/64)
// 64-bit words whose bit s says "v is in source s's frontier".  One level is
// the or-and SpMM  next[v] = (OR over in-neighbours u of front[u]) & ~seen[v]
// (the bitmask form of src/algorithms/bfs.jl:52-67 applied to k columns at
// once); each newly set bit s records dist[v, s] = level.  A row stops
// scanning as soon as every source has reached it (the or-monoid's absorbing
// all-ones), rows reached by every source are skipped, and a 1-bit
// "in any frontier" summary (n bits, L2-resident) avoids fetching the W-word
// masks of vertices that are in no frontier.
//
// dist[:, s] equals the single-source BFS levels from sources[s]: the
// bit-parallel traversal is k independent level-synchronous BFS runs.
#include <cstring>

#include "batched.hpp"

namespace gg {

constexpr int kM = 256;

// MsArgs: batched.hpp


// Bits of word w that belong to one of the k sources.
inline uint64_t word_mask(int64_t k, int w) {
  const int64_t lo = (int64_t)w * 64;
  if (lo >= k) return 0ull;
  if (lo + 64 <= k) return ~0ull;
  return (1ull << (k - lo)) - 1ull;
}

template <int W>
__device__ __forceinline__ bool finish_row(const MsArgs& a, int32_t level, int64_t rid,
                                           int64_t row, const uint64_t* acc, const uint64_t* sn,
                                           uint64_t* nw, bool claim_any = true) {
  bool any = false;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    nw[w] = acc[w] & ~sn[w] & a.live[w];
    any |= nw[w] != 0;
  }
#pragma unroll
  for (int w = 0; w < W; ++w) {
    a.next[rid * W + w] = nw[w];
    if (nw[w]) a.seen[rid * W + w] = sn[w] | nw[w];
  }
  if (any) {
    if (claim_any) bit_claim(a.next_any, rid);
    if (a.dist8) {
      // byte levels: one 8-byte write per group of 8 sources with a new bit
      // (this thread is the row's only writer in the level), read first only
      // when a source of the group was reached before: the bytes of sources
      // not reached yet still hold the 255 fill
      const uint64_t lv = (uint64_t)(uint8_t)level;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        for (int gq = 0; gq < 8; ++gq) {
          const uint32_t bits = (uint32_t)(nw[w] >> (8 * gq)) & 0xffu;
          if (!bits) continue;
          uint64_t* p = (uint64_t*)(a.dist8 + row * a.k8 + w * 64 + gq * 8);
          uint64_t v = ((sn[w] >> (8 * gq)) & 0xffu) ? *p : ~0ull;
          for (int t = 0; t < 8; ++t)
            if ((bits >> t) & 1u) v = (v & ~(0xffull << (8 * t))) | (lv << (8 * t));
          *p = v;
        }
      }
    } else if (a.dist) {
#pragma unroll
      for (int w = 0; w < W; ++w) {
        uint64_t b = nw[w];
        while (b) {
          const int s = w * 64 + __ffsll((unsigned long long)b) - 1;
          b &= b - 1;
          const int32_t d = a.start ? level - a.start[s] : level;
          if (a.layout == GG_ROW_MAJOR) a.dist[row * a.k + s] = d;
          else a.dist[row + (int64_t)s * a.m] = d;
        }
      }
    }
  }
  return any;
}

template <int W>
__device__ __forceinline__ bool saturated(const uint64_t* acc, const uint64_t* sn,
                                          const MsArgs& a) {
  bool full = true;
#pragma unroll
  for (int w = 0; w < W; ++w) full &= ((acc[w] | sn[w]) & a.live[w]) == a.live[w];
  return full;
}

// OR of the wavefront's next words into a.alive (all lanes active).
template <int W> __device__ __forceinline__ void alive_wave(const MsArgs& a, const uint64_t* nw) {
  if (!a.alive) return;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    uint64_t v = nw[w];
    for (int o = 32; o > 0; o >>= 1) v |= shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0 && v) or_new_bits(&a.alive[w], (unsigned long long)v);
  }
}

// Entries per lane per batch of msbfs_sell_kernel's light rows: more loads in
// flight per lane; with more mask words per vertex fewer, for the registers.
template <int W> constexpr int ms_batch() { return W == 1 ? 16 : 8; }

template <class TA, bool UNIT, int W>
__global__ __launch_bounds__(kM) void msbfs_sell_kernel(SellView g, int64_t n_sell_blocks, MsArgs a) {
  if (a.ctl && a.ctl[1]) return;  // a captured step after the last level
  if (a.pull_off && *a.pull_off) return;  // this level runs top-down
  const int32_t level = a.ctl ? a.ctl[0] : a.level;
  const bool test_any = !(a.dense && *a.dense);
  if ((int64_t)blockIdx.x < n_sell_blocks) {
    int64_t row, base, len;
    bool has_len;
    bool newly = false;
    uint64_t nw[W];
#pragma unroll
    for (int w = 0; w < W; ++w) nw[w] = 0;
    const int64_t slot = (int64_t)blockIdx.x * kM + threadIdx.x;
    if (sell_slot(g, slot, row, base, len, has_len)) {
      const int64_t rid = g.internal ? slot : row;
      uint64_t sn[W], acc[W];
#pragma unroll
      for (int w = 0; w < W; ++w) { sn[w] = a.seen[rid * W + w]; acc[w] = 0; }
      if (!saturated<W>(acc, sn, a)) {
        const TA* val = (const TA*)g.val;
        int64_t kk = 0;
        if (has_len) {  // batches of kMsBatch: index loads, bitmap tests, then mask loads
          constexpr int kMsBatch = ms_batch<W>();
          bool full = false;
          for (; kk + kMsBatch <= len && !full; kk += kMsBatch) {
            int32_t c[kMsBatch];
            bool on[kMsBatch];
#pragma unroll
            for (int u = 0; u < kMsBatch; ++u) c[u] = __builtin_nontemporal_load(&g.colval[base + (kk + u) * g.C]);
#pragma unroll
            for (int u = 0; u < kMsBatch; ++u)
              on[u] = (!test_any || bit_get(a.front_any, c[u] - 1)) &&
                      (UNIT || edge_on(val, base + (kk + u) * g.C));
#pragma unroll
            for (int u = 0; u < kMsBatch; ++u)
              if (on[u])
#pragma unroll
                for (int w = 0; w < W; ++w) acc[w] |= a.front[(int64_t)(c[u] - 1) * W + w];
            full = saturated<W>(acc, sn, a);
          }
          // the rest of the row (short rows: all of it) in masked batches of
          // 8: their loads are issued together where the entry-by-entry loop
          // below waits for each one (grids and other low-degree graphs)
          for (; kk < len && !full; kk += 8) {
            int32_t c[8];
            bool on[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
              c[u] = kk + u < len ? __builtin_nontemporal_load(&g.colval[base + (kk + u) * g.C]) : 0;
#pragma unroll
            for (int u = 0; u < 8; ++u)
              on[u] = c[u] > 0 && (!test_any || bit_get(a.front_any, c[u] - 1)) &&
                      (UNIT || edge_on(val, base + (kk + u) * g.C));
#pragma unroll
            for (int u = 0; u < 8; ++u)
              if (on[u])
#pragma unroll
                for (int w = 0; w < W; ++w) acc[w] |= a.front[(int64_t)(c[u] - 1) * W + w];
            full = saturated<W>(acc, sn, a);
          }
          kk = len;
        }
        for (; kk < len; ++kk) {
          const int64_t i = base + kk * g.C;
          const int32_t col = g.colval[i];
          if (!has_len && col == -1) break;
          if ((test_any && !bit_get(a.front_any, col - 1)) || !(UNIT || edge_on(val, i))) continue;
#pragma unroll
          for (int w = 0; w < W; ++w) acc[w] |= a.front[(int64_t)(col - 1) * W + w];
          if (saturated<W>(acc, sn, a)) break;
        }
        newly = finish_row<W>(a, level, rid, row, acc, sn, nw, !g.internal);
      } else {
#pragma unroll
        for (int w = 0; w < W; ++w) a.next[rid * W + w] = 0;
      }
    }
    if (g.internal) wave_set_bits(a.next_any, slot & ~(int64_t)63, newly);
    wave_flag(a.found, newly);
    alive_wave<W>(a, nw);
    return;
  }
  // one block per heavy segment: OR across the block, early exit once
  // saturated.  A row of several segments ORs them into a.hacc; the last
  // segment to arrive (a.hcnt) finishes the row and clears both, so they are
  // zero again for the next level.  A row saturated before the level is
  // settled by its first segment alone (every segment sees the same seen).
  __shared__ uint64_t red[kM / 64][W];
  __shared__ int s_full;
  const int64_t sg = (int64_t)blockIdx.x - n_sell_blocks;
  int64_t h, s, e;
  heavy_segment(g, sg, h, s, e);
  const int64_t nseg = (int64_t)g.heavy_seg[h + 1] - g.heavy_seg[h];
  const int64_t row = g.heavy_row[h] - 1;
  const int64_t rid = heavy_rid(g, h);
  uint64_t sn[W], acc[W];
#pragma unroll
  for (int w = 0; w < W; ++w) { sn[w] = a.seen[rid * W + w]; acc[w] = 0; }
  const bool full0 = saturated<W>(acc, sn, a);
  if (full0) {
    if (sg == g.heavy_seg[h] && threadIdx.x == 0)
#pragma unroll
      for (int w = 0; w < W; ++w) a.next[rid * W + w] = 0;
    return;
  }
  if (threadIdx.x == 0) {
    // segments that start after the row's earlier segments saturated it
    // (their OR is in hacc) skip the scan and only report arrival
    bool full = false;
    if (nseg > 1) {
      uint64_t cur[W];
#pragma unroll
      for (int w = 0; w < W; ++w) cur[w] = __atomic_load_n(&a.hacc[h * W + w], __ATOMIC_RELAXED);
      full = saturated<W>(cur, sn, a);
    }
    s_full = full;
  }
  __syncthreads();
  {
    const TA* hv = (const TA*)g.heavy_val;
    for (int64_t i = s + threadIdx.x; i < e; i += kM) {
      if (__atomic_load_n(&s_full, __ATOMIC_RELAXED)) break;
      const int32_t col = g.heavy_col[i];
      if ((test_any && !bit_get(a.front_any, col - 1)) || !(UNIT || edge_on(hv, i))) continue;
#pragma unroll
      for (int w = 0; w < W; ++w) acc[w] |= a.front[(int64_t)(col - 1) * W + w];
      if (saturated<W>(acc, sn, a)) __atomic_store_n(&s_full, 1, __ATOMIC_RELAXED);
    }
  }
#pragma unroll
  for (int w = 0; w < W; ++w)
    for (int o = 32; o > 0; o >>= 1) acc[w] |= shfl_xor(acc[w], o);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int w = 0; w < W; ++w) red[threadIdx.x >> 6][w] = acc[w];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < W; ++w)
      for (int q = 1; q < kM / 64; ++q) acc[w] |= red[q][w];
    if (nseg > 1) {
#pragma unroll
      for (int w = 0; w < W; ++w)
        if (acc[w]) atomicOr(&a.hacc[h * W + w], (unsigned long long)acc[w]);
      __threadfence();
      if (atomicAdd(&a.hcnt[h], 1u) != (unsigned)(nseg - 1)) return;  // not the last segment
      __threadfence();
#pragma unroll
      for (int w = 0; w < W; ++w) acc[w] = atomicExch(&a.hacc[h * W + w], 0ull);
      atomicExch(&a.hcnt[h], 0u);
    }
    uint64_t nw[W];
    if (finish_row<W>(a, level, rid, row, acc, sn, nw)) set_flag(a.found);
    if (a.alive)
#pragma unroll
      for (int w = 0; w < W; ++w)
        if (nw[w]) or_new_bits(&a.alive[w], (unsigned long long)nw[w]);
  }
}

template <class TA, bool UNIT, int W>
__global__ __launch_bounds__(kM) void msbfs_csr_kernel(const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ colval,
                                                       const TA* __restrict__ val, MsArgs a) {
  if (a.ctl && a.ctl[1]) return;
  if (a.pull_off && *a.pull_off) return;
  const int32_t level = a.ctl ? a.ctl[0] : a.level;
  const bool test_any = !(a.dense && *a.dense);
  const int64_t row = (int64_t)blockIdx.x * kM + threadIdx.x;
  bool newly = false;
  uint64_t nw[W];
#pragma unroll
  for (int w = 0; w < W; ++w) nw[w] = 0;
  if (row < a.m) {
    uint64_t sn[W], acc[W];
#pragma unroll
    for (int w = 0; w < W; ++w) { sn[w] = a.seen[row * W + w]; acc[w] = 0; }
    if (!saturated<W>(acc, sn, a)) {
      for (int32_t i = rowptr[row] - 1; i < rowptr[row + 1] - 1; ++i) {
        const int32_t col = colval[i];
        if ((test_any && !bit_get(a.front_any, col - 1)) || !(UNIT || edge_on(val, i))) continue;
#pragma unroll
        for (int w = 0; w < W; ++w) acc[w] |= a.front[(int64_t)(col - 1) * W + w];
        if (saturated<W>(acc, sn, a)) break;
      }
      newly = finish_row<W>(a, level, row, row, acc, sn, nw);
    } else {
#pragma unroll
      for (int w = 0; w < W; ++w) a.next[row * W + w] = 0;
    }
  }
  wave_flag(a.found, newly);
  alive_wave<W>(a, nw);
}

// One thread per source (several sources may share a vertex: atomic ORs);
// distances row-major m x k, bytes when dist8 is given.
// With ctl (captured levels) thread 0 also starts the level counters: level
// 1, not done, window base 0, found 0.
__global__ void msbfs_seed_kernel(const int32_t* sources, int64_t k, int64_t k8, int W,
                                  const int32_t* vinv, uint64_t* front, uint32_t* front_any,
                                  uint64_t* seen, int32_t* dist, uint8_t* dist8,
                                  unsigned int* found, int32_t* ctl) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ctl && s == 0) {
    *found = 0u;
    ctl[0] = 1;
    ctl[1] = 0;
    ctl[2] = 0;
  }
  if (s >= k) return;
  const int64_t v = sources[s] - 1;
  const int64_t r = vinv ? (int64_t)vinv[v] - 1 : v;  // state index
  atomicOr((unsigned long long*)&front[r * W + s / 64], 1ull << (s & 63));
  atomicOr((unsigned long long*)&seen[r * W + s / 64], 1ull << (s & 63));
  atomicOr(&front_any[r >> 5], 1u << (r & 31));
  if (dist8) dist8[v * k8 + s] = 0;
  else dist[v * k + s] = 0;
}

// Byte levels row-major m x k -> Int32 column-major m x k (255 = unreached ->
// typemax(Int32)) through a 64 x 64 LDS tile.  Reads are 8-byte words (k8 is
// a multiple of 8); with vec (m % 4 == 0, dst 16-byte aligned) each thread
// stores 4 rows of a column as one 16-byte non-temporal store, 16 threads per
// column, so a wavefront writes four whole 256-byte column pieces.
__device__ __forceinline__ int32_t widen_level(uint8_t d) {
  return d == 255 ? 2147483647 : (int32_t)d;
}

__global__ __launch_bounds__(256) void widen_transpose_kernel(const uint8_t* __restrict__ src,
                                                              int64_t m, int64_t k, int64_t k8,
                                                              int vec,
                                                              int32_t* __restrict__ dst) {
  __shared__ uint8_t tile[64][72];
  const int64_t tiles_k = (k + 63) / 64;
  const int64_t r0 = ((int64_t)blockIdx.x / tiles_k) * 64;
  const int64_t c0 = ((int64_t)blockIdx.x % tiles_k) * 64;
  for (int i = threadIdx.x; i < 64 * 8; i += 256) {
    const int r = i >> 3, q = i & 7;
    if (r0 + r < m && c0 + q * 8 < k8)
      *(uint64_t*)&tile[r][q * 8] = *(const uint64_t*)&src[(r0 + r) * k8 + c0 + q * 8];
  }
  __syncthreads();
  if (vec && r0 + 64 <= m) {
    const int rr = (threadIdx.x & 15) * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = (threadIdx.x >> 4) + 16 * j;
      if (c0 + c >= k) break;
      typedef int32_t v4i __attribute__((ext_vector_type(4)));
      v4i v;
      v.x = widen_level(tile[rr][c]);
      v.y = widen_level(tile[rr + 1][c]);
      v.z = widen_level(tile[rr + 2][c]);
      v.w = widen_level(tile[rr + 3][c]);
      __builtin_nontemporal_store(v, (v4i*)&dst[(r0 + rr) + (c0 + c) * m]);
    }
    return;
  }
  for (int i = threadIdx.x; i < 64 * 64; i += 256) {
    const int c = i >> 6, r = i & 63;
    if (r0 + r < m && c0 + c < k) dst[(r0 + r) + (c0 + c) * m] = widen_level(tile[r][c]);
  }
}

// Row-major m x k -> column-major m x k through a 64 x 64 LDS tile, so both
// the reads and the writes are whole lines.
__global__ __launch_bounds__(256) void transpose_rm_cm_kernel(const int32_t* __restrict__ src,
                                                              int64_t m, int64_t k,
                                                              int32_t* __restrict__ dst) {
  __shared__ int32_t tile[64][65];
  const int64_t tiles_k = (k + 63) / 64;
  const int64_t r0 = ((int64_t)blockIdx.x / tiles_k) * 64;
  const int64_t c0 = ((int64_t)blockIdx.x % tiles_k) * 64;
  for (int i = threadIdx.x; i < 64 * 64; i += 256) {
    const int r = i >> 6, c = i & 63;
    if (r0 + r < m && c0 + c < k) tile[r][c] = src[(r0 + r) * k + c0 + c];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 64 * 64; i += 256) {
    const int c = i >> 6, r = i & 63;
    if (r0 + r < m && c0 + c < k) dst[(r0 + r) + (c0 + c) * m] = tile[r][c];
  }
}

__global__ void fill_u64_kernel(uint64_t* p, int64_t n, uint64_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void ms_zero_kernel(unsigned int* p) { *p = 0u; }

// Captured levels: ctl = {level, done, window base}, started by the seed
// kernel at level 1, not done.  End of a level: done (1) when it found
// nothing, else the next level; done = 2 when the next level would reach
// `cap` (byte levels full; 0 = no cap).

// Up to 8 ranges of 32-bit words zeroed by one launch (grid-stride, 16-byte
// stores where a range allows).
struct ZeroRanges {
  uint32_t* p[8];
  int64_t n[8];
  int count;
  void add(void* ptr, int64_t words) {
    p[count] = (uint32_t*)ptr;
    n[count] = words;
    ++count;
  }
};
__global__ __launch_bounds__(256) void zero_ranges_kernel(ZeroRanges z) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x, stride = (int64_t)gridDim.x * 256;
  for (int r = 0; r < z.count; ++r) {
    uint32_t* p = z.p[r];
    const int64_t n = z.n[r];
    int64_t head = 0;
    if (((uintptr_t)p & 15) == 0) {
      typedef uint32_t u4 __attribute__((ext_vector_type(4)));
      const u4 zero = {0u, 0u, 0u, 0u};
      head = n & ~(int64_t)3;
      for (int64_t i = t; i < head / 4; i += stride) ((u4*)p)[i] = zero;
    }
    for (int64_t i = head + t; i < n; i += stride) p[i] = 0u;
  }
}

constexpr int kMsRingHost = 8;  // ring slots of a host-stepped traversal
constexpr int64_t kMsPinnedSources = 512;  // sources staged in ctx->host_pinned (2 KiB)
constexpr int kMsRingMax = 16;  // most ring slots of a captured one

// Byte t of the result is 0xff when bit t of b is set, else 0.
__device__ __forceinline__ uint64_t byte_mask8(uint32_t b) {
  uint64_t x = ((uint64_t)(b & 0xffu) * 0x0101010101010101ull) & 0x8040201008040201ull;
  x = (((x & 0x7f7f7f7f7f7f7f7full) + 0x7f7f7f7f7f7f7f7full) | x) & 0x8080808080808080ull;
  return (x >> 7) * 0xffull;
}

// Byte levels from a ring of R per-level masks.  Level L's new bits are its
// next words, kept in ring slot L % R (state order; slot L % R's any-bitmap
// tells which rows the level wrote) until a flush after every R levels moves
// the window base+1 .. base+R into the row-major byte levels.  One thread per
// (state row, mask word): it loads the R slots' words at once, composes the
// word's 8 groups of 8 byte levels and leaves them in LDS; then 8 lanes write
// each 64-byte piece of an original row together.  A group's earlier bytes
// (sources in seen but not in the window) are read first; sources not
// reached get 255.  Intermediate flushes write only groups with bits in the
// window; the last one writes every group of every row, so the byte levels
// need no fill.  Captured: base = ctl[2], the window ends at the level that
// found nothing when done, and done = 2 (overflow) records nothing;
// host-stepped: the arguments.
__global__ __launch_bounds__(256) void ms_flush_kernel(SellView g, const uint64_t* __restrict__ ring,
                                                       const uint32_t* __restrict__ ring_any,
                                                       int R, int64_t ns, int64_t words, int W,
                                                       int64_t k8, const uint64_t* __restrict__ seen,
                                                       uint8_t* __restrict__ dist8,
                                                       const int32_t* ctl, int32_t base_h,
                                                       int32_t last_h, int final_h) {
  __shared__ uint64_t sb[256][9];
  __shared__ int64_t sdst[256];
  __shared__ uint32_t swm[256];
  int32_t base = base_h, last = last_h;
  bool fin = final_h != 0;
  if (ctl) {
    if (ctl[1] == 2) return;  // uniform: every thread returns
    base = ctl[2];
    fin = ctl[1] == 1;
    last = fin ? ctl[0] : base + R;
  }
  const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int64_t dst = -1;
  uint32_t wmask = 0;  // groups of the piece to write
  if (gid < ns * W) {
    const int64_t rid = gid / W;
    const int w = (int)(gid - rid * W);
    const int gc = (int)min<int64_t>(8, k8 / 8 - 8 * w);
    const int64_t row = gc > 0 ? orig_of(g, rid) : -1;
    if (row >= 0) {  // not an empty slot, nor a word past k8
      // every slot's any-word, then every wanted mask word: the loads of a
      // phase are independent, so a wavefront waits twice, not 2R times
      uint32_t on[kMsRingMax];
#pragma unroll
      for (int j = 0; j < kMsRingMax; ++j)
        on[j] = j < R && base + (j == 0 ? R : j) <= last
                    ? (ring_any[(int64_t)j * words + (rid >> 5)] >> (rid & 31)) & 1u
                    : 0u;
      uint64_t mw[kMsRingMax];
      uint64_t win = 0;
#pragma unroll
      for (int j = 0; j < kMsRingMax; ++j) {
        mw[j] = on[j] ? ring[((int64_t)j * ns + rid) * W + w] : 0ull;
        win |= mw[j];
      }
      if (fin || win) {
        const uint64_t prev = seen[rid * W + w] & ~win;
        dst = row * k8 + 64 * w;
        const uint64_t* old = (const uint64_t*)(dist8 + dst);
        for (int q = 0; q < gc; ++q) {
          const uint32_t wq = (uint32_t)(win >> (8 * q)) & 0xffu;
          const uint32_t pq = (uint32_t)(prev >> (8 * q)) & 0xffu;
          if (!fin && !wq) continue;
          uint64_t out = 0, got = 0;
#pragma unroll
          for (int j = 0; j < kMsRingMax; ++j) {
            const uint32_t bq = (uint32_t)(mw[j] >> (8 * q)) & 0xffu;
            if (!bq) continue;
            const uint64_t m = byte_mask8(bq);
            out |= m & ((uint64_t)(uint8_t)(base + (j == 0 ? R : j)) * 0x0101010101010101ull);
            got |= m;
          }
          if (pq) {
            const uint64_t m = byte_mask8(pq);
            out |= old[q] & m;
            got |= m;
          }
          sb[threadIdx.x][q] = out | ~got;
          wmask |= 1u << q;
        }
      }
    }
  }
  sdst[threadIdx.x] = dst;
  swm[threadIdx.x] = wmask;
  __syncthreads();
  const int q = threadIdx.x & 7;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int u = i * 32 + (threadIdx.x >> 3);
    if (sdst[u] >= 0 && ((swm[u] >> q) & 1u))
      *(uint64_t*)(dist8 + sdst[u] + 8 * q) = sb[u][q];
  }
}

__global__ void ms_window_next_kernel(int32_t* ctl, int R) {
  if (ctl[1]) return;
  ctl[2] += R;
}
__global__ void ms_level_end_kernel(unsigned int* found, int32_t* ctl, int32_t cap) {
  if (ctl[1]) return;
  if (*found == 0u) {
    ctl[1] = 1;
    return;
  }
  *found = 0u;
  ctl[0] += 1;
  if (cap && ctl[0] >= cap) ctl[1] = 2;
}

// ------------------------------------------------------- top-down step
// Push when the frontier's out-edges are below 1/alpha of the matrix.
// A pull level whose front is complete (every word written by the level
// before) and whose frontier has more than 1/kDenseFrac of the edges skips
// the front_any test.  A pull writes every word of its next; a push writes
// the rows it touched, and a wide one (more than 1/kWidePush of the edges,
// likely to be followed by a dense level) clears the rest first.
constexpr double kDenseFrac = 8.0;
constexpr double kWidePush = 64.0;

// The first kernel of a level with a direction switch (one thread per word of
// front_any): clears the level's next_any word and sums the out-degrees of its
// frontier rows.  Each block adds (1 << 40) + its sum to p.scout with one
// atomic, so bits 0..39 sum the degrees (at most the entries) and bits 40..63
// count the blocks; the block whose add completes the count holds the total
// (no fence needed), decides the direction and resets the level's counters
// (queue lengths and p.scout) for the kernels after it.  p.scout is zero when
// a level begins (zeroed with the push buffers, and reset here).
constexpr int kMsSumBits = 40;

__global__ __launch_bounds__(256) void ms_begin_kernel(MsArgs a, MsPush p, SellView g, int csr,
                                                       const int32_t* __restrict__ rowptr) {
  if (a.ctl && a.ctl[1]) return;
  __shared__ unsigned long long ws[4];
  const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
  uint32_t bits = 0u;
  if (w < p.words) {
    bits = a.front_any[w];
    a.next_any[w] = 0u;
  }
  unsigned long long deg = 0;
  while (bits) {
    const int b = __ffs(bits) - 1;
    bits &= bits - 1;
    deg += (unsigned long long)state_row(g, csr, rowptr, w * 32 + b).len;
  }
  for (int o = 32; o > 0; o >>= 1) deg += __shfl_xor(deg, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = deg;
  __syncthreads();
  if (threadIdx.x != 0) return;
  const unsigned long long t = ws[0] + ws[1] + ws[2] + ws[3];
  const unsigned long long old = atomicAdd(p.scout, (1ull << kMsSumBits) + t);
  if ((old >> kMsSumBits) != (unsigned long long)gridDim.x - 1) return;
  const double e = (double)((old & ((1ull << kMsSumBits) - 1)) + t);
  const unsigned push = p.alpha > 0.0 && e * p.alpha < (double)p.nnz ? 1u : 0u;
  p.cnt[3] = push;
  p.cnt[4] = push;  // read by the pull kernels as a.pull_off
  p.cnt[5] = !push && p.cnt[6] && e * kDenseFrac > (double)p.nnz ? 1u : 0u;
  p.cnt[6] = !push || e * kWidePush > (double)p.nnz;
  p.cnt[0] = p.cnt[1] = p.cnt[2] = 0u;
  *p.scout = 0ull;
}

// On a push level: the queue of the frontier's state ids (one pass over
// front_any), and for a wide push the clearing of next (the push then writes
// the rows it touches).
__global__ __launch_bounds__(256) void ms_queue_kernel(MsArgs a, MsPush p, int64_t next_words) {
  if (a.ctl && a.ctl[1]) return;
  if (!p.cnt[3]) return;
  const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
  uint32_t bits = w < p.words ? a.front_any[w] : 0u;
  unsigned pos = block_claim((unsigned)__popc(bits), &p.cnt[0]);
  while (bits) {
    const int b = __ffs(bits) - 1;
    bits &= bits - 1;
    p.queue[pos++] = (int32_t)(w * 32 + b);
  }
  if (p.cnt[6])
    for (int64_t i = w; i < next_words; i += (int64_t)gridDim.x * 256) a.next[i] = 0ull;
}

// Edge u -> v: OR u's frontier bits that v has not seen into acc[v]; the
// first lane to touch v queues it.  With one word per vertex that lane is the
// one whose OR found acc[v] zero (acc is zero outside a push step and every
// OR adds a bit); with more words the touched bitmap decides.  All 256
// threads of the block must call it.
template <int W>
__device__ __forceinline__ void push_pair(const MsArgs& a, const MsPush& p, int64_t u, int64_t v,
                                          bool valid, bool wide) {
  bool newly = false;
  if (valid) {
    uint64_t add[W];
    bool any = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      add[w] = a.front[u * W + w] & ~a.seen[v * W + w] & a.live[w];
      any |= add[w] != 0;
    }
    if (any && wide) {
      // a wide push finishes every row afterwards: no claim, and ORs whose
      // old value nobody reads
#pragma unroll
      for (int w = 0; w < W; ++w)
        if (add[w]) __hip_atomic_fetch_or(&p.acc[v * W + w], (unsigned long long)add[w],
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (any) {
      if constexpr (W == 1) {
        newly = atomicOr(&p.acc[v], (unsigned long long)add[0]) == 0ull;
      } else {
#pragma unroll
        for (int w = 0; w < W; ++w)
          if (add[w]) atomicOr(&p.acc[v * W + w], (unsigned long long)add[w]);
        newly = bit_claim(p.touched, v);
      }
    }
  }
  block_append(p.tqueue, &p.cnt[1], v, newly);  // called block-synchronously by expand_*
}

// Frontier rows expanded edge by edge (expand_tiles / expand_pieces).
template <int W>
__global__ __launch_bounds__(256) void ms_push_tile_kernel(MsArgs a, MsPush p, SellView g, int csr,
                                                           const int32_t* __restrict__ rowptr) {
  if (a.ctl && a.ctl[1]) return;
  if (!p.cnt[3]) return;
  const int64_t qlen = p.cnt[0];
  const bool wide = p.cnt[6] != 0;
  expand_tiles(p.queue, qlen, tile_for(qlen), g, csr, rowptr, kPushBig, p.big, &p.cnt[2],
               p.big_cap,
               [&](int64_t u, int64_t v, bool valid) { push_pair<W>(a, p, u, v, valid, wide); });
}

template <int W>
__global__ __launch_bounds__(256) void ms_push_big_kernel(MsArgs a, MsPush p, SellView g, int csr,
                                                          const int32_t* __restrict__ rowptr) {
  if (a.ctl && a.ctl[1]) return;
  if (!p.cnt[3]) return;
  const bool wide = p.cnt[6] != 0;
  expand_pieces(p.big, min((int64_t)p.cnt[2], p.big_cap), g, csr, rowptr, kPushBig,
                [&](int64_t u, int64_t v, bool valid) { push_pair<W>(a, p, u, v, valid, wide); });
}

// Touched rows: new bits = pushed bits not yet seen; acc is cleared for the
// next push step.  A wide push queued no rows: every state row is visited
// and those with pushed bits finished (next was cleared before the push).
template <int W>
__global__ __launch_bounds__(256) void ms_push_finish_kernel(MsArgs a, MsPush p, SellView g,
                                                             int csr, int64_t ns) {
  if (a.ctl && a.ctl[1]) return;
  if (!p.cnt[3]) return;
  const int32_t level = a.ctl ? a.ctl[0] : a.level;
  const bool wide = p.cnt[6] != 0;
  const int64_t n = wide ? ns : (int64_t)p.cnt[1];
  for (int64_t i0 = (int64_t)blockIdx.x * 256; i0 < n; i0 += (int64_t)gridDim.x * 256) {
    const int64_t i = i0 + threadIdx.x;
    bool newly = false;
    uint64_t nw[W];
#pragma unroll
    for (int w = 0; w < W; ++w) nw[w] = 0;
    if (i < n) {
      const int64_t v = wide ? i : p.tqueue[i];
      uint64_t acc[W], any = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        acc[w] = p.acc[v * W + w];
        any |= acc[w];
      }
      if (any) {
        uint64_t sn[W];
#pragma unroll
        for (int w = 0; w < W; ++w) {
          p.acc[v * W + w] = 0ull;
          sn[w] = a.seen[v * W + w];
        }
        newly = finish_row<W>(a, level, v, csr ? v : orig_of(g, v), acc, sn, nw);
      }
    }
    wave_flag(a.found, newly);
    alive_wave<W>(a, nw);
  }
}

int msbfs_level(hipStream_t stream, const gg_matrix* A, const MsArgs& a_in, int Wc,
                const MsPush* push, int ctx_cus) {
  const bool unit = (A->flags & GG_FLAG_UNIT_PATTERN) != 0;
  const int64_t m = A->m;
  MsArgs a = a_in;
  if (push) {
    if (!ms_push_ok(A)) return set_error(GG_ERR_ARGUMENT, "top-down steps need a symmetric unit pattern");
    a.pull_off = (const int32_t*)&push->cnt[4];
    a.dense = (const int32_t*)&push->cnt[5];
    const SellView g = sell_view(A);
    const int csr = A->format == GG_CSR;
    const unsigned nbw = (unsigned)blocks_for(push->words, 256);
    hipLaunchKernelGGL(ms_begin_kernel, dim3(nbw), dim3(256), 0, stream, a, *push, g, csr, A->ptr);
    hipLaunchKernelGGL(ms_queue_kernel, dim3(nbw), dim3(256), 0, stream, a, *push,
                       state_size(A) * Wc);
  } else {
    const int64_t words = (state_size(A) + 31) / 32;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks_for(words, 256)), dim3(256), 0, stream,
                       a.next_any, words, 0u);
  }
  int st = with_type(A->val_dtype, [&](auto ta) -> int {
    typedef typename decltype(ta)::type TA;
    auto go = [&](auto ut, auto wt) -> int {
      constexpr bool U = decltype(ut)::value;
      constexpr int WW = decltype(wt)::value;
      if (A->format == GG_CSR) {
        hipLaunchKernelGGL((msbfs_csr_kernel<TA, U, WW>), dim3(blocks_for(m, kM)), dim3(kM), 0,
                           stream, A->ptr, A->idx, (const TA*)A->val, a);
      } else {
        SellView g = sell_view(A);
        const int64_t nsb = (g.n_slots + kM - 1) / kM;
        if (g.n_heavy > 0 && (!a.hacc || !a.hcnt))
          return set_error(GG_ERR_ARGUMENT, "heavy rows need segment accumulators");
        hipLaunchKernelGGL((msbfs_sell_kernel<TA, U, WW>), dim3((unsigned)(nsb + g.n_seg)),
                           dim3(kM), 0, stream, g, nsb, a);
      }
      GG_HIP(hipGetLastError());
      return GG_OK;
    };
    auto pick_w = [&](auto ut) -> int {
      switch (Wc) {
        case 1: return go(ut, std::integral_constant<int, 1>{});
        case 2: return go(ut, std::integral_constant<int, 2>{});
        case 4: return go(ut, std::integral_constant<int, 4>{});
        default: return go(ut, std::integral_constant<int, 8>{});
      }
    };
    if (unit) return pick_w(std::true_type{});
    if constexpr (std::is_floating_point<TA>::value) {
      return set_error(GG_ERR_ARGUMENT, "or-and BFS needs Bool or integer matrix values");
    } else {
      return pick_w(std::false_type{});
    }
  });
  GG_TRY(st);
  if (push) {
    const SellView g = sell_view(A);
    const int csr = A->format == GG_CSR;
    // grid-stride over at most the state rows (queue, touched rows, or all
    // rows of a wide push): a small graph needs no more blocks than its rows
    // fill, and these kernels run (returning at once) on pull levels too
    const unsigned nb = (unsigned)std::min<int64_t>((int64_t)std::max(ctx_cus, 1) * 8,
                                                    std::max<int64_t>(1, blocks_for(state_size(A), 256)));
    auto run = [&](auto wt) -> int {
      constexpr int WW = decltype(wt)::value;
      hipLaunchKernelGGL(ms_push_tile_kernel<WW>, dim3(nb), dim3(256), 0, stream, a, *push, g, csr,
                         A->ptr);
      hipLaunchKernelGGL(ms_push_big_kernel<WW>, dim3(nb), dim3(256), 0, stream, a, *push, g, csr,
                         A->ptr);
      hipLaunchKernelGGL(ms_push_finish_kernel<WW>, dim3(nb), dim3(256), 0, stream, a, *push, g,
                         csr, state_size(A));
      return GG_OK;
    };
    switch (Wc) {
      case 1: GG_TRY(run(std::integral_constant<int, 1>{})); break;
      case 2: GG_TRY(run(std::integral_constant<int, 2>{})); break;
      case 4: GG_TRY(run(std::integral_constant<int, 4>{})); break;
      default: GG_TRY(run(std::integral_constant<int, 8>{})); break;
    }
    if (Wc > 1)  // one word per vertex claims rows without the touched bitmap
      hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(push->words, 256)), dim3(256), 0, stream,
                         push->touched, push->words, 0u);
    GG_HIP(hipGetLastError());
  }
  return GG_OK;
}

// Runs one batch of k <= 512 sources.  For the column-major result the
// distances go through a row-major byte scratch, where a vertex's k entries
// share lines (a quarter of the traffic of Int32) while the levels stay below
// 255, then are widened and transposed.  The levels do not write the bytes:
// each level's next words stay in a ring of R slots and a flush after every R
// levels (ms_flush_kernel) turns a whole window into bytes with 64-byte
// writes, where per-level writes were scattered read-modify-writes of
// 8-byte words.  A traversal that reaches level 255 is run again with Int32
// levels written by the levels themselves.

static int ms_ring_slots(int U) {
  for (int r = std::min(U, kMsRingMax); r >= 2; --r)
    if (U % r == 0) return r;
  return 2;
}

static int msbfs_once(gg_ctx* ctx, const gg_matrix* A, const int32_t* sources_dev, int64_t k,
                      int32_t* dist_out, int layout, bool narrow, bool* overflow,
                      int64_t* levels_out) {
  const int64_t m = A->m;
  const int64_t k8 = (k + 7) & ~(int64_t)7;
  int32_t* dist = dist_out;
  uint8_t* dist8 = nullptr;
  *overflow = false;
  if (layout == GG_COL_MAJOR) {
    void* p;
    GG_TRY(ws_get(ctx, WS_MSDIST, narrow ? m * k8 + 8 : sizeof(int32_t) * m * k, &p));
    if (narrow) dist8 = (uint8_t*)p;
    else dist = (int32_t*)p;
  }
  const int64_t ns = state_size(A);  // rows of the traversal state
  int W = (int)((k + 63) / 64);
  int Wc = W <= 1 ? 1 : W <= 2 ? 2 : W <= 4 ? 4 : 8;
  const int64_t words = (ns + 31) / 32;
  const int U = ctx->graph_steps > 0 ? (ctx->graph_steps + 1) & ~1 : 0;  // levels per launch
  const int R = !dist8 ? 0 : U > 0 ? ms_ring_slots(U) : kMsRingHost;    // ring slots (0: none)
  const int nbuf = R > 0 ? R : 1;
  void *f, *nx = nullptr, *sn, *fa, *na = nullptr, *cnt, *hv;
  GG_TRY(ws_get(ctx, WS_MS0, sizeof(uint64_t) * ns * Wc * nbuf, &f));
  if (!R) GG_TRY(ws_get(ctx, WS_MS1, sizeof(uint64_t) * ns * Wc, &nx));
  GG_TRY(ws_get(ctx, WS_MS2, sizeof(uint64_t) * ns * Wc, &sn));
  GG_TRY(ws_get(ctx, WS_BFS1, sizeof(uint32_t) * words * nbuf, &fa));
  if (!R) GG_TRY(ws_get(ctx, WS_BFS2, sizeof(uint32_t) * words, &na));
  GG_TRY(ws_get(ctx, WS_FLAG, 64, &cnt));
  // The heavy-row accumulators and the push accumulators / touched bitmap are
  // left zero by every level; they are cleared here because their layout
  // depends on the matrix and on W (a slot reused with another layout holds
  // other data at those places).
  const size_t hb = ms_heavy_bytes(A, Wc);
  GG_TRY(ws_get(ctx, WS_MSHEAVY, hb, &hv));
  // one zeroing launch: the heavy and push accumulators, the push counters,
  // the first front (ring slot 0), seen and the first front_any
  ZeroRanges z{};
  z.add(hv, hb / 4);
  MsPush push{};
  const MsPush* pp = nullptr;
  if (ms_push_ok(A)) {
    void* pb;
    GG_TRY(ws_get(ctx, WS_MSPUSH, ms_push_bytes(A, Wc), &pb));
    push = ms_push_layout(pb, A, Wc);
    push.alpha = ctx->push_alpha;
    z.add(push.acc, 2 * ns * Wc);
    z.add(push.touched, push.words);
    z.add(push.scout, 2);
    z.add(push.cnt, 8);
    pp = &push;
  }
  z.add(f, 2 * ns * Wc);
  z.add(sn, 2 * ns * Wc);
  z.add(fa, words);
  hipLaunchKernelGGL(zero_ranges_kernel, dim3((unsigned)std::max(ctx->num_cus, 1) * 8), dim3(256),
                     0, ctx->stream, z);
  if (dist8) {
    // no fill: the last flush writes every byte
  } else {
    hipLaunchKernelGGL(fill_i32v_kernel, dim3(blocks_for(m * k, kM)), dim3(kM), 0, ctx->stream,
                       dist, m * k, (int32_t)2147483647);
  }
  const int32_t* vinv = (A->flags & GG_FLAG_INTERNAL_IDS) ? A->vinv : nullptr;
  int32_t* ctl = U > 0 ? (int32_t*)cnt + 4 : nullptr;  // captured levels: {level, done, window}
  hipLaunchKernelGGL(msbfs_seed_kernel, dim3(blocks_for(k, kM)), dim3(kM), 0, ctx->stream,
                     sources_dev, k, k8, Wc, vinv, (uint64_t*)f, (uint32_t*)fa, (uint64_t*)sn,
                     dist, dist8, (unsigned*)cnt, ctl);
  GG_HIP(hipGetLastError());
  MsArgs a{};
  a.m = m;
  a.k = k;
  a.layout = GG_ROW_MAJOR;
  a.seen = (uint64_t*)sn;
  a.dist = R ? nullptr : dist;
  a.dist8 = R ? nullptr : dist8;
  a.k8 = k8;
  a.found = (unsigned*)cnt;
  for (int w = 0; w < 8; ++w) a.live[w] = w < Wc ? word_mask(k, w) : 0ull;
  a.start = nullptr;
  a.alive = nullptr;
  a.hacc = (unsigned long long*)hv;
  a.hcnt = (unsigned*)((unsigned long long*)hv + (A->format == GG_SELL ? A->n_heavy : 0) * Wc);
  // front / next of level L: two buffers alternating, or ring slots
  // (L - 1) % R and L % R
  uint64_t* mk[2] = {(uint64_t*)f, (uint64_t*)nx};
  uint32_t* ak[2] = {(uint32_t*)fa, (uint32_t*)na};
  auto mask_of = [&](int32_t L) -> uint64_t* {
    return R ? (uint64_t*)f + (int64_t)(L % R) * ns * Wc : mk[L & 1];
  };
  auto any_of = [&](int32_t L) -> uint32_t* {
    return R ? (uint32_t*)fa + (int64_t)(L % R) * words : ak[L & 1];
  };
  const SellView gv = sell_view(A);
  const unsigned flush_blocks = (unsigned)blocks_for(ns * Wc, 256);
  auto flush = [&](hipStream_t st, const int32_t* ctl, int32_t base, int32_t last, int fin) {
    hipLaunchKernelGGL(ms_flush_kernel, dim3(flush_blocks), dim3(256), 0, st, gv,
                       (const uint64_t*)f, (const uint32_t*)fa, R, ns, words, Wc, k8,
                       (const uint64_t*)sn, dist8, ctl, base, last, fin);
  };
  int32_t level = 0;
  if (U > 0) {
    // U levels per graph launch; step i runs level (launch) * U + i + 1 (the
    // seed kernel started the counters)
    a.ctl = ctl;
    std::vector<MsArgs> ab(U, a);
    for (int i = 0; i < U; ++i) {
      ab[i].front = mask_of(i);
      ab[i].front_any = any_of(i);
      ab[i].next = mask_of(i + 1);
      ab[i].next_any = any_of(i + 1);
    }
    std::vector<uint64_t> key{2, (uint64_t)U, (uint64_t)Wc, (uint64_t)words, (uint64_t)R,
                              (uint64_t)(uintptr_t)dist8, (uint64_t)k8};
    sig_add(key, *A);
    for (int i = 0; i < U; ++i) sig_add(key, ab[i]);
    sig_add(key, push);
    auto steps = [&](hipStream_t st) -> int {
      for (int i = 0; i < U; ++i) {
        const MsArgs& x = ab[i];
        GG_TRY(msbfs_level(st, A, x, Wc, pp, ctx->num_cus));
        hipLaunchKernelGGL(ms_level_end_kernel, dim3(1), dim3(1), 0, st, (unsigned*)cnt, ctl,
                           dist8 ? 255 : 0);
        if (R && (i + 1) % R == 0) {
          flush(st, ctl, 0, 0, 0);
          hipLaunchKernelGGL(ms_window_next_kernel, dim3(1), dim3(1), 0, st, ctl, R);
        }
      }
      GG_HIP(hipGetLastError());
      return GG_OK;
    };
    // the loop's last check leaves ctl[0..1] (level, done) in host_pinned
    GG_TRY(graph_loop(ctx, key, ctl + 1, m / U + 3, steps, nullptr, ctl, 2));
    const int32_t* hctl = (const int32_t*)ctx->host_pinned;
    if (hctl[1] == 2) {
      *overflow = true;
      return GG_OK;
    }
    level = hctl[0];
  } else {
    while (true) {
      ++level;
      if (dist8 && level == 255) {  // byte levels are full
        *overflow = true;
        return GG_OK;
      }
      a.front = mask_of(level - 1);
      a.front_any = any_of(level - 1);
      a.next = mask_of(level);
      a.next_any = any_of(level);
      a.level = level;
      hipLaunchKernelGGL(ms_zero_kernel, dim3(1), dim3(1), 0, ctx->stream, (unsigned*)cnt);
      {
        HotTimer ht(ctx);
        GG_TRY(msbfs_level(ctx->stream, A, a, Wc, pp, ctx->num_cus));
      }
      GG_HIP(hipMemcpyAsync(ctx->host_pinned, cnt, sizeof(unsigned), hipMemcpyDeviceToHost,
                            ctx->stream));
      GG_HIP(hipStreamSynchronize(ctx->stream));
      if (*(unsigned*)ctx->host_pinned == 0) break;
      if (R && level % R == 0) {
        HotTimer ht(ctx, true, false);  // the levels' share of recording their distances
        flush(ctx->stream, nullptr, level - R, level, 0);
      }
    }
    if (R) {
      HotTimer ht(ctx, true, false);
      flush(ctx->stream, nullptr, (level - 1) / R * R, level, 1);
    }
  }
  if (layout == GG_COL_MAJOR && m > 0) {
    const int64_t tiles = ((m + 63) / 64) * ((k + 63) / 64);
    if (dist8)
      hipLaunchKernelGGL(widen_transpose_kernel, dim3((unsigned)tiles), dim3(256), 0,
                         ctx->stream, dist8, m, k, k8,
                         (int)(m % 4 == 0 && ((uintptr_t)dist_out & 15) == 0), dist_out);
    else
      hipLaunchKernelGGL(transpose_rm_cm_kernel, dim3((unsigned)tiles), dim3(256), 0,
                         ctx->stream, dist, m, k, dist_out);
    GG_HIP(hipGetLastError());
  }
  if (levels_out) *levels_out = level - 1;
  return GG_OK;
}

int msbfs_run(gg_ctx* ctx, const gg_matrix* A, const int32_t* sources_dev, int64_t k,
              int32_t* dist_out, int layout, int64_t* levels_out) {
  bool overflow = false;
  GG_TRY(msbfs_once(ctx, A, sources_dev, k, dist_out, layout, true, &overflow, levels_out));
  if (overflow)
    GG_TRY(msbfs_once(ctx, A, sources_dev, k, dist_out, layout, false, &overflow, levels_out));
  return GG_OK;
}

}  // namespace gg

using namespace gg;

extern "C" int gg_msbfs_distances(gg_ctx* ctx, const gg_matrix* A_T, const int32_t* sources,
                                  int64_t k, int32_t* dist, int64_t* levels) {
  if (!ctx || !dist || (k > 0 && !sources)) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  GG_TRY(check_matrix(A_T));
  if (A_T->format == GG_CSC) return set_error(GG_ERR_ARGUMENT, "BFS runs on CSR or SELL A_T");
  if (A_T->m != A_T->n) return set_error(GG_ERR_DIMENSION, "A_T must be square");
  if (k < 1) return set_error(GG_ERR_ARGUMENT, "need at least one source");
  for (int64_t i = 0; i < k; ++i)
    if (sources[i] < 1 || sources[i] > A_T->m)
      return set_error(GG_ERR_BOUNDS, "source %lld outside 1:%lld", (long long)sources[i],
                       (long long)A_T->m);
  timer_begin(ctx);
  void* src;
  GG_TRY(ws_get(ctx, WS_MS3, sizeof(int32_t) * k, &src));
  if (k <= kMsPinnedSources) {
    // staged in the pinned page's upper half (its lower half takes the
    // readbacks): an asynchronous copy; the call synchronises before it
    // returns, so the page is free again for the next call
    int32_t* stage = (int32_t*)((char*)ctx->host_pinned + 2048);
    std::memcpy(stage, sources, sizeof(int32_t) * k);
    GG_HIP(hipMemcpyAsync(src, stage, sizeof(int32_t) * k, hipMemcpyHostToDevice, ctx->stream));
  } else {
    GG_HIP(hipMemcpyAsync(src, sources, sizeof(int32_t) * k, hipMemcpyHostToDevice, ctx->stream));
  }
  // column-major m x k: batches of <= 512 sources write disjoint column blocks
  int64_t maxlev = 0;
  for (int64_t b0 = 0; b0 < k; b0 += 512) {
    const int64_t kb = std::min<int64_t>(512, k - b0);
    int64_t lv = 0;
    GG_TRY(msbfs_run(ctx, A_T, (const int32_t*)src + b0, kb, dist + b0 * A_T->m, GG_COL_MAJOR, &lv));
    maxlev = std::max(maxlev, lv);
  }
  if (levels) *levels = maxlev;
  timer_end(ctx);
  return GG_OK;
}
