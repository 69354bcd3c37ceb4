// sssp.hip -- Bellman-Ford (min, safe_add) from one or many sources.
//
// src/algorithms/shortest_path.jl:34-83 iterates next = min(next, A_T
// (min.safe_add) dist), dist .= next, until an iteration changes nothing.  Its
// fixpoint is unique: safe_add is monotone in the distance, so every fair
// relaxation order reaches the least solution of d[v] = min(d[v], min_u
// safe_add(w(u->v), d[u])), the minimum over paths of the left-fold sums from
// the source (exact for the integer-valued weights of the tests).
//
// Here one kernel per iteration relaxes every row in place and relaxes a row
// only from the in-neighbours whose distance changed in the previous iteration
// (a bitmap, L2-resident): an unchanged u cannot lower d[v] again, because v
// already read w + d[u] after u's last change.  Rows that fall set their bit
// for the next iteration; an iteration in which nothing falls is the fixpoint.
// The distances equal the reference's; the iteration count can be lower (in
// place, a change is seen within the same iteration when its row runs later).
//
// k sources (src/algorithms/shortest_path.jl:112-161) use a row-major working
// set (state rows x k): a group of lanes owns one row, one lane per source, so
// the k distances of a neighbour are one contiguous read; the reference's
// column-major result is written at the end.
#include "batched.hpp"

namespace gg {

constexpr int kT = 256;

template <class TD> __device__ __forceinline__ TD sadd(TD w, TD d) {
  const TD sat = Lim<TD>::max();
  return jmax(w, d) == sat ? sat : jadd(w, d);  // GPUGraphs_safe_add, src/algebra.jl:47-48
}


// ----------------------------------------------------------- one source, CSR
template <class TA, class TD>
__global__ __launch_bounds__(kT) void bf_csr_kernel(const int32_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ colval,
                                                    const TA* __restrict__ val, int64_t m,
                                                    int logL, TD* d,
                                                    const uint32_t* __restrict__ chp,
                                                    uint32_t* __restrict__ chn,
                                                    unsigned int* changed, const int32_t* done) {
  if (done && *done) return;  // a captured step after the fixpoint
  const int L = 1 << logL;
  const int wl = threadIdx.x & 63;
  const int lane = wl & (L - 1);
  const int64_t gpw = 64 >> logL;
  const int64_t wave = ((int64_t)blockIdx.x * kT + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kT) >> 6;
  for (int64_t w0 = wave * gpw; w0 < m; w0 += nwaves * gpw) {
    const int64_t row = w0 + (wl >> logL);
    TD acc = Lim<TD>::max();
    if (row < m) {
      for (int32_t i = rowptr[row] - 1 + lane; i < rowptr[row + 1] - 1; i += L) {
        const int32_t c = colval[i] - 1;
        if (bit_get(chp, c)) acc = jmin(acc, sadd(TD(val[i]), d[c]));
      }
    }
    for (int o = L >> 1; o > 0; o >>= 1) acc = jmin(acc, shfl_xor(acc, o));
    bool ch = false;
    if (row < m && lane == 0) {
      const TD dv = d[row];
      const TD nv = jmin(dv, acc);
      if (nv < dv) {
        d[row] = nv;
        bit_claim(chn, row);
        ch = true;
      }
    }
    wave_flag(changed, ch);
  }
}

// ---------------------------------------------------------- one source, SELL
// Light slots: one lane per row, entries in batches of 4 (index loads, bitmap
// tests, then the value / distance loads of the changed neighbours).  Heavy
// rows: one block per segment, folded into the row with an atomic min.
template <class TA, class TD>
__global__ __launch_bounds__(kT) void bf_sell_kernel(SellView g, int64_t n_sell_blocks, TD* d,
                                                     const uint32_t* __restrict__ chp,
                                                     uint32_t* __restrict__ chn,
                                                     unsigned int* changed, const int32_t* done) {
  if (done && *done) return;
  if ((int64_t)blockIdx.x < n_sell_blocks) {
    const int64_t slot = (int64_t)blockIdx.x * kT + threadIdx.x;
    int64_t row, base, len;
    bool has_len;
    bool ch = false;
    if (sell_slot(g, slot, row, base, len, has_len)) {
      const TA* val = (const TA*)g.val;
      TD acc = Lim<TD>::max();
      int64_t k = 0;
      if (has_len) {
        for (; k + 4 <= len; k += 4) {
          int32_t c[4];
          bool on[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) c[u] = g.colval[base + (k + u) * g.C] - 1;
#pragma unroll
          for (int u = 0; u < 4; ++u) on[u] = bit_get(chp, c[u]);
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (on[u]) acc = jmin(acc, sadd(TD(val[base + (k + u) * g.C]), d[c[u]]));
        }
      }
      for (; k < len; ++k) {
        const int64_t i = base + k * g.C;
        const int32_t col = g.colval[i];
        if (!has_len && col == -1) break;
        if (bit_get(chp, col - 1)) acc = jmin(acc, sadd(TD(val[i]), d[col - 1]));
      }
      const int64_t rid = g.internal ? slot : row;
      // no candidate below the unreached value: the row cannot fall, and its
      // own distance is not read
      if (acc < Lim<TD>::max()) {
        const TD dv = d[rid];
        const TD nv = jmin(dv, acc);
        if (nv < dv) {
          d[rid] = nv;
          if (!g.internal) bit_claim(chn, rid);
          ch = true;
        }
      }
    }
    if (g.internal) wave_set_bits(chn, slot & ~(int64_t)63, ch);  // the wavefront's two words
    wave_flag(changed, ch);
    return;
  }
  __shared__ TD red[kT / 64];
  int64_t h, s, e;
  heavy_segment(g, (int64_t)blockIdx.x - n_sell_blocks, h, s, e);
  const TA* hv = (const TA*)g.heavy_val;
  TD acc = Lim<TD>::max();
  for (int64_t i = s + threadIdx.x; i < e; i += kT) {
    const int32_t c = g.heavy_col[i] - 1;
    if (bit_get(chp, c)) acc = jmin(acc, sadd(TD(hv[i]), d[c]));
  }
  for (int o = 32; o > 0; o >>= 1) acc = jmin(acc, shfl_xor(acc, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kT / 64; ++w) acc = jmin(acc, red[w]);
    const int64_t rid = heavy_rid(g, h);
    if (atomic_min_lowered(&d[rid], acc)) {
      bit_claim(chn, rid);
      set_flag(changed);
    }
  }
}

// -------------------------------------------------- k sources, row-major set
// min over the changed in-neighbours c of a SELL slot of safe_add(w, d[c, j]):
// entries in batches of 4 (index loads, bitmap tests, then the value and
// distance loads), so a lane keeps four independent loads in flight.
template <class TA, class TD>
__device__ __forceinline__ TD multi_row_min(const SellView& g, const TA* __restrict__ val,
                                            int64_t base, int64_t len, bool has_len,
                                            const uint32_t* __restrict__ chp, const TD* d,
                                            int64_t k, int64_t ld, int64_t j) {
  TD acc = Lim<TD>::max();
  const bool jin = j < k;
  int64_t kk = 0;
  if (has_len) {
    for (; kk + 4 <= len; kk += 4) {
      int32_t c[4];
      bool on[4];
      TA w[4];
      TD dn[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) c[u] = g.colval[base + (kk + u) * g.C] - 1;
#pragma unroll
      for (int u = 0; u < 4; ++u) on[u] = jin && bit_get(chp, c[u]);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (on[u]) {
          w[u] = val[base + (kk + u) * g.C];
          dn[u] = d[(int64_t)c[u] * ld + j];
        }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (on[u]) acc = jmin(acc, sadd(TD(w[u]), dn[u]));
    }
  }
  for (; kk < len; ++kk) {
    const int64_t i = base + kk * g.C;
    const int32_t col = g.colval[i];
    if (!has_len && col == -1) break;
    if (jin && bit_get(chp, col - 1)) acc = jmin(acc, sadd(TD(val[i]), d[(int64_t)(col - 1) * ld + j]));
  }
  return acc;
}

// Rows of A_T as (entry list) for either format: CSR row, SELL slot, or a
// heavy row.  A group of Lk lanes (one per source, Lk >= k when k <= 64)
// walks the row; the bitmap test is uniform across a group (one column).
// Heavy rows are cut into parts of at most 256 entries (each heavy segment in
// `parts` pieces); a block per (part, column block) splits its entries over
// its groups and folds the result into the row with per-column atomic mins.
template <class TA, class TD>
__global__ __launch_bounds__(kT) void bf_multi_kernel(int format, const int32_t* __restrict__ rowptr,
                                                      SellView g, int64_t n_light_groups,
                                                      int64_t parts, int64_t k, int64_t ld, int logLk, TD* d,
                                                      const uint32_t* __restrict__ chp,
                                                      uint32_t* __restrict__ chn,
                                                      unsigned int* changed,
                                                      unsigned long long* changed_mask,
                                                      const int32_t* done,
                                                      const unsigned int* light_off) {
  if (done && *done) return;
  const int Lk = 1 << logLk;
  const int wl = threadIdx.x & 63;
  const int lane = wl & (Lk - 1);
  const int gbase = wl & ~(Lk - 1);
  const int64_t jblocks = (k + Lk - 1) >> logLk;
  const int64_t gpb = kT >> logLk;  // groups per block
  const int64_t light_blocks = (n_light_groups * jblocks + gpb - 1) / gpb;
  if ((int64_t)blockIdx.x < light_blocks) {
    if (light_off && *light_off) return;
    const int64_t q = (int64_t)blockIdx.x * gpb + (threadIdx.x >> logLk);
    const int64_t unit = q / jblocks;
    const int64_t j = (q - unit * jblocks) * Lk + lane;
    bool ch = false;
    int64_t rid = -1;
    if (unit < n_light_groups) {
      TD acc = Lim<TD>::max();
      const TA* val = (const TA*)g.val;
      if (format == GG_CSR) {
        rid = unit;
        for (int32_t i = rowptr[rid] - 1; i < rowptr[rid + 1] - 1; ++i) {
          const int32_t c = g.colval[i] - 1;
          if (bit_get(chp, c) && j < k) acc = jmin(acc, sadd(TD(val[i]), d[(int64_t)c * ld + j]));
        }
      } else {
        int64_t row, base, len;
        bool has_len;
        if (sell_slot(g, unit, row, base, len, has_len)) {
          rid = g.internal ? unit : row;
          acc = multi_row_min<TA, TD>(g, val, base, len, has_len, chp, d, k, ld, j);
        }
      }
      if (rid >= 0 && j < k && acc < Lim<TD>::max()) {
        const TD dv = d[rid * ld + j];
        const TD nv = jmin(dv, acc);
        if (nv < dv) {
          d[rid * ld + j] = nv;
          ch = true;
        }
      }
    }
    wave_flag(changed, ch);
    const unsigned long long b = __ballot(ch);
    const unsigned long long low = Lk == 64 ? ~0ull : ((1ull << Lk) - 1ull);
    if (lane == 0 && ((b >> gbase) & low)) bit_claim(chn, rid);  // this group's row fell
    if (changed_mask) {  // k <= 64: lane l of a group is column l
      unsigned long long cm = 0;
      for (int o = 0; o < 64; o += Lk) cm |= (b >> o) & low;
      if (wl == 0 && cm) or_new_bits(changed_mask, cm);
    }
    return;
  }
  // heavy rows: block (h, jblock); the block's groups split the row's entries
  extern __shared__ unsigned char smem[];
  TD* red = (TD*)smem;  // [gpb][Lk]
  __shared__ int s_any;
  const int64_t hb = (int64_t)blockIdx.x - light_blocks;
  const int64_t unit = hb / jblocks;
  const int64_t j = (hb - unit * jblocks) * Lk + lane;
  const int64_t grp = threadIdx.x >> logLk;
  int64_t h, s, e;
  heavy_segment(g, unit / parts, h, s, e);
  const int64_t plen = (g.seg_len + parts - 1) / parts;
  s += (unit % parts) * plen;
  e = min(e, s + plen);
  if (s >= e) return;  // a part past the end of the row's last segment (uniform)
  const TA* hv = (const TA*)g.heavy_val;
  TD acc = Lim<TD>::max();
  if (threadIdx.x == 0) s_any = 0;
  if (j < k)
    for (int64_t i = s + grp; i < e; i += gpb) {
      const int32_t c = g.heavy_col[i] - 1;
      if (bit_get(chp, c)) acc = jmin(acc, sadd(TD(hv[i]), d[(int64_t)c * ld + j]));
    }
  red[grp * Lk + lane] = acc;
  __syncthreads();
  if (grp == 0) {
    for (int64_t o = 1; o < gpb; ++o) acc = jmin(acc, red[o * Lk + lane]);
    if (j < k) {
      const int64_t rid = heavy_rid(g, h);
      if (atomic_min_lowered(&d[rid * ld + j], acc)) {
        set_flag(changed);
        if (changed_mask) or_new_bits(changed_mask, 1ull << j);
        s_any = 1;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_any) bit_claim(chn, heavy_rid(g, h));
}

// Light SELL slots with V columns per lane (V = 16 / sizeof(TD), k a multiple
// of V, d 16-byte aligned): a group of Lv lanes per row, each reading a
// neighbour's V distances as one 16-byte load, so a wavefront covers 64 / Lv
// rows.  Same relaxation as bf_multi_kernel's light part.
template <class TA, class TD, int V>
__device__ __forceinline__ void bf_vec_light(SellView g, int64_t n_light, int64_t k, int64_t ld,
    int logLv, TD* d, const uint32_t* __restrict__ chp, uint32_t* __restrict__ chn,
    unsigned int* changed, unsigned long long* changed_mask, const int32_t* done,
    const unsigned int* light_off, ColGroups cg, int64_t bid) {
  typedef TD VT __attribute__((ext_vector_type(V)));
  if (done && *done) return;
  if (light_off && *light_off) return;
  const int Lv = 1 << logLv;
  const int wl = threadIdx.x & 63;
  const int lane = wl & (Lv - 1);
  const int gbase = wl & ~(Lv - 1);
  const int64_t kq = cg.n ? cg.n : k / V;
  const int64_t jblocks = (kq + Lv - 1) >> logLv;
  const int64_t gpb = kT >> logLv;
  const int64_t q = bid * gpb + (threadIdx.x >> logLv);
  const int64_t unit = q / jblocks;
  const int64_t jq = (q - unit * jblocks) * Lv + lane;
  const bool jin = jq < kq;
  const int64_t j0 = (cg.n && jin ? (int64_t)cg.g[jq] : jq) * V;
  unsigned chbits = 0;
  int64_t rid = -1;
  if (unit < n_light) {
    int64_t row, base, len;
    bool has_len;
    if (sell_slot(g, unit, row, base, len, has_len)) {
      rid = g.internal ? unit : row;
      if (jin) {
        const TA* val = (const TA*)g.val;
        VT acc;
#pragma unroll
        for (int t = 0; t < V; ++t) acc[t] = Lim<TD>::max();
        int64_t kk = 0;
        if (has_len) {
          for (; kk + 4 <= len; kk += 4) {
            int32_t c[4];
            bool on[4];
            TA w[4];
            VT dn[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) c[u] = g.colval[base + (kk + u) * g.C] - 1;
#pragma unroll
            for (int u = 0; u < 4; ++u) on[u] = bit_get(chp, c[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (on[u]) {
                w[u] = val[base + (kk + u) * g.C];
                dn[u] = *(const VT*)&d[(int64_t)c[u] * ld + j0];
              }
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (on[u])
#pragma unroll
                for (int t = 0; t < V; ++t) acc[t] = jmin(acc[t], sadd(TD(w[u]), dn[u][t]));
          }
        }
        for (; kk < len; ++kk) {
          const int64_t i = base + kk * g.C;
          const int32_t col = g.colval[i];
          if (!has_len && col == -1) break;
          if (!bit_get(chp, col - 1)) continue;
          const TD w = TD(val[i]);
          const VT dn = *(const VT*)&d[(int64_t)(col - 1) * ld + j0];
#pragma unroll
          for (int t = 0; t < V; ++t) acc[t] = jmin(acc[t], sadd(w, dn[t]));
        }
        // the own distances are read only when some candidate is below the
        // unreached value (otherwise none of the lane's columns can fall)
        bool cand = false;
#pragma unroll
        for (int t = 0; t < V; ++t) cand |= acc[t] < Lim<TD>::max();
        if (cand) {
          VT own = *(const VT*)&d[rid * ld + j0];
#pragma unroll
          for (int t = 0; t < V; ++t)
            if (acc[t] < own[t]) {
              own[t] = acc[t];
              chbits |= 1u << t;
            }
          if (chbits) *(VT*)&d[rid * ld + j0] = own;
        }
      }
    }
  }
  const bool ch = chbits != 0;
  wave_flag(changed, ch);
  const unsigned long long b = __ballot(ch);
  const unsigned long long low = Lv == 64 ? ~0ull : ((1ull << Lv) - 1ull);
  if (lane == 0 && ((b >> gbase) & low)) bit_claim(chn, rid);  // this group's row fell
  if (changed_mask) {  // k <= 64: bit j of the mask is column j
    unsigned long long cm = ch ? (unsigned long long)chbits << j0 : 0ull;
    for (int o = 32; o > 0; o >>= 1) cm |= __shfl_xor(cm, o, 64);
    if (wl == 0 && cm) or_new_bits(changed_mask, cm);
  }
}

template <class TA, class TD, int V>
__global__ __launch_bounds__(kT) void bf_multi_vec_kernel(SellView g, int64_t n_light, int64_t k,
    int64_t ld, int logLv, TD* d, const uint32_t* __restrict__ chp, uint32_t* __restrict__ chn,
    unsigned int* changed, unsigned long long* changed_mask, const int32_t* done,
    const unsigned int* light_off, ColGroups cg) {
  bf_vec_light<TA, TD, V>(g, n_light, k, ld, logLv, d, chp, chn, changed, changed_mask, done,
                          light_off, cg, (int64_t)blockIdx.x);
}

// Heavy rows with V columns per lane: a block per (part of kHeavyVecPart
// entries of a heavy segment, column block); its groups of Lv lanes split the
// part's entries (batches of 4 per group), fold through LDS and lower the row
// with per-column atomic mins.  Whole segments of the usual lengths per block:
// fewer, longer blocks and fewer atomics (1024-entry parts ran at 0.8x the
// rate of 4096-entry ones on R-MAT 22 with 32 sources).
constexpr int64_t kHeavyVecPart = 8192;

template <class TA, class TD, int V>
__device__ __forceinline__ void bf_vec_heavy(SellView g, int64_t parts, int64_t k, int64_t ld,
    int logLv, TD* d, const uint32_t* __restrict__ chp, uint32_t* __restrict__ chn,
    unsigned int* changed, unsigned long long* changed_mask, const int32_t* done, ColGroups cg,
    int64_t bid) {
  typedef TD VT __attribute__((ext_vector_type(V)));
  __shared__ VT red[kT];
  __shared__ int s_any;
  if (done && *done) return;
  const int Lv = 1 << logLv;
  const int lane = threadIdx.x & (Lv - 1);
  const int64_t grp = threadIdx.x >> logLv;
  const int64_t gpb = kT >> logLv;
  const int64_t kq = cg.n ? cg.n : k / V;
  const int64_t jblocks = (kq + Lv - 1) >> logLv;
  const int64_t unit = bid / jblocks;
  const int64_t jq = (bid - unit * jblocks) * Lv + lane;
  const bool jin = jq < kq;
  const int64_t j0 = (cg.n && jin ? (int64_t)cg.g[jq] : jq) * V;
  int64_t h, s, e;
  heavy_segment(g, unit / parts, h, s, e);
  s += (unit % parts) * kHeavyVecPart;
  e = min(e, s + kHeavyVecPart);
  if (s >= e) return;  // a part past the end of the row's last segment (uniform)
  const TA* hv = (const TA*)g.heavy_val;
  VT acc;
#pragma unroll
  for (int t = 0; t < V; ++t) acc[t] = Lim<TD>::max();
  if (threadIdx.x == 0) s_any = 0;
  if (jin) {
    int64_t i = s + grp;
    for (; i + 3 * gpb < e; i += 4 * gpb) {
      int32_t c[4];
      bool on[4];
      TA w[4];
      VT dn[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) c[u] = g.heavy_col[i + u * gpb] - 1;
#pragma unroll
      for (int u = 0; u < 4; ++u) on[u] = bit_get(chp, c[u]);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (on[u]) {
          w[u] = hv[i + u * gpb];
          dn[u] = *(const VT*)&d[(int64_t)c[u] * ld + j0];
        }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (on[u])
#pragma unroll
          for (int t = 0; t < V; ++t) acc[t] = jmin(acc[t], sadd(TD(w[u]), dn[u][t]));
    }
    for (; i < e; i += gpb) {
      const int32_t c = g.heavy_col[i] - 1;
      if (!bit_get(chp, c)) continue;
      const TD w = TD(hv[i]);
      const VT dn = *(const VT*)&d[(int64_t)c * ld + j0];
#pragma unroll
      for (int t = 0; t < V; ++t) acc[t] = jmin(acc[t], sadd(w, dn[t]));
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (grp == 0) {
    for (int64_t o = 1; o < gpb; ++o) {
      const VT r = red[o * Lv + lane];
#pragma unroll
      for (int t = 0; t < V; ++t) acc[t] = jmin(acc[t], r[t]);
    }
    if (jin) {
      const int64_t rid = heavy_rid(g, h);
      unsigned long long cm = 0;
#pragma unroll
      for (int t = 0; t < V; ++t)
        if (atomic_min_lowered(&d[rid * ld + j0 + t], acc[t])) cm |= 1ull << t;
      if (cm) {
        set_flag(changed);
        if (changed_mask) or_new_bits(changed_mask, cm << j0);
        s_any = 1;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_any) bit_claim(chn, heavy_rid(g, h));
}

template <class TA, class TD, int V>
__global__ __launch_bounds__(kT) void bf_multi_heavy_vec_kernel(SellView g, int64_t parts,
    int64_t k, int64_t ld, int logLv, TD* d, const uint32_t* __restrict__ chp,
    uint32_t* __restrict__ chn, unsigned int* changed, unsigned long long* changed_mask,
    const int32_t* done, ColGroups cg) {
  bf_vec_heavy<TA, TD, V>(g, parts, k, ld, logLv, d, chp, chn, changed, changed_mask, done, cg,
                          (int64_t)blockIdx.x);
}

// Light slots and heavy parts in one launch, the heavy parts first: a heavy
// part is a long chain of dependent batches on few blocks, and in a launch of
// their own they ran after the light slots with the device mostly idle; here
// the light slots' blocks fill the device around them.
template <class TA, class TD, int V>
__global__ __launch_bounds__(kT) void bf_multi_fused_vec_kernel(
    SellView g, int64_t n_light, int64_t parts, int64_t nh, int64_t k, int64_t ld, int logLv,
    TD* d, const uint32_t* __restrict__ chp, uint32_t* __restrict__ chn, unsigned int* changed,
    unsigned long long* changed_mask, const int32_t* done, const unsigned int* light_off,
    ColGroups cg) {
  if ((int64_t)blockIdx.x < nh)
    bf_vec_heavy<TA, TD, V>(g, parts, k, ld, logLv, d, chp, chn, changed, changed_mask, done, cg,
                            (int64_t)blockIdx.x);
  else
    bf_vec_light<TA, TD, V>(g, n_light, k, ld, logLv, d, chp, chn, changed, changed_mask, done,
                            light_off, cg, (int64_t)blockIdx.x - nh);
}

// ----------------------------------------------------- sparse iterations
// On the symmetric pattern of a relabelled SELL, row u of A_T also lists u's
// out-neighbours.  A k-source iteration whose fallen rows have few out-edges
// (edges * alpha < entries) relaxes only the light rows their edges reach -- exactly the rows with
// an in-neighbour that fell -- instead of every light slot, a group of lanes
// per row over its k contiguous distances; the heavy rows keep their segment
// blocks.  A device-side switch decides, so the iteration is the same launches
// either way (and can be captured).  One source keeps the dense sweep: a lone
// row of a SELL slice reads a cache line per entry, and the sweep's prologue
// costs more than the rows it skips.

__global__ void bf_sp_start_kernel(BfSparse s, const int32_t* done) {
  if (done && *done) return;
  s.cnt[0] = s.cnt[1] = s.cnt[2] = 0u;
  *s.scout = 0ull;
}

// Out-edges of the fallen rows (QUEUE = false: their degrees summed; the
// pattern is symmetric, so a row's in-degree is its out-degree), or their
// queue (QUEUE = true, sparse iterations only).
template <bool QUEUE>
__global__ __launch_bounds__(kT) void bf_sp_fallen_kernel(BfSparse s, const uint32_t* __restrict__ chp,
                                                          SellView g, const int32_t* done) {
  if (done && *done) return;
  if (QUEUE && !s.cnt[3]) return;
  int64_t w[kScanWords];
  uint32_t bits[kScanWords];
  unsigned c = 0;
#pragma unroll
  for (int i = 0; i < kScanWords; ++i) {
    w[i] = (int64_t)blockIdx.x * (kT * kScanWords) + i * kT + threadIdx.x;
    bits[i] = w[i] < s.words ? chp[w[i]] : 0u;
    c += __popc(bits[i]);
  }
  if (!QUEUE) {
    unsigned long long deg = 0;
#pragma unroll
    for (int i = 0; i < kScanWords; ++i)
      for (uint32_t b = bits[i]; b; b &= b - 1)
        deg += (unsigned long long)state_row(g, false, nullptr, w[i] * 32 + __ffs(b) - 1).len;
    block_add(deg, s.scout);
    return;
  }
  {
    unsigned pos = block_claim(c, &s.cnt[0]);
#pragma unroll
    for (int i = 0; i < kScanWords; ++i)
      while (bits[i]) {
        const int b = __ffs(bits[i]) - 1;
        bits[i] &= bits[i] - 1;
        s.queue[pos++] = (int32_t)(w[i] * 32 + b);
      }
  }
}

// Sparse when the fallen rows' out-edges times alpha are below the matrix's
// entries (counting their rows alone, at the average degree, took sparse
// iterations for fallen sets that held hubs, where the sweep is cheaper).
__global__ void bf_sp_decide_kernel(BfSparse s, const int32_t* done) {
  if (done && *done) return;
  const unsigned sp = (double)(*s.scout) * s.alpha < (double)s.nnz ? 1u : 0u;
  s.cnt[3] = sp;
  s.cnt[4] = sp;
}

__device__ __forceinline__ void bf_sp_mark(const BfSparse& s, const SellView& g, int64_t v,
                                           bool valid) {
  const bool add = valid && v < g.slot_base && bit_claim(s.mark, v);
  block_append(s.dq, &s.cnt[1], v, add);  // called block-synchronously by expand_*
}

__global__ __launch_bounds__(256) void bf_sp_mark_tile_kernel(BfSparse s, SellView g,
                                                              const int32_t* done) {
  if (done && *done) return;
  if (!s.cnt[3]) return;
  const int64_t qlen = s.cnt[0];
  expand_tiles(s.queue, qlen, tile_for(qlen), g, false, nullptr, kSparseBig, s.big, &s.cnt[2],
               s.big_cap, [&](int64_t, int64_t v, bool valid) { bf_sp_mark(s, g, v, valid); });
}

__global__ __launch_bounds__(256) void bf_sp_mark_big_kernel(BfSparse s, SellView g,
                                                             const int32_t* done) {
  if (done && *done) return;
  if (!s.cnt[3]) return;
  expand_pieces(s.big, min((int64_t)s.cnt[2], s.big_cap), g, false, nullptr, kSparseBig,
                [&](int64_t, int64_t v, bool valid) { bf_sp_mark(s, g, v, valid); });
}

// The marked light rows of a k-column working set: a group of Lk lanes per
// row, one lane per column (as bf_multi_kernel's light part).
template <class TA, class TD>
__global__ __launch_bounds__(kT) void bf_sp_rows_multi_kernel(BfSparse s, SellView g, int64_t k, int64_t ld,
                                                              int logLk, TD* d,
                                                              const uint32_t* __restrict__ chp,
                                                              uint32_t* __restrict__ chn,
                                                              unsigned int* changed,
                                                              unsigned long long* changed_mask,
                                                              const int32_t* done) {
  if (done && *done) return;
  if (!s.cnt[3]) return;
  const int Lk = 1 << logLk;
  const int wl = threadIdx.x & 63;
  const int lane = wl & (Lk - 1);
  const int gbase = wl & ~(Lk - 1);
  const int64_t jblocks = (k + Lk - 1) >> logLk;
  const int64_t gpb = kT >> logLk;
  const int64_t ngroups = (int64_t)s.cnt[1] * jblocks;
  const TA* val = (const TA*)g.val;
  for (int64_t q0 = (int64_t)blockIdx.x * gpb; q0 < ngroups; q0 += (int64_t)gridDim.x * gpb) {
    const int64_t q = q0 + (threadIdx.x >> logLk);
    bool ch = false;
    int64_t v = -1;
    if (q < ngroups) {
      const int64_t unit = q / jblocks;
      const int64_t j = (q - unit * jblocks) * Lk + lane;
      v = s.dq[unit];
      if (j < k) {
        const StateRow r = state_row(g, false, nullptr, v);
        TD acc = Lim<TD>::max();
        int64_t kk = 0;
        for (; kk + 4 <= r.len; kk += 4) {  // batches of 4, as multi_row_min
          int32_t c[4];
          bool on[4];
          TA w[4];
          TD dn[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) c[u] = r.col[r.base + (kk + u) * r.stride] - 1;
#pragma unroll
          for (int u = 0; u < 4; ++u) on[u] = bit_get(chp, c[u]);
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (on[u]) {
              w[u] = val[r.base + (kk + u) * r.stride];
              dn[u] = d[(int64_t)c[u] * ld + j];
            }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (on[u]) acc = jmin(acc, sadd(TD(w[u]), dn[u]));
        }
        for (; kk < r.len; ++kk) {
          const int64_t e = r.base + kk * r.stride;
          const int32_t c = r.col[e] - 1;
          if (bit_get(chp, c)) acc = jmin(acc, sadd(TD(val[e]), d[(int64_t)c * ld + j]));
        }
        if (acc < Lim<TD>::max()) {
          const TD dv = d[v * ld + j];
          const TD nv = jmin(dv, acc);
          if (nv < dv) {
            d[v * ld + j] = nv;
            ch = true;
          }
        }
      }
    }
    wave_flag(changed, ch);
    const unsigned long long b = __ballot(ch);
    const unsigned long long low = Lk == 64 ? ~0ull : ((1ull << Lk) - 1ull);
    if (lane == 0 && ((b >> gbase) & low)) bit_claim(chn, v);
    if (changed_mask) {
      unsigned long long cm = 0;
      for (int o = 0; o < 64; o += Lk) cm |= (b >> o) & low;
      if (wl == 0 && cm) or_new_bits(changed_mask, cm);
    }
  }
}

// The marked light rows with V columns per lane (as bf_multi_vec_kernel).
template <class TA, class TD, int V>
__global__ __launch_bounds__(kT) void bf_sp_rows_vec_kernel(BfSparse s, SellView g, int64_t k, int64_t ld,
                                                            int logLv, TD* d,
                                                            const uint32_t* __restrict__ chp,
                                                            uint32_t* __restrict__ chn,
                                                            unsigned int* changed,
                                                            unsigned long long* changed_mask,
                                                            const int32_t* done, ColGroups cg) {
  typedef TD VT __attribute__((ext_vector_type(V)));
  if (done && *done) return;
  if (!s.cnt[3]) return;
  const int Lv = 1 << logLv;
  const int wl = threadIdx.x & 63;
  const int lane = wl & (Lv - 1);
  const int gbase = wl & ~(Lv - 1);
  const int64_t kq = cg.n ? cg.n : k / V;
  const int64_t jblocks = (kq + Lv - 1) >> logLv;
  const int64_t gpb = kT >> logLv;
  const int64_t ngroups = (int64_t)s.cnt[1] * jblocks;
  const TA* val = (const TA*)g.val;
  for (int64_t q0 = (int64_t)blockIdx.x * gpb; q0 < ngroups; q0 += (int64_t)gridDim.x * gpb) {
    const int64_t q = q0 + (threadIdx.x >> logLv);
    unsigned chbits = 0;
    int64_t v = -1;
    int64_t j0 = 0;
    if (q < ngroups) {
      const int64_t unit = q / jblocks;
      const int64_t jq = (q - unit * jblocks) * Lv + lane;
      j0 = (cg.n && jq < kq ? (int64_t)cg.g[jq] : jq) * V;
      v = s.dq[unit];
      if (jq < kq) {
        const StateRow r = state_row(g, false, nullptr, v);
        VT acc;
#pragma unroll
        for (int t = 0; t < V; ++t) acc[t] = Lim<TD>::max();
        int64_t kk = 0;
        for (; kk + 4 <= r.len; kk += 4) {
          int32_t c[4];
          bool on[4];
          TA w[4];
          VT dn[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) c[u] = r.col[r.base + (kk + u) * r.stride] - 1;
#pragma unroll
          for (int u = 0; u < 4; ++u) on[u] = bit_get(chp, c[u]);
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (on[u]) {
              w[u] = val[r.base + (kk + u) * r.stride];
              dn[u] = *(const VT*)&d[(int64_t)c[u] * ld + j0];
            }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (on[u])
#pragma unroll
              for (int t = 0; t < V; ++t) acc[t] = jmin(acc[t], sadd(TD(w[u]), dn[u][t]));
        }
        for (; kk < r.len; ++kk) {
          const int64_t e = r.base + kk * r.stride;
          const int32_t c = r.col[e] - 1;
          if (!bit_get(chp, c)) continue;
          const TD w = TD(val[e]);
          const VT dn = *(const VT*)&d[(int64_t)c * ld + j0];
#pragma unroll
          for (int t = 0; t < V; ++t) acc[t] = jmin(acc[t], sadd(w, dn[t]));
        }
        bool cand = false;
#pragma unroll
        for (int t = 0; t < V; ++t) cand |= acc[t] < Lim<TD>::max();
        if (cand) {
          VT own = *(const VT*)&d[v * ld + j0];
#pragma unroll
          for (int t = 0; t < V; ++t)
            if (acc[t] < own[t]) {
              own[t] = acc[t];
              chbits |= 1u << t;
            }
          if (chbits) *(VT*)&d[v * ld + j0] = own;
        }
      }
    }
    const bool ch = chbits != 0;
    wave_flag(changed, ch);
    const unsigned long long b = __ballot(ch);
    const unsigned long long low = Lv == 64 ? ~0ull : ((1ull << Lv) - 1ull);
    if (lane == 0 && ((b >> gbase) & low)) bit_claim(chn, v);
    if (changed_mask) {  // k <= 64
      unsigned long long cm = ch ? (unsigned long long)chbits << j0 : 0ull;
      for (int o = 32; o > 0; o >>= 1) cm |= __shfl_xor(cm, o, 64);
      if (wl == 0 && cm) or_new_bits(changed_mask, cm);
    }
  }
}

// Enqueues the sparse-iteration prologue: fallen-row edges, the switch, the
// fallen queue and the marking of the rows to relax.
static void bf_sparse_prologue(hipStream_t st, const BfSparse& s, const SellView& g,
                               const uint32_t* chp, const int32_t* done, int cus) {
  static_assert(kT == kBlockClaimThreads, "block_claim / block_add assume 256 threads");
  const unsigned nbw = blocks_for(s.words, kT * kScanWords), nb = (unsigned)std::max(cus, 1) * 8;
  hipLaunchKernelGGL(bf_sp_start_kernel, dim3(1), dim3(1), 0, st, s, done);
  hipLaunchKernelGGL(bf_sp_fallen_kernel<false>, dim3(nbw), dim3(kT), 0, st, s, chp, g, done);
  hipLaunchKernelGGL(bf_sp_decide_kernel, dim3(1), dim3(1), 0, st, s, done);
  hipLaunchKernelGGL(bf_sp_fallen_kernel<true>, dim3(nbw), dim3(kT), 0, st, s, chp, g, done);
  hipLaunchKernelGGL(bf_sp_mark_tile_kernel, dim3(nb), dim3(256), 0, st, s, g, done);
  hipLaunchKernelGGL(bf_sp_mark_big_kernel, dim3(nb), dim3(256), 0, st, s, g, done);
}

template <class TD>
__global__ void bf_multi_seed_kernel(TD* d, const int32_t* sources, int64_t k,
                                     const int32_t* vinv, uint32_t* chp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const int64_t v = sources[i] - 1;
  const int64_t r = vinv ? (int64_t)vinv[v] - 1 : v;
  d[r * k + i] = TD(0);
  atomicOr(&chp[r >> 5], 1u << (r & 31));
}

template <class TD>
__global__ void fill_td_kernel(TD* p, int64_t n, TD v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <class TD>
__global__ void seed_one_kernel(TD* p, int64_t s, const int32_t* vinv, uint32_t* chp) {
  const int64_t r = vinv ? (int64_t)vinv[s] - 1 : s;
  p[r] = TD(0);
  chp[r >> 5] |= 1u << (r & 31);
}

// Row-major working set (state rows, k columns) -> result at original rows,
// column-major m x k (or row-major when rowmajor).  State row of original v:
// vinv[v] - 1 for a relabelled SELL, else v.
template <class TD>
__global__ void state_to_result_kernel(const TD* __restrict__ src, const int32_t* __restrict__ vinv,
                                       int64_t m, int64_t k, int rowmajor, TD* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m * k) return;
  const int64_t v = i / k, j = i - v * k;
  const int64_t r = vinv ? (int64_t)vinv[v] - 1 : v;
  dst[rowmajor ? i : v + j * m] = src[r * k + j];
}

// The same into the column-major result through an LDS tile of 4096 values:
// TC = k rounded up to a power of two in 4..64 columns by TR = 4096 / TC rows
// (rows padded to TC + 1 values, so a column is read without bank conflicts).
// A block reads its state rows' TC values (contiguous in each row) and writes
// each column's TR consecutive vertices contiguously, where one thread per
// value wrote the k values of a row k columns apart.
template <class TD>
__global__ __launch_bounds__(256) void state_to_cols_kernel(const TD* __restrict__ src,
                                                            const int32_t* __restrict__ vinv,
                                                            int64_t m, int64_t k, int logTC,
                                                            TD* __restrict__ dst) {
  __shared__ TD t[4096 + 1024];
  const int TC = 1 << logTC, TR = 4096 >> logTC;
  const int64_t v0 = (int64_t)blockIdx.x * TR, j0 = (int64_t)blockIdx.y * TC;
  const int64_t jn = min((int64_t)TC, k - j0);
  for (int idx = threadIdx.x; idx < 4096; idx += 256) {
    const int rr = idx >> logTC, jj = idx & (TC - 1);
    const int64_t v = v0 + rr;
    if (v < m && jj < jn) {
      const int64_t r = vinv ? (int64_t)vinv[v] - 1 : v;
      t[rr * (TC + 1) + jj] = src[r * k + j0 + jj];
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 4096; idx += 256) {
    const int jj = idx >> (12 - logTC), rr = idx & (TR - 1);
    const int64_t v = v0 + rr;
    if (v < m && jj < jn) dst[(j0 + jj) * m + v] = t[rr * (TC + 1) + jj];
  }
}

__global__ void zero_u32_kernel(unsigned int* p) { *p = 0u; }

// Captured iterations: ctl = {iterations, done}.  After an iteration: done
// when nothing fell, else clear the flag for the next one.
__global__ void iter_begin_kernel(unsigned int* changed, int32_t* ctl) {
  *changed = 0u;
  ctl[0] = 0;
  ctl[1] = 0;
}
__global__ void iter_end_kernel(unsigned int* changed, int32_t* ctl) {
  if (ctl[1]) return;
  ctl[0] += 1;
  if (*changed == 0u) ctl[1] = 1;
  else *changed = 0u;
}

// Dist type of shortest_path: Int32 for Bool values, else the value type
// (src/algorithms/shortest_path.jl:18).
static int dist_type_for(int val_dtype) { return val_dtype == GG_BOOL ? GG_INT32 : val_dtype; }

template <class F> static int with_ad(int ta, int td, F&& f) {
  if (td != dist_type_for(ta))
    return set_error(GG_ERR_ARGUMENT, "dist element type must be %s of the matrix values",
                     ta == GG_BOOL ? "Int32 for Bool" : "the type");
  return with_type(ta, [&](auto a) -> int {
    typedef typename decltype(a)::type TA;
    if constexpr (std::is_same<TA, jbool>::value) return f(a, Tag<int32_t>{});
    else return f(a, Tag<TA>{});
  });
}

static int read_u32(gg_ctx* ctx, const unsigned int* d, unsigned int* h) {
  GG_HIP(hipMemcpyAsync(ctx->host_pinned, d, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  GG_HIP(hipStreamSynchronize(ctx->stream));
  *h = *(unsigned*)ctx->host_pinned;
  return GG_OK;
}

// Changed-row bitmaps of an iteration: prev (read), next (written, cleared
// before the step).  Both live in one workspace slot.
struct Changed {
  uint32_t* prev;
  uint32_t* next;
  int64_t words;
};

static int changed_bitmaps(gg_ctx* ctx, int64_t ns, Changed& c) {
  c.words = (ns + 31) / 32;
  void* p;
  GG_TRY(ws_get(ctx, WS_SSSP2, sizeof(uint32_t) * 2 * (c.words + 64), &p));
  c.prev = (uint32_t*)p;
  c.next = c.prev + c.words + 64;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(c.words, kT)), dim3(kT), 0, ctx->stream,
                     c.prev, c.words, 0u);
  return GG_OK;
}

// Iterate until an iteration lowers nothing; the check reads a device flag
// every iteration for the first 16 and then every 8 (an iteration after the
// fixpoint reads no neighbour and changes nothing).  With graph steps, U
// iterations are captured (the bitmaps alternate, U even) and relaunched
// until the device-side flag says done.  step(stream, prev, next, done)
// enqueues one iteration; `key` identifies everything its kernels read.
template <class Step>
static int iterate(gg_ctx* ctx, unsigned int* changed, Changed& c, int64_t max_iter, Step&& step,
                   std::vector<uint64_t> key, int64_t* iters) {
  if (ctx->graph_steps > 0) {
    const int U = (ctx->graph_steps + 1) & ~1;
    int32_t* ctl = (int32_t*)changed + 4;
    hipLaunchKernelGGL(iter_begin_kernel, dim3(1), dim3(1), 0, ctx->stream, changed, ctl);
    key.push_back((uint64_t)U);
    sig_add(key, c);
    sig_add(key, changed);
    auto steps = [&](hipStream_t st) -> int {
      for (int i = 0; i < U; ++i) {
        uint32_t* p = (i & 1) ? c.next : c.prev;
        uint32_t* nx = (i & 1) ? c.prev : c.next;
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(c.words, kT)), dim3(kT), 0, st, nx,
                           c.words, 0u);
        GG_TRY(step(st, p, nx, ctl + 1));
        hipLaunchKernelGGL(iter_end_kernel, dim3(1), dim3(1), 0, st, changed, ctl);
      }
      GG_HIP(hipGetLastError());
      return GG_OK;
    };
    const int st = graph_loop(ctx, key, ctl + 1, max_iter / U + 2, steps, nullptr);
    if (st == GG_ERR_NOT_CONVERGED)
      return set_error(st, "no fixpoint after %lld iterations (negative cycle?)",
                       (long long)(max_iter / U + 2) * U);
    GG_TRY(st);
    GG_HIP(hipMemcpyAsync(ctx->host_pinned, ctl, sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    GG_HIP(hipStreamSynchronize(ctx->stream));
    if (iters) *iters = *(const int32_t*)ctx->host_pinned;
    return GG_OK;
  }
  int64_t it = 0;
  while (true) {
    ++it;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, ctx->stream, changed);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(c.words, kT)), dim3(kT), 0, ctx->stream,
                       c.next, c.words, 0u);
    GG_TRY(step(ctx->stream, c.prev, c.next, nullptr));
    std::swap(c.prev, c.next);
    if (it <= 16 || it % 8 == 0) {
      unsigned h = 0;
      GG_TRY(read_u32(ctx, changed, &h));
      if (h == 0) break;
    }
    if (max_iter > 0 && it >= max_iter) {
      if (iters) *iters = it;
      return set_error(GG_ERR_NOT_CONVERGED, "no fixpoint after %lld iterations (negative cycle?)",
                       (long long)it);
    }
  }
  if (iters) *iters = it;
  return GG_OK;
}

int sssp_one(gg_ctx* ctx, const gg_matrix* A, int64_t source, void* dist, int dist_dtype,
             int64_t* iterations) {
  return with_ad(A->val_dtype, dist_dtype, [&](auto a, auto dt) -> int {
    typedef typename decltype(a)::type TA;
    typedef typename decltype(dt)::type TD;
    const int64_t m = A->m;
    // a relabelled SELL iterates in internal order in a workspace and
    // scatters the fixpoint to dist; captured iterations use the workspace
    // too (a graph keeps its pointers); otherwise dist is the working set
    const int32_t* vinv = (A->flags & GG_FLAG_INTERNAL_IDS) ? A->vinv : nullptr;
    const bool in_ws = vinv || ctx->graph_steps > 0;
    const int64_t ns = state_size(A);
    void *tmp = nullptr, *cnt;
    if (in_ws) GG_TRY(ws_get(ctx, WS_SSSP0, sizeof(TD) * ns, &tmp));
    GG_TRY(ws_get(ctx, WS_FLAG, 64, &cnt));
    Changed c;
    GG_TRY(changed_bitmaps(ctx, ns, c));
    TD* d = in_ws ? (TD*)tmp : (TD*)dist;
    hipLaunchKernelGGL(fill_td_kernel<TD>, dim3(blocks_for(ns, kT)), dim3(kT), 0, ctx->stream, d,
                       ns, Lim<TD>::max());
    hipLaunchKernelGGL(seed_one_kernel<TD>, dim3(1), dim3(1), 0, ctx->stream, d, source - 1, vinv,
                       c.prev);
    const int lg = csr_log_lanes(ctx, A);
    SellView g = sell_view(A);
    const int64_t nsb = (g.n_slots + kT - 1) / kT;
    auto step = [&](hipStream_t st, const uint32_t* chp, uint32_t* chn,
                    const int32_t* done) -> int {
      HotTimer ht(ctx, st == ctx->stream);
      if (A->format == GG_CSR) {
        unsigned nb = blocks_for(m << lg, kT);
        const unsigned cap = (unsigned)ctx->num_cus * 32;
        if (nb > cap) nb = cap;
        hipLaunchKernelGGL((bf_csr_kernel<TA, TD>), dim3(nb), dim3(kT), 0, st, A->ptr, A->idx,
                           (const TA*)A->val, m, lg, d, chp, chn, (unsigned*)cnt, done);
      } else if (nsb + g.n_seg > 0) {
        hipLaunchKernelGGL((bf_sell_kernel<TA, TD>), dim3((unsigned)(nsb + g.n_seg)), dim3(kT),
                           0, st, g, nsb, d, chp, chn, (unsigned*)cnt, done);
      }
      GG_HIP(hipGetLastError());
      return GG_OK;
    };
    std::vector<uint64_t> key{3, (uint64_t)dist_dtype};
    sig_add(key, *A);
    sig_add(key, d);
    int st = iterate(ctx, (unsigned*)cnt, c, m + 16, step, key, iterations);
    if (st == GG_OK && in_ws) {
      hipLaunchKernelGGL(state_to_result_kernel<TD>, dim3(blocks_for(m, kT)), dim3(kT), 0,
                         ctx->stream, d, vinv, m, (int64_t)1, 1, (TD*)dist);
      GG_HIP(hipGetLastError());
    }
    GG_HIP(hipStreamSynchronize(ctx->stream));
    return st;
  });
}

int bf_multi_step(hipStream_t stream, const gg_matrix* A, int dist_dtype, int64_t k, void* d,
                  const uint32_t* chp, uint32_t* chn, unsigned int* changed,
                  unsigned long long* changed_mask, const int32_t* done) {
  return bf_multi_step_sp(stream, A, dist_dtype, k, d, chp, chn, changed, changed_mask, done,
                          nullptr, 256);
}

int bf_multi_step_sp(hipStream_t stream, const gg_matrix* A, int dist_dtype, int64_t k, void* d,
                     const uint32_t* chp, uint32_t* chn, unsigned int* changed,
                     unsigned long long* changed_mask, const int32_t* done, const BfSparse* sp,
                     int cus, int64_t ld, const ColGroups* cols) {
  if (ld < 0) ld = k;
  if (ld < k) return set_error(GG_ERR_ARGUMENT, "row stride below the columns");
  if (changed_mask && k > 64) return set_error(GG_ERR_ARGUMENT, "column masks need k <= 64");
  return with_ad(A->val_dtype, dist_dtype, [&](auto a, auto dt) -> int {
    typedef typename decltype(a)::type TA;
    typedef typename decltype(dt)::type TD;
    int lg = 0;
    while (lg < 6 && (1LL << lg) < k) ++lg;
    const int64_t Lk = 1LL << lg;
    const int64_t jblocks = (k + Lk - 1) / Lk;
    const int64_t gpb = kT / Lk;
    SellView g = sell_view(A);
    if (A->format == GG_CSR) {
      g.colval = A->idx;
      g.val = A->val;
      g.n_heavy = 0;
      g.n_seg = 0;
    }
    const int64_t n_light = A->format == GG_CSR ? A->m : g.n_slots;
    const int64_t light_blocks = (n_light * jblocks + gpb - 1) / gpb;
    const int64_t parts = g.n_seg > 0 ? (g.seg_len + 255) / 256 : 1;
    const int64_t nb = light_blocks + g.n_seg * parts * jblocks;
    if (nb > 0x7fffffffLL) return set_error(GG_ERR_ARGUMENT, "multi-source launch too large");
    if (sp) bf_sparse_prologue(stream, *sp, g, chp, done, cus);
    constexpr int V = 16 / (int)sizeof(TD);
    const bool vec = A->format == GG_SELL && k % V == 0 && ld % V == 0 && ((uintptr_t)d & 15) == 0;
    ColGroups cg{};
    if (cols && cols->n > 0) {
      if (!vec) return set_error(GG_ERR_ARGUMENT, "column groups need the 16-byte column kernels");
      if (cols->n > 32 || cols->n > k / V)
        return set_error(GG_ERR_ARGUMENT, "more column groups than columns");
      for (int i = 0; i < cols->n; ++i)
        if ((int64_t)cols->g[i] >= k / V) return set_error(GG_ERR_ARGUMENT, "column group out of range");
      cg = *cols;
    }
    // light slots, heavy parts and sparse rows VV = V columns (16 bytes) per
    // lane (32 bytes per lane, half the lanes per row, ran at 0.6x on R-MAT)
    auto run_vec = [&](auto vt) -> int {
      constexpr int VV = decltype(vt)::value;
      const int64_t nq = cg.n ? cg.n : k / VV;  // VV-column groups each row relaxes
      int lv = 0;
      while (lv < 6 && (1LL << lv) < nq) ++lv;
      const int64_t jbv = (nq + (1LL << lv) - 1) >> lv;
      const int64_t gpbv = kT >> lv;
      const int64_t lbv = (n_light * jbv + gpbv - 1) / gpbv;
      if (lbv > 0x7fffffffLL) return set_error(GG_ERR_ARGUMENT, "multi-source launch too large");
      const int64_t pv = g.n_seg > 0 ? (g.seg_len + kHeavyVecPart - 1) / kHeavyVecPart : 1;
      const int64_t nh = g.n_seg * pv * jbv;
      if (nh + lbv > 0x7fffffffLL) return set_error(GG_ERR_ARGUMENT, "multi-source launch too large");
      if (lbv > 0 && nh > 0)
        hipLaunchKernelGGL((bf_multi_fused_vec_kernel<TA, TD, VV>), dim3((unsigned)(nh + lbv)),
                           dim3(kT), 0, stream, g, n_light, pv, nh, k, ld, lv, (TD*)d, chp, chn,
                           changed, changed_mask, done, sp ? &sp->cnt[4] : nullptr, cg);
      else if (lbv > 0)
        hipLaunchKernelGGL((bf_multi_vec_kernel<TA, TD, VV>), dim3((unsigned)lbv), dim3(kT), 0,
                           stream, g, n_light, k, ld, lv, (TD*)d, chp, chn, changed, changed_mask,
                           done, sp ? &sp->cnt[4] : nullptr, cg);
      else if (nh > 0)
        hipLaunchKernelGGL((bf_multi_heavy_vec_kernel<TA, TD, VV>), dim3((unsigned)nh), dim3(kT), 0,
                           stream, g, pv, k, ld, lv, (TD*)d, chp, chn, changed, changed_mask, done,
                           cg);
      if (sp)
        hipLaunchKernelGGL((bf_sp_rows_vec_kernel<TA, TD, VV>), dim3((unsigned)std::max(cus, 1) * 8),
                           dim3(kT), 0, stream, *sp, g, k, ld, lv, (TD*)d, chp, chn, changed,
                           changed_mask, done, cg);
      return GG_OK;
    };
    if (vec) {
      GG_TRY(run_vec(std::integral_constant<int, V>{}));
    } else {
      if (nb > 0)
        hipLaunchKernelGGL((bf_multi_kernel<TA, TD>), dim3((unsigned)nb), dim3(kT), sizeof(TD) * kT,
                           stream, A->format, A->ptr, g, n_light, parts, k, ld, lg, (TD*)d, chp,
                           chn, changed, changed_mask, done, sp ? &sp->cnt[4] : nullptr);
      if (sp)
        hipLaunchKernelGGL((bf_sp_rows_multi_kernel<TA, TD>), dim3((unsigned)std::max(cus, 1) * 8),
                           dim3(kT), 0, stream, *sp, g, k, ld, lg, (TD*)d, chp, chn, changed,
                           changed_mask, done);
    }
    if (sp) {
      hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(sp->words, kT)), dim3(kT), 0, stream,
                         sp->mark, sp->words, 0u);
    }
    GG_HIP(hipGetLastError());
    return GG_OK;
  });
}

// k-source Bellman-Ford in a row-major working set; result written
// column-major into dist (or row-major when rowmajor_out).
int sssp_multi(gg_ctx* ctx, const gg_matrix* A, const int32_t* sources_host, int64_t k,
               void* dist, int dist_dtype, bool rowmajor_out, int64_t* iterations) {
  return with_ad(A->val_dtype, dist_dtype, [&](auto, auto dt) -> int {
    typedef typename decltype(dt)::type TD;
    const int64_t m = A->m;
    const int64_t ns = state_size(A);
    const int32_t* vinv = (A->flags & GG_FLAG_INTERNAL_IDS) ? A->vinv : nullptr;
    void *b0, *cnt, *src;
    GG_TRY(ws_get(ctx, WS_MS0, sizeof(TD) * ns * k, &b0));
    GG_TRY(ws_get(ctx, WS_MS2, sizeof(int32_t) * k, &src));
    GG_TRY(ws_get(ctx, WS_FLAG, 64, &cnt));
    Changed c;
    GG_TRY(changed_bitmaps(ctx, ns, c));
    GG_HIP(hipMemcpyAsync(src, sources_host, sizeof(int32_t) * k, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(fill_td_kernel<TD>, dim3(blocks_for(ns * k, kT)), dim3(kT), 0, ctx->stream,
                       (TD*)b0, ns * k, Lim<TD>::max());
    hipLaunchKernelGGL(bf_multi_seed_kernel<TD>, dim3(blocks_for(k, kT)), dim3(kT), 0,
                       ctx->stream, (TD*)b0, (const int32_t*)src, k, vinv, c.prev);
    BfSparse sp{};
    const bool sparse = bf_sparse_ok(A) && ctx->sssp_alpha > 0.0;
    if (sparse) {
      void* sb;
      GG_TRY(ws_get(ctx, WS_SSSP1, bf_sparse_bytes(A), &sb));
      sp = bf_sparse_layout(sb, A, ctx->sssp_alpha);
      hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(sp.words, kT)), dim3(kT), 0, ctx->stream,
                         sp.mark, sp.words, 0u);
    }
    auto step = [&](hipStream_t st, const uint32_t* chp, uint32_t* chn,
                    const int32_t* done) -> int {
      HotTimer ht(ctx, st == ctx->stream);
      return bf_multi_step_sp(st, A, dist_dtype, k, b0, chp, chn, (unsigned*)cnt, nullptr, done,
                              sparse ? &sp : nullptr, ctx->num_cus);
    };
    std::vector<uint64_t> key{4, (uint64_t)dist_dtype, (uint64_t)k};
    sig_add(key, *A);
    sig_add(key, b0);
    sig_add(key, sp);
    int st = iterate(ctx, (unsigned*)cnt, c, m + 16, step, key, iterations);
    if (st == GG_OK) {
      if (rowmajor_out && !vinv) {
        GG_HIP(hipMemcpyAsync(dist, b0, sizeof(TD) * m * k, hipMemcpyDeviceToDevice, ctx->stream));
      } else if (rowmajor_out) {
        hipLaunchKernelGGL(state_to_result_kernel<TD>, dim3(blocks_for(m * k, kT)), dim3(kT), 0,
                           ctx->stream, (const TD*)b0, vinv, m, k, 1, (TD*)dist);
        GG_HIP(hipGetLastError());
      } else {
        int ltc = 2;  // columns per tile: k rounded up to a power of two in 4..64
        while (ltc < 6 && (1LL << ltc) < k) ++ltc;
        hipLaunchKernelGGL(state_to_cols_kernel<TD>,
                           dim3((unsigned)blocks_for(m, 4096 >> ltc), (unsigned)blocks_for(k, 1 << ltc)),
                           dim3(256), 0, ctx->stream, (const TD*)b0, vinv, m, k, ltc, (TD*)dist);
        GG_HIP(hipGetLastError());
      }
    }
    GG_HIP(hipStreamSynchronize(ctx->stream));
    return st;
  });
}

static int check_sp(const gg_matrix* A, int64_t source) {
  GG_TRY(check_matrix(A));
  if (A->format == GG_CSC) return set_error(GG_ERR_ARGUMENT, "Bellman-Ford runs on CSR or SELL A_T");
  if (A->m != A->n) return set_error(GG_ERR_DIMENSION, "A_T must be square");
  if (source < 1 || source > A->m)
    return set_error(GG_ERR_BOUNDS, "source %lld outside 1:%lld", (long long)source, (long long)A->m);
  return GG_OK;
}

}  // namespace gg

using namespace gg;

extern "C" {

int gg_shortest_path(gg_ctx* ctx, const gg_matrix* A_T, int64_t source, void* dist,
                     int dist_dtype, int64_t* iterations) {
  if (!ctx || !dist) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  GG_TRY(check_sp(A_T, source));
  timer_begin(ctx);
  int st = sssp_one(ctx, A_T, source, dist, dist_dtype, iterations);
  timer_end(ctx);
  return st;
}

int gg_shortest_path_multi(gg_ctx* ctx, const gg_matrix* A_T, const int32_t* sources, int64_t k,
                           void* dist, int dist_dtype, int64_t* iterations) {
  if (!ctx || !dist || (k > 0 && !sources)) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  if (k < 1) return set_error(GG_ERR_ARGUMENT, "need at least one source");
  for (int64_t i = 0; i < k; ++i) GG_TRY(check_sp(A_T, sources[i]));
  timer_begin(ctx);
  int st = sssp_multi(ctx, A_T, sources, k, dist, dist_dtype, false, iterations);
  timer_end(ctx);
  return st;
}

int gg_mssp_distances(gg_ctx* ctx, const gg_matrix* A_T, const int32_t* sources, int64_t k,
                      void* dist, int dist_dtype, int64_t* iterations) {
  return gg_shortest_path_multi(ctx, A_T, sources, k, dist, dist_dtype, iterations);
}

}  // extern "C"
