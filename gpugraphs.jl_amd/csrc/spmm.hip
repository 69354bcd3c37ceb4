// spmm.hip -- semiring SpMM with a dense operand (gg_spmm).
//
// A group of Lk lanes (Lk = next power of two >= min(k, 64)) owns one row and
// Lk consecutive columns of B / C; every lane walks the row's entries in
// storage order, so each (row, column) result is accumulated exactly as the
// reference's work-item does, including its absorbing exit:
//   CSR  csr_spmm_kernel!  src/spmm/CSR.jl:16-34 (break after the update),
//   SELL sell_spmm_kernel! src/spmm/SELL.jl:2-39 (break on pad -1 or on
//        acc == terminal before the entry).
// The lanes of a group read the same colval / nzval entry (one broadcast
// load) and, with a row-major operand, k consecutive B values of the same
// vertex (one coalesced line); the column-major layout of the reference's
// Julia matrices is accepted as well.
#include "frontier.hpp"

namespace gg {

constexpr int kBlk = 256;

template <class S, class TA, class TB>
struct MmTypes {
  typedef typename CompT<S, TA, TB>::type CT;
  typedef typename AccOf<S::ADD, CT>::type A;
  typedef typename OutT<S, TA, TB>::type TC;
  typedef Monoid<S::ADD, A> MO;
  typedef Mul<S::MUL, CT> MU;
};

struct MmArgs {
  int format;
  const int32_t* ptr;
  const int32_t* idx;
  const void* val;
  const int32_t* perm_inv;  // unused (rows are located through slot search below)
  int64_t C, m, n, k;
  // SELL hybrid
  const int32_t* slot_of_row;  // [m] 0-based slot of each light row, -1 for heavy rows
  const int32_t* slot_len;
  const int32_t* heavy_of_row;  // [m] heavy index or -1
  const int64_t* heavy_ptr;
  const int32_t* heavy_col;
  const void* heavy_val;
  int b_rowmajor, c_rowmajor;
  int64_t r0, nrows;
  // internal ids: columns are slot / slot_base + heavy index; secondi maps
  // them back through perm / heavy_row
  int internal;
  int64_t slot_base;
  const int32_t* perm;
  const int32_t* heavy_row;
};

template <class S>
__device__ __forceinline__ int32_t mm_col(const MmArgs& g, int32_t c) {
  if constexpr (S::MUL == GG_OP_SECONDI) {
    if (g.internal)
      return c - 1 < g.slot_base ? g.perm[c - 1] : g.heavy_row[c - 1 - g.slot_base];
  }
  return c;
}

template <class S, class TA, class TB>
__global__ __launch_bounds__(kBlk) void spmm_kernel(typename MmTypes<S, TA, TB>::TC* __restrict__ Cm,
                                                    const TB* __restrict__ B, MmArgs g, int accum,
                                                    const void* __restrict__ mask, int mdt,
                                                    int logLk) {
  typedef MmTypes<S, TA, TB> T;
  typedef typename T::A A;
  typedef typename T::CT CT;
  const int Lk = 1 << logLk;
  const int lane = threadIdx.x & (Lk - 1);
  const int64_t jblocks = (g.k + Lk - 1) >> logLk;
  const int64_t ngroups = g.nrows * jblocks;
  const int64_t G = ((int64_t)gridDim.x * kBlk) >> logLk;
  for (int64_t q = ((int64_t)blockIdx.x * kBlk + threadIdx.x) >> logLk; q < ngroups; q += G) {
    const int64_t row = g.r0 + q / jblocks;
    const int64_t j = (q % jblocks) * Lk + lane;
    if (j >= g.k || !mask_on(mask, mdt, row)) continue;
    A acc = T::MO::neutral();
    auto bval = [&](int32_t col) -> CT {
      return CT(g.b_rowmajor ? B[(int64_t)(col - 1) * g.k + j] : B[(int64_t)(col - 1) + j * g.n]);
    };
    if (g.format == GG_CSR) {
      const TA* val = (const TA*)g.val;
      for (int32_t i = g.ptr[row] - 1; i < g.ptr[row + 1] - 1; ++i) {
        const int32_t col = g.idx[i];
        acc = T::MO::op(acc, A(T::MU::ap(CT(val[i]), bval(col), col)));
        if (T::MO::has_absorb && acc == T::MO::absorb()) break;
      }
    } else {
      const TA* val = (const TA*)g.val;
      const int32_t h = g.heavy_of_row ? g.heavy_of_row[row] : -1;
      if (h >= 0) {
        const TA* hv = (const TA*)g.heavy_val;
        for (int64_t i = g.heavy_ptr[h]; i < g.heavy_ptr[h + 1]; ++i) {
          if (T::MO::has_absorb && acc == T::MO::absorb()) break;
          const int32_t col = g.heavy_col[i];
          acc = T::MO::op(acc, A(T::MU::ap(CT(hv[i]), bval(col), mm_col<S>(g, col))));
        }
      } else {
        const int64_t slot = g.slot_of_row ? g.slot_of_row[row] : row;
        const int64_t sl = slot / g.C, r = slot - sl * g.C;
        const int64_t base = (int64_t)g.ptr[sl] - 1 + r;
        const int64_t len = g.slot_len ? (int64_t)g.slot_len[slot]
                                       : ((int64_t)g.ptr[sl + 1] - g.ptr[sl]) / g.C;
        for (int64_t kk = 0; kk < len; ++kk) {
          const int64_t i = base + kk * g.C;
          const int32_t col = g.idx[i];
          if (col == -1 || (T::MO::has_absorb && acc == T::MO::absorb())) break;
          acc = T::MO::op(acc, A(T::MU::ap(CT(val[i]), bval(col), mm_col<S>(g, col))));
        }
      }
    }
    const int64_t ci = g.c_rowmajor ? row * g.k + j : row + j * g.m;
    CT v = CT(acc);
    if (accum_reads_c(accum)) v = accum_apply<CT>(accum, CT(Cm[ci]), v);
    Cm[ci] = typename T::TC(v);
  }
}

// ------------------------------------------------ SELL, 16-byte column groups
// SELL with a row-major B (k a multiple of VB = 16 / sizeof(TB), 16-byte
// aligned): a group of Lv lanes per slot, each lane VB consecutive columns
// read as one 16-byte load per entry, entries in batches of 4 (index / value
// loads, then the B loads, then the folds in storage order).  Each column
// folds its terms in the reference's order and stops at the absorbing value
// (src/spmm/SELL.jl:22-27), so light rows give the one-lane-per-column
// result exactly.  Heavy rows: one block per (segment, column block), its
// groups splitting the segment's entries; a row of one segment is written
// directly, the others fold their segment partials in segment order
// (spmm_combine_kernel) -- exact for the idempotent monoids (min, max, or,
// and, any) and a reassociation of the sum for plus on floats.
template <class A, class MO>
__device__ __forceinline__ void mm_fold(A& acc, A term) {
  if (!(MO::has_absorb && acc == MO::absorb())) acc = MO::op(acc, term);
}

template <class TC, class CT>
__device__ __forceinline__ void mm_store(TC* Cm, const MmArgs& g, int64_t row, int64_t j, int accum,
                                         CT v) {
  const int64_t ci = g.c_rowmajor ? row * g.k + j : row + j * g.m;
  if (accum_reads_c(accum)) v = accum_apply<CT>(accum, CT(Cm[ci]), v);
  Cm[ci] = TC(v);
}

template <class S, class TA, class TB, int VB>
__global__ __launch_bounds__(kBlk) void spmm_sell_vec_kernel(
    typename MmTypes<S, TA, TB>::TC* __restrict__ Cm, const TB* __restrict__ B, MmArgs g,
    const int32_t* __restrict__ perm, int64_t n_slots, int accum, int logLv) {
  typedef MmTypes<S, TA, TB> T;
  typedef typename T::A A;
  typedef typename T::CT CT;
  typedef TB VT __attribute__((ext_vector_type(VB)));
  const int Lv = 1 << logLv;
  const int lane = threadIdx.x & (Lv - 1);
  const int64_t kq = g.k / VB;
  const int64_t jblocks = (kq + Lv - 1) >> logLv;
  const int64_t q = ((int64_t)blockIdx.x * kBlk + threadIdx.x) >> logLv;
  const int64_t slot = q / jblocks;
  const int64_t jq = (q - slot * jblocks) * Lv + lane;
  if (slot >= n_slots || jq >= kq) return;
  int64_t row = slot;
  if (perm) {
    if (perm[slot] == 0) return;
    row = perm[slot] - 1;
  } else if (row >= g.m) {
    return;
  }
  const int64_t j0 = jq * VB;
  const int64_t sl = slot / g.C, r = slot - sl * g.C;
  const int64_t base = (int64_t)g.ptr[sl] - 1 + r;
  const int64_t len = g.slot_len ? (int64_t)g.slot_len[slot]
                                 : ((int64_t)g.ptr[sl + 1] - g.ptr[sl]) / g.C;
  const TA* __restrict__ val = (const TA*)g.val;
  A acc[VB];
#pragma unroll
  for (int t = 0; t < VB; ++t) acc[t] = T::MO::neutral();
  int64_t kk = 0;
  bool end = false;
  for (; kk + 4 <= len && !end; kk += 4) {
    int32_t c[4];
    TA v[4];
    VT b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = g.idx[base + (kk + u) * g.C];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = val[base + (kk + u) * g.C];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (c[u] != -1) b[u] = *(const VT*)&B[(int64_t)(c[u] - 1) * g.k + j0];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (c[u] == -1) end = true;  // the reference's pad break
      if (!end)
#pragma unroll
        for (int t = 0; t < VB; ++t)
          mm_fold<A, typename T::MO>(acc[t],
                                     A(T::MU::ap(CT(v[u]), CT(b[u][t]), mm_col<S>(g, c[u]))));
    }
  }
  for (; kk < len && !end; ++kk) {
    const int64_t i = base + kk * g.C;
    const int32_t col = g.idx[i];
    if (col == -1) break;
    const VT b = *(const VT*)&B[(int64_t)(col - 1) * g.k + j0];
#pragma unroll
    for (int t = 0; t < VB; ++t)
      mm_fold<A, typename T::MO>(acc[t], A(T::MU::ap(CT(val[i]), CT(b[t]), mm_col<S>(g, col))));
  }
#pragma unroll
  for (int t = 0; t < VB; ++t) mm_store<typename T::TC, CT>(Cm, g, row, j0 + t, accum, CT(acc[t]));
}

template <class S, class TA, class TB, int VB>
__global__ __launch_bounds__(kBlk) void spmm_heavy_vec_kernel(
    typename MmTypes<S, TA, TB>::TC* __restrict__ Cm, const TB* __restrict__ B, MmArgs g,
    const int32_t* __restrict__ heavy_row, const int32_t* __restrict__ heavy_seg,
    int64_t n_heavy, int64_t seg_len, void* partial, int accum, int logLv) {
  typedef MmTypes<S, TA, TB> T;
  typedef typename T::A A;
  typedef typename T::CT CT;
  typedef TB VT __attribute__((ext_vector_type(VB)));
  __shared__ A red[kBlk * VB];
  const int Lv = 1 << logLv;
  const int lane = threadIdx.x & (Lv - 1);
  const int64_t grp = threadIdx.x >> logLv;
  const int64_t gpb = kBlk >> logLv;
  const int64_t kq = g.k / VB;
  const int64_t jblocks = (kq + Lv - 1) >> logLv;
  const int64_t sg = (int64_t)blockIdx.x / jblocks;
  const int64_t jq = ((int64_t)blockIdx.x - sg * jblocks) * Lv + lane;
  const bool jin = jq < kq;
  const int64_t j0 = jq * VB;
  int64_t lo = 0, hi = n_heavy;  // heavy_seg[lo] <= sg < heavy_seg[lo+1]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (heavy_seg[mid] <= sg) lo = mid; else hi = mid;
  }
  const int64_t nseg = heavy_seg[lo + 1] - heavy_seg[lo];
  const int64_t s = g.heavy_ptr[lo] + (sg - heavy_seg[lo]) * seg_len;
  const int64_t e = min(s + seg_len, g.heavy_ptr[lo + 1]);
  const TA* __restrict__ hv = (const TA*)g.heavy_val;
  A acc[VB];
#pragma unroll
  for (int t = 0; t < VB; ++t) acc[t] = T::MO::neutral();
  if (jin) {
    int64_t i = s + grp;
    for (; i + 3 * gpb < e; i += 4 * gpb) {
      int32_t c[4];
      TA v[4];
      VT b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) c[u] = g.heavy_col[i + u * gpb];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = hv[i + u * gpb];
#pragma unroll
      for (int u = 0; u < 4; ++u) b[u] = *(const VT*)&B[(int64_t)(c[u] - 1) * g.k + j0];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int t = 0; t < VB; ++t)
          mm_fold<A, typename T::MO>(acc[t],
                                     A(T::MU::ap(CT(v[u]), CT(b[u][t]), mm_col<S>(g, c[u]))));
    }
    for (; i < e; i += gpb) {
      const int32_t c = g.heavy_col[i];
      const VT b = *(const VT*)&B[(int64_t)(c - 1) * g.k + j0];
#pragma unroll
      for (int t = 0; t < VB; ++t)
        mm_fold<A, typename T::MO>(acc[t], A(T::MU::ap(CT(hv[i]), CT(b[t]), mm_col<S>(g, c))));
    }
  }
#pragma unroll
  for (int t = 0; t < VB; ++t) red[threadIdx.x * VB + t] = acc[t];
  __syncthreads();
  if (grp != 0 || !jin) return;
  for (int64_t o = 1; o < gpb; ++o)
#pragma unroll
    for (int t = 0; t < VB; ++t) mm_fold<A, typename T::MO>(acc[t], red[(o * Lv + lane) * VB + t]);
  if (nseg == 1) {
    const int64_t row = heavy_row[lo] - 1;
#pragma unroll
    for (int t = 0; t < VB; ++t) mm_store<typename T::TC, CT>(Cm, g, row, j0 + t, accum, CT(acc[t]));
  } else {
#pragma unroll
    for (int t = 0; t < VB; ++t) ((A*)partial)[sg * g.k + j0 + t] = acc[t];
  }
}

// CSR rows with VB columns per lane (the SELL kernel's column groups; the
// group's lanes read the same entries of the row, a broadcast): rows
// r0 .. r0 + nrows - 1, mask applied per row (src/spmm/CSR.jl:111-204).
template <class S, class TA, class TB, int VB>
__global__ __launch_bounds__(kBlk) void spmm_csr_vec_kernel(
    typename MmTypes<S, TA, TB>::TC* __restrict__ Cm, const TB* __restrict__ B, MmArgs g,
    int accum, const void* __restrict__ mask, int mdt, int logLv) {
  typedef MmTypes<S, TA, TB> T;
  typedef typename T::A A;
  typedef typename T::CT CT;
  typedef TB VT __attribute__((ext_vector_type(VB)));
  const int Lv = 1 << logLv;
  const int lane = threadIdx.x & (Lv - 1);
  const int64_t kq = g.k / VB;
  const int64_t jblocks = (kq + Lv - 1) >> logLv;
  const int64_t G = ((int64_t)gridDim.x * kBlk) >> logLv;
  const TA* __restrict__ val = (const TA*)g.val;
  for (int64_t q = ((int64_t)blockIdx.x * kBlk + threadIdx.x) >> logLv; q < g.nrows * jblocks;
       q += G) {
    const int64_t row = g.r0 + q / jblocks;
    const int64_t jq = (q % jblocks) * Lv + lane;
    if (jq >= kq || !mask_on(mask, mdt, row)) continue;
    const int64_t j0 = jq * VB;
    A acc[VB];
#pragma unroll
    for (int t = 0; t < VB; ++t) acc[t] = T::MO::neutral();
    int64_t i = (int64_t)g.ptr[row] - 1;
    const int64_t e = (int64_t)g.ptr[row + 1] - 1;
    for (; i + 4 <= e; i += 4) {
      int32_t c[4];
      TA v[4];
      VT b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) c[u] = g.idx[i + u];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = val[i + u];
#pragma unroll
      for (int u = 0; u < 4; ++u) b[u] = *(const VT*)&B[(int64_t)(c[u] - 1) * g.k + j0];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int t = 0; t < VB; ++t)
          mm_fold<A, typename T::MO>(acc[t], A(T::MU::ap(CT(v[u]), CT(b[u][t]), c[u])));
    }
    for (; i < e; ++i) {
      const int32_t col = g.idx[i];
      const VT b = *(const VT*)&B[(int64_t)(col - 1) * g.k + j0];
#pragma unroll
      for (int t = 0; t < VB; ++t)
        mm_fold<A, typename T::MO>(acc[t], A(T::MU::ap(CT(val[i]), CT(b[t]), col)));
    }
#pragma unroll
    for (int t = 0; t < VB; ++t)
      mm_store<typename T::TC, CT>(Cm, g, row, j0 + t, accum, CT(acc[t]));
  }
}

// Column-major n x k -> row-major through a 64 x 64 LDS tile (both sides
// whole lines), for the SELL kernel's row-major operand.
template <class T>
__global__ __launch_bounds__(kBlk) void to_rowmajor_kernel(const T* __restrict__ src, int64_t n,
                                                           int64_t k, T* __restrict__ dst) {
  __shared__ T tile[64][65];
  const int64_t tiles_n = (n + 63) / 64;
  const int64_t r0 = ((int64_t)blockIdx.x % tiles_n) * 64;
  const int64_t c0 = ((int64_t)blockIdx.x / tiles_n) * 64;
  for (int i = threadIdx.x; i < 64 * 64; i += kBlk) {
    const int c = i >> 6, r = i & 63;  // consecutive threads: consecutive rows of a column
    if (r0 + r < n && c0 + c < k) tile[r][c] = src[(r0 + r) + (c0 + c) * n];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 64 * 64; i += kBlk) {
    const int r = i >> 6, c = i & 63;  // consecutive threads: consecutive columns of a row
    if (r0 + r < n && c0 + c < k) dst[(r0 + r) * k + c0 + c] = tile[r][c];
  }
}

// Row-major m x k result -> the caller's column-major C, applying accum
// (the same accum_apply the kernels use on a direct store).
template <class TC>
__global__ __launch_bounds__(kBlk) void from_rowmajor_accum_kernel(const TC* __restrict__ src,
                                                                   int64_t m, int64_t k,
                                                                   int accum, TC* __restrict__ dst) {
  __shared__ TC tile[64][65];
  const int64_t tiles_m = (m + 63) / 64;
  const int64_t r0 = ((int64_t)blockIdx.x % tiles_m) * 64;
  const int64_t c0 = ((int64_t)blockIdx.x / tiles_m) * 64;
  for (int i = threadIdx.x; i < 64 * 64; i += kBlk) {
    const int r = i >> 6, c = i & 63;
    if (r0 + r < m && c0 + c < k) tile[r][c] = src[(r0 + r) * k + c0 + c];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 64 * 64; i += kBlk) {
    const int c = i >> 6, r = i & 63;
    if (r0 + r < m && c0 + c < k) {
      const int64_t ci = (r0 + r) + (c0 + c) * m;
      TC v = tile[r][c];
      if (accum_reads_c(accum)) v = accum_apply<TC>(accum, dst[ci], v);
      dst[ci] = v;
    }
  }
}

// Rows of several heavy segments: their partials folded in segment order.
template <class S, class TA, class TB>
__global__ __launch_bounds__(kBlk) void spmm_combine_kernel(
    typename MmTypes<S, TA, TB>::TC* __restrict__ Cm, MmArgs g,
    const int32_t* __restrict__ heavy_row, const int32_t* __restrict__ heavy_seg,
    int64_t n_heavy, const void* partial, int accum) {
  typedef MmTypes<S, TA, TB> T;
  typedef typename T::A A;
  typedef typename T::CT CT;
  const int64_t id = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  const int64_t h = id / g.k, j = id - h * g.k;
  if (h >= n_heavy) return;
  const int64_t s0 = heavy_seg[h], s1 = heavy_seg[h + 1];
  if (s1 - s0 < 2) return;
  A acc = T::MO::neutral();
  for (int64_t sg = s0; sg < s1; ++sg) mm_fold<A, typename T::MO>(acc, ((const A*)partial)[sg * g.k + j]);
  mm_store<typename T::TC, CT>(Cm, g, heavy_row[h] - 1, j, accum, CT(acc));
}

// Inverse maps of a SELL hybrid (row -> slot, row -> heavy index), built once
// per matrix view and cached in the context by device pointer of perm.
__global__ void sell_row_maps_kernel(const int32_t* perm, int64_t n_slots, int32_t* slot_of_row) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_slots) return;
  const int32_t p = perm[s];
  if (p > 0) slot_of_row[p - 1] = (int32_t)s;
}
__global__ void heavy_row_map_kernel(const int32_t* heavy_row, int64_t n_heavy, int32_t* heavy_of_row) {
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h < n_heavy) heavy_of_row[heavy_row[h] - 1] = (int32_t)h;
}
__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Row -> slot / heavy maps of a SELL hybrid in workspace slots s0, s1.
int sell_row_maps(gg_ctx* ctx, const gg_matrix* A, int s0, int s1, const int32_t** slot_of_row,
                  const int32_t** heavy_of_row) {
  *slot_of_row = nullptr;
  *heavy_of_row = nullptr;
  if (A->format != GG_SELL || (!A->perm && A->n_heavy == 0) || A->m == 0) return GG_OK;
  void *p0 = nullptr, *p1 = nullptr;
  GG_TRY(ws_get(ctx, s0, sizeof(int32_t) * A->m, &p0));
  GG_TRY(ws_get(ctx, s1, sizeof(int32_t) * A->m, &p1));
  const unsigned nb = blocks_for(A->m, 256);
  hipLaunchKernelGGL(fill_i32_kernel, dim3(nb), dim3(256), 0, ctx->stream, (int32_t*)p0, A->m, -1);
  hipLaunchKernelGGL(fill_i32_kernel, dim3(nb), dim3(256), 0, ctx->stream, (int32_t*)p1, A->m, -1);
  if (A->perm) {
    const int64_t n_slots = A->nslices * A->slice_size;
    hipLaunchKernelGGL(sell_row_maps_kernel, dim3(blocks_for(n_slots, 256)), dim3(256), 0,
                       ctx->stream, A->perm, n_slots, (int32_t*)p0);
  }
  if (A->n_heavy > 0)
    hipLaunchKernelGGL(heavy_row_map_kernel, dim3(blocks_for(A->n_heavy, 256)), dim3(256), 0,
                       ctx->stream, A->heavy_row, A->n_heavy, (int32_t*)p1);
  GG_HIP(hipGetLastError());
  *slot_of_row = A->perm ? (const int32_t*)p0 : nullptr;
  *heavy_of_row = A->n_heavy ? (const int32_t*)p1 : nullptr;
  return GG_OK;
}

int launch_spmm(gg_ctx* ctx, void* C, int c_dtype, int c_layout, const gg_matrix* A,
                const void* B, int b_dtype, int b_layout, int64_t k, int add, int mul, int accum,
                const void* mask, int mask_dtype, int64_t r0, int64_t nrows, int ws0, int ws1) {
  if (nrows <= 0 || k <= 0) return GG_OK;
  MmArgs g{};
  g.format = A->format;
  g.ptr = A->ptr;
  g.idx = A->idx;
  g.val = A->val;
  g.C = A->slice_size;
  g.m = A->m;
  g.n = A->n;
  g.k = k;
  g.slot_len = A->slot_len;
  g.heavy_ptr = A->heavy_ptr;
  g.heavy_col = A->heavy_col;
  g.heavy_val = A->heavy_val;
  g.b_rowmajor = b_layout == GG_ROW_MAJOR;
  g.c_rowmajor = c_layout == GG_ROW_MAJOR;
  g.r0 = r0;
  g.nrows = nrows;
  GG_TRY(sell_row_maps(ctx, A, ws0, ws1, &g.slot_of_row, &g.heavy_of_row));
  int lg = 0;
  while (lg < 6 && (1LL << lg) < k) ++lg;
  return dispatch(add, mul, A->val_dtype, b_dtype, [&](auto s, auto ta, auto tb) -> int {
    typedef decltype(s) S;
    typedef typename decltype(ta)::type TA;
    typedef typename decltype(tb)::type TB;
    typedef typename MmTypes<S, TA, TB>::TC TC;
    if (c_dtype != code_of<TC>())
      return set_error(GG_ERR_ARGUMENT, "result element type must equal promote_type(eltype(A), eltype(B))");
    if (!accum_ok(accum, std::is_floating_point<TC>::value))
      return set_error(GG_ERR_UNSUPPORTED, "accum operator %d not supported for this type", accum);
    if (A->format == GG_SELL && (A->flags & GG_FLAG_INTERNAL_IDS)) {
      // columns are internal ids: B rows gathered into internal order, row-major;
      // a column-major B is made row-major first (LDS-tiled), so that the
      // gather moves whole rows instead of one element per line
      const int64_t n_int = A->nslices * A->slice_size + A->n_heavy;
      const TB* bsrc = (const TB*)B;
      if (!g.b_rowmajor && A->n > 0 && k > 1) {
        void* tp;
        GG_TRY(ws_get(ctx, WS_BROWS, sizeof(TB) * (size_t)(A->n * k), &tp));
        const int64_t tiles = ((A->n + 63) / 64) * ((k + 63) / 64);
        if (tiles > 0x7fffffffLL) return set_error(GG_ERR_ARGUMENT, "SpMM launch too large");
        hipLaunchKernelGGL(to_rowmajor_kernel<TB>, dim3((unsigned)tiles), dim3(kBlk), 0,
                           ctx->stream, (const TB*)B, A->n, k, (TB*)tp);
        bsrc = (const TB*)tp;
        g.b_rowmajor = 1;
      }
      void* bp;
      GG_TRY(ws_get(ctx, WS_XPERM, sizeof(TB) * (size_t)(n_int * k), &bp));
      hipLaunchKernelGGL(gather_internal_kernel<TB>, dim3(blocks_for(n_int * k, kBlk)), dim3(kBlk),
                         0, ctx->stream, bsrc, A->perm, A->nslices * A->slice_size,
                         A->heavy_row, A->n_heavy, A->n, k, g.b_rowmajor, (TB*)bp);
      B = bp;
      g.b_rowmajor = 1;
      g.n = n_int;
      g.internal = 1;
      g.slot_base = A->nslices * A->slice_size;
      g.perm = A->perm;
      g.heavy_row = A->heavy_row;
    }
    constexpr int VB = 16 / (int)sizeof(TB);
    // SELL with a row-major B of 16-byte column groups; integer plus keeps the
    // one-lane-per-column kernel (its absorbing exit depends on the order)
    const bool int_plus = S::ADD == GG_OP_ADD && is_int<typename MmTypes<S, TA, TB>::CT>();
    const bool sell_vec = A->format == GG_SELL && k % VB == 0 && !int_plus && !mask &&
                          nrows == A->m;
    const bool csr_vec = A->format == GG_CSR && k % VB == 0 && !int_plus;
    const bool vec = sell_vec || csr_vec;
    if (vec && !g.b_rowmajor) {  // the reference's column-major operand, made row-major once
      void* bp;
      GG_TRY(ws_get(ctx, WS_XPERM, sizeof(TB) * (size_t)(A->n * k), &bp));
      const int64_t tiles = ((A->n + 63) / 64) * ((k + 63) / 64);
      if (tiles > 0x7fffffffLL) return set_error(GG_ERR_ARGUMENT, "SpMM launch too large");
      if (tiles > 0)
        hipLaunchKernelGGL(to_rowmajor_kernel<TB>, dim3((unsigned)tiles), dim3(kBlk), 0,
                           ctx->stream, (const TB*)B, A->n, k, (TB*)bp);
      B = bp;
      g.b_rowmajor = 1;
    }
    if (csr_vec && ((uintptr_t)B & 15) == 0) {
      int lv = 0;
      while (lv < 6 && (1LL << lv) < k / VB) ++lv;
      const int64_t jb = (k / VB + (1LL << lv) - 1) >> lv;
      unsigned nb = blocks_for((nrows * jb) << lv, kBlk);
      const unsigned cap = (unsigned)ctx->num_cus * 64;
      if (nb > cap) nb = cap;
      HotTimer ht(ctx);
      hipLaunchKernelGGL((spmm_csr_vec_kernel<S, TA, TB, VB>), dim3(nb), dim3(kBlk), 0,
                         ctx->stream, (TC*)C, (const TB*)B, g, accum, mask, mask_dtype, lv);
      GG_HIP(hipGetLastError());
      return GG_OK;
    }
    if (sell_vec && ((uintptr_t)B & 15) == 0) {
      int lv = 0;
      while (lv < 6 && (1LL << lv) < k / VB) ++lv;
      const int64_t jb = (k / VB + (1LL << lv) - 1) >> lv;
      const int64_t n_slots = A->nslices * A->slice_size;
      const int64_t nb_l = (((n_slots * jb) << lv) + kBlk - 1) / kBlk;
      const int64_t nb_h = A->n_heavy > 0 ? A->heavy_nseg * jb : 0;
      if (nb_l > 0x7fffffffLL || nb_h > 0x7fffffffLL)
        return set_error(GG_ERR_ARGUMENT, "SpMM launch too large");
      void* part = nullptr;
      if (A->n_heavy > 0)
        GG_TRY(ws_get(ctx, WS_PARTIAL, sizeof(typename MmTypes<S, TA, TB>::A) *
                                           (size_t)(A->heavy_nseg * k), &part));
      // a column-major C is written row-major to a workspace first (the
      // slots' rows are scattered, so direct column stores are short) and
      // turned with accum applied by from_rowmajor_accum_kernel
      void* const C_user = C;
      const int accum_user = accum;
      const bool via_rm = !g.c_rowmajor &&
                          std::is_same<TC, typename MmTypes<S, TA, TB>::CT>::value;
      if (via_rm) {
        GG_TRY(ws_get(ctx, WS_MS1, sizeof(TC) * (size_t)(A->m * k), &C));
        g.c_rowmajor = 1;
        accum = GG_OP_SECOND;
      }
      HotTimer ht(ctx);
      if (n_slots > 0)
        hipLaunchKernelGGL((spmm_sell_vec_kernel<S, TA, TB, VB>), dim3((unsigned)nb_l), dim3(kBlk),
                           0, ctx->stream, (TC*)C, (const TB*)B, g, A->perm, n_slots, accum, lv);
      if (nb_h > 0) {
        hipLaunchKernelGGL((spmm_heavy_vec_kernel<S, TA, TB, VB>), dim3((unsigned)nb_h),
                           dim3(kBlk), 0, ctx->stream, (TC*)C, (const TB*)B, g, A->heavy_row,
                           A->heavy_seg, A->n_heavy, A->heavy_seg_len, part, accum, lv);
        if (A->heavy_nseg > A->n_heavy)
          hipLaunchKernelGGL((spmm_combine_kernel<S, TA, TB>),
                             dim3(blocks_for(A->n_heavy * k, kBlk)), dim3(kBlk), 0, ctx->stream,
                             (TC*)C, g, A->heavy_row, A->heavy_seg, A->n_heavy, part, accum);
      }
      if (via_rm) {
        const int64_t tiles = ((A->m + 63) / 64) * ((k + 63) / 64);
        if (tiles > 0x7fffffffLL) return set_error(GG_ERR_ARGUMENT, "SpMM launch too large");
        if (tiles > 0)
          hipLaunchKernelGGL(from_rowmajor_accum_kernel<TC>, dim3((unsigned)tiles), dim3(kBlk), 0,
                             ctx->stream, (const TC*)C, A->m, k, accum_user, (TC*)C_user);
      }
      GG_HIP(hipGetLastError());
      return GG_OK;
    }
    const int64_t groups = nrows * ((k + (1 << lg) - 1) >> lg);
    unsigned nb = blocks_for(groups << lg, kBlk);
    const unsigned cap = (unsigned)ctx->num_cus * 64;
    if (nb > cap) nb = cap;
    HotTimer ht(ctx);
    hipLaunchKernelGGL((spmm_kernel<S, TA, TB>), dim3(nb), dim3(kBlk), 0, ctx->stream, (TC*)C,
                       (const TB*)B, g, accum, mask, mask_dtype, lg);
    GG_HIP(hipGetLastError());
    return GG_OK;
  });
}

// CSC SpMM by atomic scatter, the k-column form of csc_spmv_kernel!
// (src/spmv/CSC.jl:2-21; src/spmm/CSC.jl is a placeholder): thread (column c
// of A, column j of B) adds A[r,c] * B[c,j] into C[r,j] for every entry of
// column c.  Element (i, j) of B / C sits at i * rs + j * cs.
template <class T>
__global__ __launch_bounds__(256) void csc_spmm_kernel(T* __restrict__ c, int64_t c_rs, int64_t c_cs,
                                                       const int32_t* __restrict__ colptr,
                                                       const int32_t* __restrict__ rowval,
                                                       const T* __restrict__ val,
                                                       const T* __restrict__ b, int64_t b_rs,
                                                       int64_t b_cs, int64_t n, int64_t k) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n * k) return;
  const int64_t col = t / k, j = t - col * k;
  const T bc = b[col * b_rs + j * b_cs];
  for (int32_t i = colptr[col] - 1; i < colptr[col + 1] - 1; ++i) {
    const T v = jmul(bc, val[i]);
    T* dst = &c[(int64_t)(rowval[i] - 1) * c_rs + j * c_cs];
    if constexpr (std::is_same<T, int32_t>::value) atomicAdd((int*)dst, (int)v);
    else atomicAdd(dst, v);
  }
}

}  // namespace gg

using namespace gg;

extern "C" int gg_spmm(gg_ctx* ctx, void* C, int c_dtype, int64_t c_rows, int64_t c_cols,
                       int c_layout, const gg_matrix* A, const void* B, int b_dtype,
                       int64_t b_rows, int64_t b_cols, int b_layout, int add, int mul, int accum,
                       const void* mask, int mask_dtype, int64_t mask_len, int64_t range_start,
                       int64_t range_len) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  clear_error();
  GG_TRY(check_matrix(A));
  // src/spmm/CSR.jl:131-133
  if (A->n != b_rows || c_rows != A->m || c_cols != b_cols)
    return set_error(GG_ERR_ASSERTION, "SpMM dimensions must agree");
  if ((c_layout != GG_COL_MAJOR && c_layout != GG_ROW_MAJOR) ||
      (b_layout != GG_COL_MAJOR && b_layout != GG_ROW_MAJOR))
    return set_error(GG_ERR_ARGUMENT, "unknown layout");
  if (A->format == GG_CSC) {
    // as the CSC SpMV: plus-times by atomic adds into C, add and accum not
    // applied, one element type, neither mask nor range
    if (mask || range_len >= 0)
      return set_error(GG_ERR_ARGUMENT, "CSC SpMM takes neither mask nor range");
    if (A->val_dtype != b_dtype || c_dtype != b_dtype)
      return set_error(GG_ERR_ARGUMENT, "CSC SpMM needs one element type for A, B and C");
    if (A->n == 0 || b_cols == 0) return GG_OK;
    const int64_t b_rs = b_layout == GG_ROW_MAJOR ? b_cols : 1;
    const int64_t b_cs = b_layout == GG_ROW_MAJOR ? 1 : b_rows;
    const int64_t c_rs = c_layout == GG_ROW_MAJOR ? c_cols : 1;
    const int64_t c_cs = c_layout == GG_ROW_MAJOR ? 1 : c_rows;
    timer_begin(ctx);
    int st = with_type(c_dtype, [&](auto t) -> int {
      typedef typename decltype(t)::type T;
      if constexpr (std::is_same<T, float>::value || std::is_same<T, double>::value ||
                    std::is_same<T, int32_t>::value) {
        hipLaunchKernelGGL(csc_spmm_kernel<T>, dim3(blocks_for(A->n * b_cols, 256)), dim3(256), 0,
                           ctx->stream, (T*)C, c_rs, c_cs, A->ptr, A->idx, (const T*)A->val,
                           (const T*)B, b_rs, b_cs, A->n, b_cols);
        GG_HIP(hipGetLastError());
        return GG_OK;
      } else {
        return set_error(GG_ERR_UNSUPPORTED, "CSC SpMM supports Float32, Float64, Int32");
      }
    });
    timer_end(ctx);
    return st;
  }
  if (A->format == GG_SELL && (mask || range_len >= 0))
    return set_error(GG_ERR_ARGUMENT, "SELL SpMM takes neither mask nor range (src/spmm/SELL.jl:84-98)");
  int64_t r0 = 0, nrows = A->m;
  if (range_len >= 0) {
    if (mask) return set_error(GG_ERR_ASSERTION, "range and mask together (src/spmm/CSR.jl:160)");
    if (range_start < 1 || range_start + range_len - 1 > A->m)
      return set_error(GG_ERR_ASSERTION, "range outside the rows of A (src/spmm/CSR.jl:161)");
    r0 = range_start - 1;
    nrows = range_len;
  }
  if (mask && mask_len != A->m)
    return set_error(GG_ERR_ASSERTION, "length(mask) == size(A, 1) (src/spmm/CSR.jl:183)");
  timer_begin(ctx);
  int st = launch_spmm(ctx, C, c_dtype, c_layout, A, B, b_dtype, b_layout, b_cols, add, mul,
                       accum, mask, mask_dtype, r0, nrows, WS_MS3, WS_MS4);
  timer_end(ctx);
  if (st == GG_ERR_UNSUPPORTED && gg_last_error()[0] == 0)
    set_error(st, "no kernel for add=%d mul=%d with value types %d, %d", add, mul, A->val_dtype, b_dtype);
  return st;
}
