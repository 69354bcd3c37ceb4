// spmv.hip -- semiring SpMV over CSR, SELL-C-sigma and CSC (gg_spmv).
//
// CSR (src/spmv/CSR.jl): groups of L lanes own a row (L a power of two chosen
// from the mean row length), stream the row's colval/nzval contiguously and
// combine with a cross-lane butterfly.  Each lane stops at the absorbing
// element as the reference's loop does (src/spmv/CSR.jl:49-51); for integer
// plus, whose absorbing exit changes the value, L is 1 so the exit happens at
// exactly the reference's position.  The any-with-mask fast path
// (src/spmv/CSR.jl:147-171) picks the first entry in row order whose b value
// is nonzero: lanes take their first hit and the group keeps the smallest
// position.
//
// SELL (src/spmv/SELL.jl): one lane per slot; the 64 lanes of a wavefront
// cover one C=64 slice and read entry k of all rows as one contiguous 256-byte
// line, with non-temporal loads (the entries are read once; x stays cached).  With slot lengths (sigma-sorted layout) a lane stops at its row's
// end, so padding is never fetched; without them it stops at the -1 pad like
// the reference.  Rows of the heavy partition are cut into segments of
// heavy_seg_len entries, one 256-thread block per segment, in the same launch;
// rows with several segments are finished by a small combine kernel that
// folds the partials in segment order (deterministic).
//
// CSC (src/spmv/CSC.jl:2-21): column-parallel atomic scatter
// c[row] += b[col] * a, plus-times only, add / accum ignored as in the
// reference.
#include "frontier.hpp"

namespace gg {

constexpr int kBlock = 256;
constexpr int kU = 4;  // entries per batch of a SELL slot walk
constexpr int kH = 8;  // entries per thread per batch of a heavy segment

template <class S, class TA, class TX>
struct Types {
  typedef typename CompT<S, TA, TX>::type CT;      // operator type
  typedef typename AccOf<S::ADD, CT>::type A;     // accumulator type
  typedef typename OutT<S, TA, TX>::type TC;      // result array type
  typedef Monoid<S::ADD, A> MO;
  typedef Mul<S::MUL, CT> MU;
};

template <class S, class TA, class TX>
__device__ __forceinline__ typename Types<S, TA, TX>::A term(TA a, TX x, int32_t col) {
  typedef Types<S, TA, TX> T;
  return typename T::A(T::MU::ap(typename T::CT(a), typename T::CT(x), col));
}

template <class TC, class CT, class A>
__device__ __forceinline__ void store_acc(TC* c, int64_t row, int accum, A acc) {
  CT v = CT(acc);
  if (accum_reads_c(accum)) v = accum_apply<CT>(accum, CT(c[row]), v);
  c[row] = TC(v);
}

// ------------------------------------------------------------------ CSR
template <class S, class TA, class TX>
__global__ __launch_bounds__(kBlock) void csr_spmv_kernel(
    typename Types<S, TA, TX>::TC* __restrict__ c, const int32_t* __restrict__ rowptr,
    const int32_t* __restrict__ colval, const TA* __restrict__ val, const TX* __restrict__ x,
    int accum, const void* __restrict__ mask, int mdt, int64_t r0, int64_t nrows, int logL,
    int any_fast, const int32_t* __restrict__ rows, int64_t m) {
  typedef Types<S, TA, TX> T;
  typedef typename T::A A;
  const int L = 1 << logL;
  const int lane = threadIdx.x & (L - 1);
  const int64_t G = ((int64_t)gridDim.x * kBlock) >> logL;
  for (int64_t q = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> logL; q < nrows; q += G) {
    // a row list (1-based; entries outside 1..m are skipped), or r0 + q
    const int64_t row = rows ? (int64_t)rows[q] - 1 : r0 + q;
    if (rows && (row < 0 || row >= m)) continue;
    if (!mask_on(mask, mdt, row)) continue;
    const int32_t s = rowptr[row] - 1, e = rowptr[row + 1] - 1;
    if (any_fast) {
      int32_t first = 0x7fffffff;
      for (int32_t i = s + lane; i < e; i += L)
        if (x[colval[i] - 1] != TX(0)) { first = i; break; }
      for (int o = L >> 1; o > 0; o >>= 1) first = min(first, __shfl_xor(first, o, 64));
      if (lane == 0 && first != 0x7fffffff) {
        const int32_t col = colval[first];
        store_acc<typename T::TC, typename T::CT>(c, row, accum, term<S, TA, TX>(val[first], x[col - 1], col));
      }
      continue;
    }
    A acc = T::MO::neutral();
    for (int32_t i = s + lane; i < e; i += L) {
      const int32_t col = colval[i];
      acc = T::MO::op(acc, term<S, TA, TX>(val[i], x[col - 1], col));
      if (T::MO::has_absorb && acc == T::MO::absorb()) break;
    }
    acc = group_reduce<typename T::MO>(acc, L);
    if (lane == 0) store_acc<typename T::TC, typename T::CT>(c, row, accum, acc);
  }
}

// ------------------------------------------------------------------ SELL
struct SellArgs {
  const int32_t* slice_ptr;
  const int32_t* colval;
  const void* val;
  const int32_t* perm;
  const int32_t* slot_len;
  int64_t C, logC, n_slots, m;
  int64_t r_lo, r_hi;  // rows [r_lo, r_hi) (0-based) are computed
  int64_t n_sell_blocks;
  int64_t n_heavy;
  const int32_t* heavy_row;
  const int64_t* heavy_ptr;
  const int32_t* heavy_seg;
  int64_t seg_len;
  const int32_t* heavy_col;
  const void* heavy_val;
  void* partial;  // [n_seg] accumulators of multi-segment rows
  // Relabelled layouts: each row's value (the operator type, before accum)
  // goes to stage[internal id], coalesced, and stage_to_rows_kernel moves it
  // to the row's original id afterwards (NULL: results go straight to c)
  void* stage;
  int internal;       // columns are internal ids (slot, then slot_base + heavy index)
  int64_t slot_base;  // nslices * C
};

// Column handed to the multiply: secondi reads it, and with internal ids it
// must be the original column (perm / heavy_row of the internal id).
template <class S>
__device__ __forceinline__ int32_t mul_col(const SellArgs& g, int32_t c) {
  if constexpr (S::MUL == GG_OP_SECONDI) {
    if (g.internal)
      return c - 1 < g.slot_base ? g.perm[c - 1] : g.heavy_row[c - 1 - g.slot_base];
  }
  return c;
}

template <class MO, class A>
__device__ __forceinline__ A block_reduce(A v, A* lds) {
  v = group_reduce<MO>(v, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kBlock / 64; ++i) v = MO::op(v, lds[i]);
  }
  return v;
}

template <class S, class TA, class TX>
__global__ __launch_bounds__(kBlock) void sell_spmv_kernel(
    typename Types<S, TA, TX>::TC* __restrict__ c, const TX* __restrict__ x, SellArgs g,
    int accum, const void* __restrict__ mask, int mdt) {
  typedef Types<S, TA, TX> T;
  typedef typename T::A A;
  if ((int64_t)blockIdx.x < g.n_sell_blocks) {
    const int64_t slot = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (slot >= g.n_slots) return;
    int64_t row;
    if (g.perm) {
      const int32_t p = g.perm[slot];
      if (p == 0) return;
      row = p - 1;
    } else {
      row = slot;
      if (row >= g.m) return;
    }
    if (row < g.r_lo || row >= g.r_hi || !mask_on(mask, mdt, row)) return;
    int64_t sl, r;
    if (g.logC >= 0) { sl = slot >> g.logC; r = slot & (g.C - 1); }
    else { sl = slot / g.C; r = slot - sl * g.C; }
    const int64_t base = (int64_t)g.slice_ptr[sl] - 1 + r;
    const TA* __restrict__ val = (const TA*)g.val;
    A acc = T::MO::neutral();
    if (g.slot_len) {
      // batches of kU entries: all index / value loads, then all gathers, then
      // the folds in storage order, so kU loads of each kind are in flight
      const int32_t len = g.slot_len[slot];
      int64_t i = base;
      int32_t k = 0;
      for (; k + kU <= len; k += kU, i += kU * g.C) {
        int32_t c[kU];
        TA v[kU];
        TX xv[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) c[u] = __builtin_nontemporal_load(&g.colval[i + u * g.C]);
#pragma unroll
        for (int u = 0; u < kU; ++u) v[u] = __builtin_nontemporal_load(&val[i + u * g.C]);
#pragma unroll
        for (int u = 0; u < kU; ++u) xv[u] = x[c[u] - 1];
#pragma unroll
        for (int u = 0; u < kU; ++u)
          acc = T::MO::op(acc, term<S, TA, TX>(v[u], xv[u], mul_col<S>(g, c[u])));
      }
      for (; k < len; ++k, i += g.C) {
        const int32_t col = g.colval[i];
        acc = T::MO::op(acc, term<S, TA, TX>(val[i], x[col - 1], mul_col<S>(g, col)));
      }
    } else {
      const int64_t end = (int64_t)g.slice_ptr[sl + 1] - 1;
      for (int64_t i = base; i < end; i += g.C) {
        const int32_t col = g.colval[i];
        if (col == -1) break;
        acc = T::MO::op(acc, term<S, TA, TX>(val[i], x[col - 1], mul_col<S>(g, col)));
      }
    }
    if (g.stage) ((typename T::CT*)g.stage)[slot] = typename T::CT(acc);
    else store_acc<typename T::TC, typename T::CT>(c, row, accum, acc);
    return;
  }
  // heavy segment
  __shared__ A lds[kBlock / 64];
  const int64_t sg = (int64_t)blockIdx.x - g.n_sell_blocks;
  int64_t lo = 0, hi = g.n_heavy;  // heavy_seg[lo] <= sg < heavy_seg[lo+1]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (g.heavy_seg[mid] <= sg) lo = mid; else hi = mid;
  }
  const int64_t row = g.heavy_row[lo] - 1;
  if (row < g.r_lo || row >= g.r_hi || !mask_on(mask, mdt, row)) return;
  const int64_t nseg = g.heavy_seg[lo + 1] - g.heavy_seg[lo];
  const int64_t s = g.heavy_ptr[lo] + (sg - g.heavy_seg[lo]) * g.seg_len;
  const int64_t e = min(s + g.seg_len, g.heavy_ptr[lo + 1]);
  const TA* __restrict__ hv = (const TA*)g.heavy_val;
  A acc = T::MO::neutral();
  for (int64_t i0 = s + threadIdx.x; i0 < e; i0 += kH * kBlock) {
    int32_t c[kH];
    TA v[kH];
    TX xv[kH];
#pragma unroll
    for (int u = 0; u < kH; ++u) {
      const int64_t i = i0 + (int64_t)u * kBlock;
      c[u] = i < e ? __builtin_nontemporal_load(&g.heavy_col[i]) : 1;  // column 1: valid, unused
      v[u] = i < e ? __builtin_nontemporal_load(&hv[i]) : TA(0);
    }
#pragma unroll
    for (int u = 0; u < kH; ++u) xv[u] = x[c[u] - 1];
#pragma unroll
    for (int u = 0; u < kH; ++u)
      if (i0 + (int64_t)u * kBlock < e)
        acc = T::MO::op(acc, term<S, TA, TX>(v[u], xv[u], mul_col<S>(g, c[u])));
  }
  acc = block_reduce<typename T::MO>(acc, lds);
  if (threadIdx.x == 0) {
    if (nseg > 1) ((A*)g.partial)[sg] = acc;
    else if (g.stage) ((typename T::CT*)g.stage)[g.slot_base + lo] = typename T::CT(acc);
    else store_acc<typename T::TC, typename T::CT>(c, row, accum, acc);
  }
}

// Folds the partials of multi-segment heavy rows in segment order.
template <class S, class TA, class TX>
__global__ __launch_bounds__(kBlock) void heavy_combine_kernel(
    typename Types<S, TA, TX>::TC* __restrict__ c, SellArgs g, int accum,
    const void* __restrict__ mask, int mdt) {
  typedef Types<S, TA, TX> T;
  typedef typename T::A A;
  const int64_t h = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (h >= g.n_heavy) return;
  const int64_t s0 = g.heavy_seg[h], s1 = g.heavy_seg[h + 1];
  if (s1 - s0 < 2) return;
  const int64_t row = g.heavy_row[h] - 1;
  if (row < g.r_lo || row >= g.r_hi || !mask_on(mask, mdt, row)) return;
  const A* p = (const A*)g.partial;
  A acc = T::MO::neutral();
  for (int64_t j = s0; j < s1; ++j) acc = T::MO::op(acc, p[j]);
  if (g.stage) ((typename T::CT*)g.stage)[g.slot_base + h] = typename T::CT(acc);
  else store_acc<typename T::TC, typename T::CT>(c, row, accum, acc);
}

// The staged values (internal ids) to the rows they belong to: one thread per
// original row in the range whose mask is on, so c is read and written
// coalesced and the staged values are gathered (the same rows the SpMV
// kernels computed; accum is applied here, as store_acc would have).
template <class TC, class CT>
__global__ __launch_bounds__(kBlock) void stage_to_rows_kernel(TC* __restrict__ c,
                                                               const CT* __restrict__ stage,
                                                               const int32_t* __restrict__ vinv,
                                                               int64_t r_lo, int64_t r_hi,
                                                               int accum, const void* __restrict__ mask,
                                                               int mdt) {
  const int64_t row = r_lo + (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (row >= r_hi || !mask_on(mask, mdt, row)) return;
  store_acc<TC, CT>(c, row, accum, stage[vinv[row] - 1]);
}

// ------------------------------------------------------------------ CSC
template <class T>
__global__ __launch_bounds__(kBlock) void csc_spmv_kernel(T* __restrict__ c,
                                                           const int32_t* __restrict__ colptr,
                                                           const int32_t* __restrict__ rowval,
                                                           const T* __restrict__ val,
                                                           const T* __restrict__ b, int64_t n) {
  const int64_t col = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (col >= n) return;
  const T bc = b[col];
  for (int32_t i = colptr[col] - 1; i < colptr[col + 1] - 1; ++i) {
    const T v = jmul(bc, val[i]);
    if constexpr (std::is_same<T, int32_t>::value) {
      atomicAdd((int*)&c[rowval[i] - 1], (int)v);
    } else {
      atomicAdd(&c[rowval[i] - 1], v);
    }
  }
}

static int ilog2_exact(int64_t v) {
  if (v <= 0 || (v & (v - 1))) return -1;
  int l = 0;
  while ((1LL << l) < v) ++l;
  return l;
}

int launch_spmv(gg_ctx* ctx, void* C, int c_dtype, const gg_matrix* A, const void* B,
                int b_dtype, int add, int mul, int accum, const void* mask, int mask_dtype,
                int64_t r0, int64_t nrows, bool has_range, const int32_t* rows) {
  if (nrows <= 0) return GG_OK;
  return dispatch(add, mul, A->val_dtype, b_dtype, [&](auto s, auto ta, auto tx) -> int {
    typedef decltype(s) S;
    typedef typename decltype(ta)::type TA;
    typedef typename decltype(tx)::type TX;
    typedef Types<S, TA, TX> T;
    typedef typename T::TC TC;
    if (c_dtype != code_of<TC>())
      return set_error(GG_ERR_ARGUMENT, "Result type must be able to hold the result of the multiplication");
    if (!accum_ok(accum, std::is_floating_point<TC>::value))
      return set_error(GG_ERR_UNSUPPORTED, "accum operator %d not supported for this type", accum);
    if (A->format == GG_CSR) {
      int lg = csr_log_lanes(ctx, A);
      if (S::ADD == GG_OP_ADD && is_int<typename T::CT>()) lg = 0;  // exact absorb position
      // mask without range only (_dense_mask_gpu_spmv!, src/spmv/CSR.jl:191-205);
      // range + mask runs the generic kernel, src/spmv/CSR.jl:338-353
      const bool any_fast = S::ADD == GG_OP_ANY && mask != nullptr && !has_range && !rows;
      const int64_t groups = nrows;
      const int64_t threads = groups << lg;
      unsigned nb = blocks_for(threads, kBlock);
      const unsigned cap = (unsigned)ctx->num_cus * 64;
      if (nb > cap) nb = cap;
      {
        HotTimer ht(ctx);
        hipLaunchKernelGGL((csr_spmv_kernel<S, TA, TX>), dim3(nb), dim3(kBlock), 0, ctx->stream,
                           (TC*)C, A->ptr, A->idx, (const TA*)A->val, (const TX*)B, accum, mask,
                           mask_dtype, r0, nrows, lg, (int)any_fast, rows, A->m);
      }
      GG_HIP(hipGetLastError());
      return GG_OK;
    }
    // SELL
    const TX* xin = (const TX*)B;
    if (A->flags & GG_FLAG_INTERNAL_IDS) {  // columns are internal ids: gather x first
      const int64_t n_int = A->nslices * A->slice_size + A->n_heavy;
      void* xp;
      GG_TRY(ws_get(ctx, WS_XPERM, sizeof(TX) * (size_t)n_int, &xp));
      hipLaunchKernelGGL(gather_internal_kernel<TX>, dim3(blocks_for(n_int, kBlock)), dim3(kBlock),
                         0, ctx->stream, (const TX*)B, A->perm, A->nslices * A->slice_size,
                         A->heavy_row, A->n_heavy, A->m, (int64_t)1, 1, (TX*)xp);
      xin = (const TX*)xp;
    }
    SellArgs g{};
    g.slice_ptr = A->ptr;
    g.colval = A->idx;
    g.val = A->val;
    g.perm = A->perm;
    g.slot_len = A->slot_len;
    g.C = A->slice_size;
    g.logC = ilog2_exact(A->slice_size);
    g.n_slots = sell_slots(A);
    g.m = A->m;
    g.internal = (A->flags & GG_FLAG_INTERNAL_IDS) ? 1 : 0;
    g.slot_base = A->nslices * A->slice_size;
    g.r_lo = r0;
    g.r_hi = r0 + nrows;
    g.n_sell_blocks = (g.n_slots + kBlock - 1) / kBlock;
    g.n_heavy = A->n_heavy;
    int64_t n_seg = 0;
    if (A->n_heavy > 0) {
      n_seg = A->heavy_nseg;
      g.heavy_row = A->heavy_row;
      g.heavy_ptr = A->heavy_ptr;
      g.heavy_seg = A->heavy_seg;
      g.seg_len = A->heavy_seg_len;
      g.heavy_col = A->heavy_col;
      g.heavy_val = A->heavy_val;
      GG_TRY(ws_get(ctx, WS_PARTIAL, sizeof(typename T::A) * (size_t)n_seg, &g.partial));
    }
    if (g.internal) {
      void* st;
      GG_TRY(ws_get(ctx, WS_SPMSPV2, sizeof(typename T::CT) * (size_t)(g.slot_base + A->n_heavy),
                    &st));
      g.stage = st;
    }
    const int64_t nb = g.n_sell_blocks + n_seg;
    if (nb > 0x7fffffff) return set_error(GG_ERR_ARGUMENT, "SELL launch too large");
    {
      HotTimer ht(ctx);
      hipLaunchKernelGGL((sell_spmv_kernel<S, TA, TX>), dim3((unsigned)nb), dim3(kBlock), 0,
                         ctx->stream, (TC*)C, xin, g, accum, mask, mask_dtype);
    }
    GG_HIP(hipGetLastError());
    if (n_seg > A->n_heavy) {
      hipLaunchKernelGGL((heavy_combine_kernel<S, TA, TX>), dim3(blocks_for(A->n_heavy, kBlock)),
                         dim3(kBlock), 0, ctx->stream, (TC*)C, g, accum, mask, mask_dtype);
      GG_HIP(hipGetLastError());
    }
    if (g.stage && nrows > 0) {
      hipLaunchKernelGGL((stage_to_rows_kernel<TC, typename T::CT>),
                         dim3(blocks_for(nrows, kBlock)), dim3(kBlock), 0, ctx->stream, (TC*)C,
                         (const typename T::CT*)g.stage, A->vinv, g.r_lo, g.r_hi, accum, mask,
                         mask_dtype);
      GG_HIP(hipGetLastError());
    }
    return GG_OK;
  });
}

}  // namespace gg

using namespace gg;

extern "C" int gg_spmv(gg_ctx* ctx, void* C, int c_dtype, int64_t c_len, const gg_matrix* A,
                       const void* B, int b_dtype, int64_t b_len, int add, int mul, int accum,
                       const void* mask, int mask_dtype, int64_t mask_len, int64_t range_start,
                       int64_t range_len) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  clear_error();
  GG_TRY(check_matrix(A));
  // _validate_args order, src/spmv/CSR.jl:374-402
  if (A->n != b_len) return set_error(GG_ERR_DIMENSION, "Matrix dimensions must agree");
  if (c_len != A->m) return set_error(GG_ERR_DIMENSION, "Matrix dimensions must agree");
  if (A->format == GG_CSC) {
    // gpu_spmv!(C, ::SparseGPUMatrixCSC, B) has neither mask nor range, src/spmv/CSC.jl:23-37
    if (mask || range_len >= 0)
      return set_error(GG_ERR_ARGUMENT, "CSC SpMV takes neither mask nor range");
    if (A->val_dtype != b_dtype || c_dtype != b_dtype)
      return set_error(GG_ERR_ARGUMENT, "CSC SpMV needs one element type for A, B and C");
    if (A->n == 0) return GG_OK;
    timer_begin(ctx);
    int st = with_type(c_dtype, [&](auto t) -> int {
      typedef typename decltype(t)::type T;
      if constexpr (std::is_same<T, float>::value || std::is_same<T, double>::value ||
                    std::is_same<T, int32_t>::value) {
        hipLaunchKernelGGL(csc_spmv_kernel<T>, dim3(blocks_for(A->n, kBlock)), dim3(kBlock), 0,
                           ctx->stream, (T*)C, A->ptr, A->idx, (const T*)A->val, (const T*)B, A->n);
        GG_HIP(hipGetLastError());
        return GG_OK;
      } else {
        return set_error(GG_ERR_UNSUPPORTED, "CSC SpMV supports Float32, Float64, Int32");
      }
    });
    timer_end(ctx);
    return st;
  }
  if (mask && mask_len != A->m)
    return set_error(GG_ERR_DIMENSION, "Mask length must be equal to the number of rows in A");
  if (mask && (dtype_size(mask_dtype) == 0 || mask_dtype == GG_UINT64))
    return set_error(GG_ERR_ARGUMENT, "mask element type must be Bool or an integer/float code");
  int64_t r0 = 0, nrows = A->m;
  if (range_len >= 0) {
    if (range_start < 1 || range_start + range_len - 1 > A->m)
      return set_error(GG_ERR_BOUNDS, "range %lld:%lld outside 1:%lld", (long long)range_start,
                       (long long)(range_start + range_len - 1), (long long)A->m);
    r0 = range_start - 1;
    nrows = range_len;
  }
  timer_begin(ctx);
  int st = launch_spmv(ctx, C, c_dtype, A, B, b_dtype, add, mul, accum, mask, mask_dtype, r0, nrows,
                       range_len >= 0);
  timer_end(ctx);
  if (st == GG_ERR_UNSUPPORTED && gg_last_error()[0] == 0)
    set_error(st, "no kernel for add=%d mul=%d with value types %d, %d", add, mul, A->val_dtype, b_dtype);
  return st;
}

// The sparse-mask SpMV of sparse_masked_csr_spmv_kernel! (src/spmv/CSR.jl:90-115):
// C[r] = accum(C[r], fold of row r) for each row r of the list.
extern "C" int gg_spmv_rows(gg_ctx* ctx, void* C, int c_dtype, int64_t c_len, const gg_matrix* A,
                            const void* B, int b_dtype, int64_t b_len, int add, int mul,
                            int accum, const int32_t* rows, int64_t n_rows) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  clear_error();
  GG_TRY(check_matrix(A));
  if (A->n != b_len) return set_error(GG_ERR_DIMENSION, "Matrix dimensions must agree");
  if (c_len != A->m) return set_error(GG_ERR_DIMENSION, "Matrix dimensions must agree");
  if (A->format != GG_CSR)
    return set_error(GG_ERR_UNSUPPORTED, "a row-list mask is a CSR kernel (src/spmv/CSR.jl:90-115)");
  if (n_rows < 0 || (n_rows > 0 && !rows)) return set_error(GG_ERR_ARGUMENT, "bad row list");
  timer_begin(ctx);
  int st = launch_spmv(ctx, C, c_dtype, A, B, b_dtype, add, mul, accum, nullptr, 0, 0, n_rows,
                       false, rows);
  timer_end(ctx);
  if (st == GG_ERR_UNSUPPORTED && gg_last_error()[0] == 0)
    set_error(st, "no kernel for add=%d mul=%d with value types %d, %d", add, mul, A->val_dtype, b_dtype);
  return st;
}
