// spmm.hip -- semiring SpMM with a dense operand (gg_spmm).
//
// A group of Lk lanes (Lk = next power of two >= min(k, 64)) owns one row and
// Lk consecutive columns of B / C; every lane walks the row's entries in
// storage order, so each (row, column) result is accumulated exactly as the
// reference's work-item does, including its absorbing exit:
//   CSR  csr_spmm_kernel!  src/spmm/CSR.jl:16-34 (break after the update),
//   SELL sell_spmm_kernel! src/spmm/SELL.jl:221-242 (break on pad -1 or on
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
      // columns are internal ids: B rows gathered into internal order, row-major
      const int64_t n_int = A->nslices * A->slice_size + A->n_heavy;
      void* bp;
      GG_TRY(ws_get(ctx, WS_XPERM, sizeof(TB) * (size_t)(n_int * k), &bp));
      hipLaunchKernelGGL(gather_internal_kernel<TB>, dim3(blocks_for(n_int * k, kBlk)), dim3(kBlk),
                         0, ctx->stream, (const TB*)B, A->perm, A->nslices * A->slice_size,
                         A->heavy_row, A->n_heavy, A->n, k, g.b_rowmajor, (TB*)bp);
      B = bp;
      g.b_rowmajor = 1;
      g.n = n_int;
      g.internal = 1;
      g.slot_base = A->nslices * A->slice_size;
      g.perm = A->perm;
      g.heavy_row = A->heavy_row;
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
  if (A->format == GG_CSC) return set_error(GG_ERR_UNSUPPORTED, "CSC SpMM is not provided (src/spmm/CSC.jl:1)");
  // src/spmm/CSR.jl:131-133
  if (A->n != b_rows || c_rows != A->m || c_cols != b_cols)
    return set_error(GG_ERR_ASSERTION, "SpMM dimensions must agree");
  if ((c_layout != GG_COL_MAJOR && c_layout != GG_ROW_MAJOR) ||
      (b_layout != GG_COL_MAJOR && b_layout != GG_ROW_MAJOR))
    return set_error(GG_ERR_ARGUMENT, "unknown layout");
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
