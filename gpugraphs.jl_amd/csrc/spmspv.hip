// spmspv.hip -- masked sparse-matrix x sparse-vector product (gg_spmspv).
//
// The reference's csc_spmspv_kernel! (src/spmspv.jl:2-21) does not compile
// (its body names a_row_ptr and b, which are undefined), so the semantics are
// defined here: for each stored b_j and each stored A[i,j] of column j,
// c_i = add(c_i, mul(A[i,j], b_j, B_i = j)) from the monoid neutral, over the
// rows the mask admits; the result holds the rows that received a
// contribution, indices ascending.
//
// Expansion: a group of lanes per stored b_j walks column j of A (CSC) and
// folds each product into a dense accumulator with a compare-and-swap loop
// applying the monoid exactly (min / max / or / and / any are order-free; plus
// over floats is the one order-dependent case).  A touched-row bitmap is then
// compacted with an ordered two-level scan, so the output is sorted.
#include "frontier.hpp"

namespace gg {

constexpr int kS = 256;


// Apply acc[i] = op(acc[i], v) atomically for 4- and 8-byte accumulators.
template <class MO, class A> __device__ __forceinline__ void atomic_fold(A* p, A v) {
  if constexpr (sizeof(A) == 8) {
    unsigned long long* q = (unsigned long long*)p;
    unsigned long long old = *q, assumed;
    do {
      assumed = old;
      A cur = *reinterpret_cast<A*>(&assumed);
      A nv = MO::op(cur, v);
      unsigned long long nb = *reinterpret_cast<unsigned long long*>(&nv);
      if (nb == assumed) return;
      old = atomicCAS(q, assumed, nb);
    } while (old != assumed);
  } else if constexpr (sizeof(A) == 4) {
    unsigned int* q = (unsigned int*)p;
    unsigned int old = *q, assumed;
    do {
      assumed = old;
      A cur = *reinterpret_cast<A*>(&assumed);
      A nv = MO::op(cur, v);
      unsigned int nb = *reinterpret_cast<unsigned int*>(&nv);
      if (nb == assumed) return;
      old = atomicCAS(q, assumed, nb);
    } while (old != assumed);
  } else {  // one-byte Bool inside its 32-bit word
    unsigned int* q = (unsigned int*)((uintptr_t)p & ~(uintptr_t)3);
    const int sh = ((uintptr_t)p & 3) * 8;
    unsigned int old = *q, assumed;
    do {
      assumed = old;
      A cur = (A)((assumed >> sh) & 0xffu);
      A nv = MO::op(cur, v);
      unsigned int nb = (assumed & ~(0xffu << sh)) | ((unsigned int)(uint8_t)nv << sh);
      if (nb == assumed) return;
      old = atomicCAS(q, assumed, nb);
    } while (old != assumed);
  }
}

template <class A> __global__ void fill_acc_kernel(A* p, int64_t n, A v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <class S, class TA, class TX>
__global__ __launch_bounds__(kS) void spmspv_expand_kernel(
    const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
    const TA* __restrict__ val, const int32_t* __restrict__ b_idx, const TX* __restrict__ b_val,
    int64_t b_nnz, int64_t ncols, int logL, const void* __restrict__ mask, int mdt, int comp,
    typename AccOf<S::ADD, typename CompT<S, TA, TX>::type>::type* acc, uint32_t* touched) {
  typedef typename CompT<S, TA, TX>::type CT;
  typedef typename AccOf<S::ADD, CT>::type A;
  typedef Monoid<S::ADD, A> MO;
  const int L = 1 << logL;
  const int lane = threadIdx.x & (L - 1);
  const int64_t G = ((int64_t)gridDim.x * kS) >> logL;
  for (int64_t q = ((int64_t)blockIdx.x * kS + threadIdx.x) >> logL; q < b_nnz; q += G) {
    const int32_t j = b_idx[q];
    if (j < 1 || j > ncols) continue;  // indices outside 1:n contribute nothing
    const CT bj = CT(b_val[q]);
    for (int32_t i = colptr[j - 1] - 1 + lane; i < colptr[j] - 1; i += L) {
      const int32_t r = rowval[i] - 1;
      bool on = mask_on(mask, mdt, r);
      if (mask && comp) on = !on;
      if (!on) continue;
      const A v = A(Mul<S::MUL, CT>::ap(CT(val[i]), bj, j));
      atomic_fold<MO, A>(&acc[r], v);
      const uint32_t bit = 1u << (r & 31);
      if (!(touched[r >> 5] & bit)) atomicOr(&touched[r >> 5], bit);
    }
  }
}

// Ordered compaction: per-block popcounts, then block offsets, then writes.
__global__ __launch_bounds__(kS) void block_popcount_kernel(const uint32_t* __restrict__ bm,
                                                            int64_t words,
                                                            unsigned int* __restrict__ sums) {
  __shared__ unsigned int red[kS / 64];
  const int64_t w = (int64_t)blockIdx.x * kS + threadIdx.x;
  unsigned int c = w < words ? (unsigned)__popc(bm[w]) : 0u;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = 0;
    for (int i = 0; i < kS / 64; ++i) t += red[i];
    sums[blockIdx.x] = t;
  }
}

__global__ void exclusive_scan_small_kernel(unsigned int* sums, int64_t n, unsigned int* total) {
  // one thread; n = number of blocks (words / 256), small
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned acc = 0;
  for (int64_t i = 0; i < n; ++i) {
    unsigned v = sums[i];
    sums[i] = acc;
    acc += v;
  }
  *total = acc;
}

template <class A, class TC>
__global__ __launch_bounds__(kS) void compact_kernel(const uint32_t* __restrict__ bm, int64_t words,
                                                     const unsigned int* __restrict__ offs,
                                                     const A* __restrict__ acc,
                                                     int32_t* __restrict__ out_idx,
                                                     TC* __restrict__ out_val) {
  __shared__ unsigned int wsum[kS / 64];
  const int64_t w = (int64_t)blockIdx.x * kS + threadIdx.x;
  const int wl = threadIdx.x & 63;
  uint32_t bits = w < words ? bm[w] : 0u;
  const int c = __popc(bits);
  int incl = c;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (wl >= o) incl += y;
  }
  if (wl == 63) wsum[threadIdx.x >> 6] = incl;
  __syncthreads();
  unsigned base = offs[blockIdx.x];
  for (int i = 0; i < (int)(threadIdx.x >> 6); ++i) base += wsum[i];
  unsigned pos = base + incl - c;
  while (bits) {
    const int b = __ffs(bits) - 1;
    bits &= bits - 1;
    const int64_t r = w * 32 + b;
    out_idx[pos] = (int32_t)(r + 1);
    out_val[pos] = TC(acc[r]);
    ++pos;
  }
}

}  // namespace gg

using namespace gg;

extern "C" int gg_spmspv(gg_ctx* ctx, int32_t* c_nzind, void* c_nzval, int64_t* c_nnz, int c_dtype,
                         int64_t c_len, const gg_matrix* A, const int32_t* b_nzind,
                         const void* b_nzval, int64_t b_nnz, int b_dtype, int64_t b_len, int add,
                         int mul, const void* mask, int mask_dtype, int64_t mask_len,
                         int mask_complement) {
  if (!ctx || !c_nnz) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  GG_TRY(check_matrix(A));
  if (A->format != GG_CSC) return set_error(GG_ERR_ARGUMENT, "SpMSpV expands columns of a CSC matrix");
  if (A->n != b_len || A->m != c_len) return set_error(GG_ERR_DIMENSION, "Matrix dimensions must agree");
  if (mask && mask_len != A->m)
    return set_error(GG_ERR_DIMENSION, "Mask length must be equal to the number of rows in A");
  if (b_nnz < 0 || b_nnz > b_len) return set_error(GG_ERR_ARGUMENT, "b has more entries than its length");
  timer_begin(ctx);
  int st = dispatch(add, mul, A->val_dtype, b_dtype, [&](auto s, auto ta, auto tx) -> int {
    typedef decltype(s) S;
    typedef typename decltype(ta)::type TA;
    typedef typename decltype(tx)::type TX;
    typedef typename CompT<S, TA, TX>::type CT;
    typedef typename AccOf<S::ADD, CT>::type Acc;
    typedef typename OutT<S, TA, TX>::type TC;
    if (c_dtype != code_of<TC>())
      return set_error(GG_ERR_ARGUMENT, "Result type must be able to hold the result of the multiplication");
    const int64_t m = A->m;
    const int64_t words = (m + 31) / 32;
    const int64_t nblk = (words + kS - 1) / kS;
    void *accp, *tp, *sums;
    GG_TRY(ws_get(ctx, WS_SPMSPV0, sizeof(Acc) * std::max<int64_t>(m, 1), &accp));
    GG_TRY(ws_get(ctx, WS_SPMSPV1, sizeof(uint32_t) * std::max<int64_t>(words, 1), &tp));
    GG_TRY(ws_get(ctx, WS_SPMSPV2, sizeof(unsigned) * (nblk + 2), &sums));
    hipLaunchKernelGGL(fill_acc_kernel<Acc>, dim3(blocks_for(m, kS)), dim3(kS), 0, ctx->stream,
                       (Acc*)accp, m, Monoid<S::ADD, Acc>::neutral());
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(words, kS)), dim3(kS), 0, ctx->stream,
                       (uint32_t*)tp, words, 0u);
    if (b_nnz > 0) {
      int lg = 0;
      const double avg = A->n ? (double)A->nnz / (double)A->n : 1.0;
      while (lg < 6 && (double)(1 << lg) < avg) ++lg;
      unsigned nb = blocks_for(b_nnz << lg, kS);
      const unsigned cap = (unsigned)ctx->num_cus * 32;
      if (nb > cap) nb = cap;
      hipLaunchKernelGGL((spmspv_expand_kernel<S, TA, TX>), dim3(nb), dim3(kS), 0, ctx->stream,
                         A->ptr, A->idx, (const TA*)A->val, b_nzind, (const TX*)b_nzval, b_nnz,
                         A->n, lg, mask, mask_dtype, mask_complement, (Acc*)accp, (uint32_t*)tp);
    }
    unsigned* total = (unsigned*)sums + nblk + 1;
    hipLaunchKernelGGL(block_popcount_kernel, dim3((unsigned)std::max<int64_t>(nblk, 1)), dim3(kS), 0,
                       ctx->stream, (const uint32_t*)tp, words, (unsigned*)sums);
    hipLaunchKernelGGL(exclusive_scan_small_kernel, dim3(1), dim3(1), 0, ctx->stream,
                       (unsigned*)sums, nblk, total);
    hipLaunchKernelGGL((compact_kernel<Acc, TC>), dim3((unsigned)std::max<int64_t>(nblk, 1)),
                       dim3(kS), 0, ctx->stream, (const uint32_t*)tp, words, (const unsigned*)sums,
                       (const Acc*)accp, c_nzind, (TC*)c_nzval);
    GG_HIP(hipGetLastError());
    GG_HIP(hipMemcpyAsync(ctx->host_pinned, total, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
    GG_HIP(hipStreamSynchronize(ctx->stream));
    *c_nnz = *(unsigned*)ctx->host_pinned;
    return GG_OK;
  });
  timer_end(ctx);
  if (st == GG_ERR_UNSUPPORTED && gg_last_error()[0] == 0)
    set_error(st, "no kernel for add=%d mul=%d with value types %d, %d", add, mul, A->val_dtype, b_dtype);
  return st;
}
