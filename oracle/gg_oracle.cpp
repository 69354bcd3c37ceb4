// gg_oracle.cpp -- CPU restatement of GPUGraphs.jl's semiring kernels and
// graph loops.  TEST INFRASTRUCTURE ONLY: the product never links or calls
// this file; tests, the smoke check and the bench's CPU-baseline leg use it as
// the checker.
//
// Each function follows the reference file:line it cites, with the same
// per-row accumulation order, sentinels (typemax, parent 0, pad -1) and
// early exits.  Indices are 1-based Int32 as in the reference structs.
// OpenMP splits rows across threads the way KernelAbstractions' CPU backend
// splits an ndrange.  Built with -ffp-contract=off so a*b+c is never fused.
//
// Every exported function is pinned independently of this restatement by
// tests/test_oracle_pinning.py: the SELL known-answer test and CSR error
// cases of the reference's test/structs.jl (tests/golden/), FP64 products,
// closed forms of masks / ranges / saturation, scipy's BFS and Dijkstra,
// BFS-tree validity and brute force on tiny graphs.

#include <stdint.h>
#include <string.h>
#include <math.h>
#include <limits>
#include <vector>
#include <algorithm>
#include <type_traits>
#include <omp.h>

namespace {

enum { S_OK = 0, S_DIM = 1, S_ARG = 2, S_ASSERT = 3, S_BOUNDS = 4, S_UNSUP = 5,
       S_NOCONV = 8 };
enum { D_BOOL = 1, D_I32 = 2, D_I64 = 3, D_F32 = 4, D_F64 = 5 };
enum { OP_FIRST = 1, OP_SECOND, OP_ANY, OP_PAIR, OP_ADD, OP_SAFE_ADD, OP_MUL, OP_MIN,
       OP_MAX, OP_LOR, OP_LAND, OP_BOR, OP_BAND, OP_SECONDI };
enum { F_CSR = 1, F_SELL = 2, F_CSC = 3 };

typedef uint8_t jbool;  // Julia Bool: one byte, 0 or 1

// ---------------------------------------------------------------- limits
// typemax/typemin as Julia defines them (Bool: true/false, floats: +-Inf).
template <class T> struct J {
  static T tmax() {
    if (std::is_floating_point<T>::value) return std::numeric_limits<T>::infinity();
    return std::numeric_limits<T>::max();
  }
  static T tmin() {
    if (std::is_floating_point<T>::value) return -std::numeric_limits<T>::infinity();
    return std::numeric_limits<T>::min();
  }
};
template <> struct J<jbool> {
  static jbool tmax() { return 1; }
  static jbool tmin() { return 0; }
};

template <class T> inline bool isnan_(T x) {
  if constexpr (std::is_floating_point<T>::value) return x != x; else return false;
}
// Julia's min/max: NaN-propagating, -0.0 < 0.0 (Base.min/max for IEEE floats).
template <class T> inline T jmin(T x, T y) {
  if constexpr (std::is_floating_point<T>::value) {
    if (isnan_(x)) return x;
    if (isnan_(y)) return y;
    if (x < y) return x;
    if (y < x) return y;
    return std::signbit(x) ? x : y;
  } else {
    return x < y ? x : y;
  }
}
template <class T> inline T jmax(T x, T y) {
  if constexpr (std::is_floating_point<T>::value) {
    if (isnan_(x)) return x;
    if (isnan_(y)) return y;
    if (x > y) return x;
    if (y > x) return y;
    return std::signbit(x) ? y : x;
  } else {
    return x > y ? x : y;
  }
}
// Wrapping integer add/mul (Julia Int arithmetic wraps).
template <class T> inline T jadd(T x, T y) {
  if constexpr (std::is_same<T, int32_t>::value) return (int32_t)((uint32_t)x + (uint32_t)y);
  else if constexpr (std::is_same<T, int64_t>::value) return (int64_t)((uint64_t)x + (uint64_t)y);
  else if constexpr (std::is_same<T, jbool>::value) return (jbool)(x | y);  // unused
  else return x + y;
}
template <class T> inline T jmul(T x, T y) {
  if constexpr (std::is_same<T, int32_t>::value) return (int32_t)((uint32_t)x * (uint32_t)y);
  else if constexpr (std::is_same<T, int64_t>::value) return (int64_t)((uint64_t)x * (uint64_t)y);
  else if constexpr (std::is_same<T, jbool>::value) return (jbool)(x & y);  // Bool * Bool = x & y
  else return x * y;
}

// ---------------------------------------------------------------- monoids
// monoid_neutral / monoid_absorb, src/algebra.jl:10-23,36-37,44-45,66-70,78-82,96-100.
template <class T> struct Monoid {
  int op;
  T neutral;
  bool has_absorb;
  T absorb;
};
template <class T> bool make_monoid(int op, Monoid<T>& m) {
  m.op = op;
  m.has_absorb = true;
  switch (op) {
    case OP_ADD: m.neutral = T(0); m.absorb = J<T>::tmax(); return true;        // :44-45
    case OP_MIN: m.neutral = J<T>::tmax(); m.absorb = J<T>::tmin(); return true; // :66-67
    case OP_MAX: m.neutral = J<T>::tmin(); m.absorb = J<T>::tmax(); return true; // :69-70
    case OP_ANY: m.neutral = T(0); m.has_absorb = false; m.absorb = T(0); return true; // :36-37 (absorb = Any never compares equal)
    case OP_BOR:
      if constexpr (std::is_floating_point<T>::value) return false;
      else { m.neutral = T(0); m.absorb = std::is_same<T, jbool>::value ? T(1) : T(~T(0)); return true; } // :96-97
    case OP_BAND:
      if constexpr (std::is_floating_point<T>::value) return false;
      else { m.neutral = std::is_same<T, jbool>::value ? T(1) : T(~T(0)); m.absorb = T(0); return true; } // :99-100
    case OP_LOR:
      if (!std::is_same<T, jbool>::value) return false;
      m.neutral = T(0); m.absorb = T(1); return true;  // :78-79 (Bool only)
    case OP_LAND:
      if (!std::is_same<T, jbool>::value) return false;
      m.neutral = T(1); m.absorb = T(0); return true;  // :81-82 (Bool only)
    default: return false;
  }
}
template <class T> inline T madd(int op, T x, T y) {
  switch (op) {
    case OP_ADD: return jadd(x, y);
    case OP_MIN: return jmin(x, y);
    case OP_MAX: case OP_ANY: return jmax(x, y);  // GPUGraphs_any = max, algebra.jl:35
    case OP_BOR: case OP_LOR:
      if constexpr (std::is_floating_point<T>::value) return x; else return T(x | y);
    case OP_BAND: case OP_LAND:
      if constexpr (std::is_floating_point<T>::value) return x; else return T(x & y);
  }
  return x;
}

// Binary "mul" operators, src/algebra.jl:31-111, evaluated in the compute
// type T = promote_type(Tv, InputType); col is the 1-based column (B_i).
template <class T> inline bool mul_supported(int op) {
  switch (op) {
    case OP_FIRST: case OP_SECOND: case OP_PAIR: case OP_MIN: case OP_MAX: case OP_SECONDI:
      return true;
    case OP_ADD: case OP_SAFE_ADD: case OP_MUL:
      return !std::is_same<T, jbool>::value || op == OP_MUL;
    case OP_BOR: case OP_BAND: case OP_LOR: case OP_LAND:
      return !std::is_floating_point<T>::value;
  }
  return false;
}
template <class T> inline T bmul(int op, T a, T x, int64_t col) {
  switch (op) {
    case OP_FIRST: return a;                                     // :31
    case OP_SECOND: return x;                                    // :32
    case OP_PAIR: return x != T(0) ? T(1) : T(0);                // :40
    case OP_ADD: return jadd(a, x);                              // :43
    case OP_SAFE_ADD: {                                          // :47-48
      T sat = J<T>::tmax();
      return jmax(a, x) == sat ? sat : jadd(a, x);
    }
    case OP_MUL: return jmul(a, x);                              // :55
    case OP_MIN: return jmin(a, x);                              // :65
    case OP_MAX: return jmax(a, x);                              // :68
    case OP_BOR: case OP_LOR:
      if constexpr (std::is_floating_point<T>::value) return a; else return T(a | x);  // :77,95
    case OP_BAND: case OP_LAND:
      if constexpr (std::is_floating_point<T>::value) return a; else return T(a & x);  // :80,98
    case OP_SECONDI: return jmul(T(col), x);                     // :109 B_i * y
  }
  return a;
}

// ---------------------------------------------------------------- dtype glue
template <class T> constexpr int dcode();
template <> constexpr int dcode<jbool>() { return D_BOOL; }
template <> constexpr int dcode<int32_t>() { return D_I32; }
template <> constexpr int dcode<int64_t>() { return D_I64; }
template <> constexpr int dcode<float>() { return D_F32; }
template <> constexpr int dcode<double>() { return D_F64; }

// promote_type for the pairs the reference's paths use.
int promote(int a, int b) {
  if (a == b) return a;
  if (a > b) std::swap(a, b);  // a < b in code order BOOL<I32<I64<F32<F64
  if (b == D_F64) return D_F64;
  if (b == D_F32) return D_F32;  // Julia: promote_type(Int32, Float32) = Float32
  return b;                      // Bool < Int32 < Int64
}

inline bool mask_on(const void* mask, int mdt, int64_t r) {
  if (!mask) return true;
  switch (mdt) {
    case D_BOOL: return ((const jbool*)mask)[r] != 0;
    case D_I32: return ((const int32_t*)mask)[r] != 0;
    case D_I64: return ((const int64_t*)mask)[r] != 0;
    case D_F32: return ((const float*)mask)[r] != 0.0f;
    case D_F64: return ((const double*)mask)[r] != 0.0;
  }
  return true;
}

template <class T> inline T load_as(const void* p, int dt, int64_t i) {
  switch (dt) {
    case D_BOOL: return T(((const jbool*)p)[i]);
    case D_I32: return T(((const int32_t*)p)[i]);
    case D_I64: return T(((const int64_t*)p)[i]);
    case D_F32: return T(((const float*)p)[i]);
    case D_F64: return T(((const double*)p)[i]);
  }
  return T(0);
}

}  // namespace

// ---------------------------------------------------------------- matrix view
struct orc_matrix {
  int32_t format, val_dtype;
  int64_t m, n, nnz;
  const int32_t* ptr;
  const int32_t* idx;
  const void* val;
  int64_t slice_size, nslices, n_stored;
};

namespace {

// accum(c, acc), evaluated in T; the reference calls accum(c[row], acc, ...)
// (src/spmv/CSR.jl:53).  Supported: first, second, add, mul, min, max,
// bor/lor, band/land.
template <class T> inline bool accum_supported(int op) {
  switch (op) {
    case OP_FIRST: case OP_SECOND: case OP_ADD: case OP_MUL: case OP_MIN: case OP_MAX:
      return true;
    case OP_BOR: case OP_BAND: case OP_LOR: case OP_LAND:
      return !std::is_floating_point<T>::value;
  }
  return false;
}
template <class T> inline T accum_apply(int op, T c, T acc) {
  return bmul<T>(op, c, acc, 0);
}

// -------------------------------------------------------- SpMV, one row
// The per-row loop of csr_spmv_kernel! src/spmv/CSR.jl:44-53 (and the masked
// / range kernels :132-143, :238-247, :301-312): acc starts at the monoid
// neutral, break when acc == absorb.
template <class T, class TA>
inline T csr_row(const orc_matrix& A, int64_t row, const T* b, const Monoid<T>& mo, int mul) {
  T acc = mo.neutral;
  const TA* av = (const TA*)A.val;
  for (int64_t i = A.ptr[row - 1]; i <= A.ptr[row] - 1; ++i) {
    int64_t col = A.idx[i - 1];
    acc = madd<T>(mo.op, acc, bmul<T>(mul, T(av[i - 1]), b[col - 1], col));
    if (mo.has_absorb && acc == mo.absorb) break;
  }
  return acc;
}

// The per-row loop of sell_spmv_kernel! src/spmv/SELL.jl:47-61: lane stride
// slice_size, stop at col == -1, no absorb exit.
template <class T, class TA>
inline T sell_row(const orc_matrix& A, int64_t row, const T* b, const Monoid<T>& mo, int mul) {
  const int64_t C = A.slice_size;
  const int64_t slice = (row - 1) / C + 1;
  const int64_t offset = (row - 1) % C;
  const int64_t start = A.ptr[slice - 1] + offset;
  const int64_t stop = A.ptr[slice] - 1;
  const TA* av = (const TA*)A.val;
  T acc = mo.neutral;
  for (int64_t i = start; i <= stop; i += C) {
    int64_t col = A.idx[i - 1];
    if (col == -1) break;
    acc = madd<T>(mo.op, acc, bmul<T>(mul, T(av[i - 1]), b[col - 1], col));
  }
  return acc;
}

template <class T, class TA, class TC>
int spmv_typed(TC* c, const orc_matrix& A, const void* bv, int bdt, int add, int mul, int accum,
               const void* mask, int mdt, int64_t r0, int64_t nrows, int nthreads, bool fast) {
  // B converted to the compute type once (promote_type(Tv, InputType)).
  const int64_t n = A.n;
  std::vector<T> b(n);
  for (int64_t j = 0; j < n; ++j) b[j] = load_as<T>(bv, bdt, j);
  Monoid<T> mo;
  if (!make_monoid<T>(add, mo)) return S_UNSUP;
  if (!mul_supported<T>(mul) || !accum_supported<T>(accum)) return S_UNSUP;
  // any_dense_masked_csr_spmv_kernel! runs only for a mask without a range
  // (_dense_mask_gpu_spmv!, src/spmv/CSR.jl:191-205); range + mask takes the
  // generic kernel (src/spmv/CSR.jl:338-353) where any = max.
  const bool any_fast = fast && (add == OP_ANY) && mask && A.format == F_CSR;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (int64_t k = 0; k < nrows; ++k) {
    int64_t row = r0 + k;
    if (!mask_on(mask, mdt, row - 1)) continue;
    if (any_fast) {
      // any_dense_masked_csr_spmv_kernel! src/spmv/CSR.jl:159-170: the first
      // entry whose b value is nonzero writes accum(c, mul(...)) and stops;
      // a row without one is left untouched.
      const TA* av = (const TA*)A.val;
      for (int64_t i = A.ptr[row - 1]; i <= A.ptr[row] - 1; ++i) {
        int64_t col = A.idx[i - 1];
        T bval = b[col - 1];
        if (bval != T(0)) {
          T v = bmul<T>(mul, T(av[i - 1]), bval, col);
          c[row - 1] = (TC)accum_apply<T>(accum, (T)c[row - 1], v);
          break;
        }
      }
      continue;
    }
    T acc = A.format == F_CSR ? csr_row<T, TA>(A, row, b.data(), mo, mul)
                              : sell_row<T, TA>(A, row, b.data(), mo, mul);
    c[row - 1] = (TC)accum_apply<T>(accum, (T)c[row - 1], acc);
  }
  return S_OK;
}

// Dispatch on (value type of A, compute type).  The result array has the
// compute type (the reference requires promote_type(Tv, Tin) <: Tres,
// src/spmv/CSR.jl:381-387); ANY.SECONDI writes the Int32 index type.
template <class T>
int spmv_dispatch_a(void* c, const orc_matrix& A, const void* b, int bdt, int add, int mul,
                    int accum, const void* mask, int mdt, int64_t r0, int64_t nrows, int nth,
                    int cdt, bool fast) {
  if (cdt != dcode<T>()) return S_ARG;
  switch (A.val_dtype) {
    case D_BOOL: return spmv_typed<T, jbool, T>((T*)c, A, b, bdt, add, mul, accum, mask, mdt, r0, nrows, nth, fast);
    case D_I32: return spmv_typed<T, int32_t, T>((T*)c, A, b, bdt, add, mul, accum, mask, mdt, r0, nrows, nth, fast);
    case D_I64: return spmv_typed<T, int64_t, T>((T*)c, A, b, bdt, add, mul, accum, mask, mdt, r0, nrows, nth, fast);
    case D_F32: return spmv_typed<T, float, T>((T*)c, A, b, bdt, add, mul, accum, mask, mdt, r0, nrows, nth, fast);
    case D_F64: return spmv_typed<T, double, T>((T*)c, A, b, bdt, add, mul, accum, mask, mdt, r0, nrows, nth, fast);
  }
  return S_UNSUP;
}

int spmv_any(void* c, int cdt, const orc_matrix& A, const void* b, int bdt, int add, int mul,
             int accum, const void* mask, int mdt, int64_t r0, int64_t nrows, int nth, bool fast) {
  int tc = promote(A.val_dtype, bdt);
  if (mul == OP_SECONDI && cdt == D_I32 && tc != D_F32 && tc != D_F64) tc = D_I32;
  switch (tc) {
    case D_BOOL: return spmv_dispatch_a<jbool>(c, A, b, bdt, add, mul, accum, mask, mdt, r0, nrows, nth, cdt, fast);
    case D_I32: return spmv_dispatch_a<int32_t>(c, A, b, bdt, add, mul, accum, mask, mdt, r0, nrows, nth, cdt, fast);
    case D_I64: return spmv_dispatch_a<int64_t>(c, A, b, bdt, add, mul, accum, mask, mdt, r0, nrows, nth, cdt, fast);
    case D_F32: return spmv_dispatch_a<float>(c, A, b, bdt, add, mul, accum, mask, mdt, r0, nrows, nth, cdt, fast);
    case D_F64: return spmv_dispatch_a<double>(c, A, b, bdt, add, mul, accum, mask, mdt, r0, nrows, nth, cdt, fast);
  }
  return S_UNSUP;
}

// ---------------------------------------------------------------- SpMM
// csr_spmm_kernel! src/spmm/CSR.jl:16-34 (+ range :51-70, row mask :89-108)
// and sell_spmm_kernel! src/spmm/SELL.jl:2-39 (break on -1 or on
// acc == terminal before the entry).  B and C are column-major.
template <class T, class TA>
int spmm_typed(T* Cm, int64_t k, const orc_matrix& A, const void* Bv, int bdt, int add, int mul,
               int accum, const void* mask, int mdt, int64_t r0, int64_t nrows, int nth) {
  Monoid<T> mo;
  if (!make_monoid<T>(add, mo)) return S_UNSUP;
  if (!mul_supported<T>(mul) || !accum_supported<T>(accum)) return S_UNSUP;
  const int64_t n = A.n, m = A.m;
  std::vector<T> B((size_t)n * k);
  for (int64_t j = 0; j < n * k; ++j) B[j] = load_as<T>(Bv, bdt, j);
  const TA* av = (const TA*)A.val;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nth)
  for (int64_t q = 0; q < nrows; ++q) {
    int64_t row = r0 + q;
    if (!mask_on(mask, mdt, row - 1)) continue;
    for (int64_t j = 1; j <= k; ++j) {
      const T* bcol = B.data() + (j - 1) * n;
      T acc = mo.neutral;
      if (A.format == F_CSR) {
        for (int64_t i = A.ptr[row - 1]; i <= A.ptr[row] - 1; ++i) {
          int64_t col = A.idx[i - 1];
          acc = madd<T>(mo.op, acc, bmul<T>(mul, T(av[i - 1]), bcol[col - 1], col));
          if (mo.has_absorb && acc == mo.absorb) break;
        }
      } else {
        const int64_t Cs = A.slice_size;
        const int64_t slice = (row - 1) / Cs + 1, off = (row - 1) % Cs;
        for (int64_t i = A.ptr[slice - 1] + off; i <= A.ptr[slice] - 1; i += Cs) {
          int64_t col = A.idx[i - 1];
          if (col == -1 || (mo.has_absorb && acc == mo.absorb)) break;
          acc = madd<T>(mo.op, acc, bmul<T>(mul, T(av[i - 1]), bcol[col - 1], col));
        }
      }
      T& cr = Cm[(row - 1) + (j - 1) * m];
      cr = accum_apply<T>(accum, cr, acc);
    }
  }
  return S_OK;
}

template <class T>
int spmm_dispatch_a(void* C, int64_t k, const orc_matrix& A, const void* B, int bdt, int add,
                    int mul, int accum, const void* mask, int mdt, int64_t r0, int64_t nrows,
                    int nth) {
  switch (A.val_dtype) {
    case D_BOOL: return spmm_typed<T, jbool>((T*)C, k, A, B, bdt, add, mul, accum, mask, mdt, r0, nrows, nth);
    case D_I32: return spmm_typed<T, int32_t>((T*)C, k, A, B, bdt, add, mul, accum, mask, mdt, r0, nrows, nth);
    case D_I64: return spmm_typed<T, int64_t>((T*)C, k, A, B, bdt, add, mul, accum, mask, mdt, r0, nrows, nth);
    case D_F32: return spmm_typed<T, float>((T*)C, k, A, B, bdt, add, mul, accum, mask, mdt, r0, nrows, nth);
    case D_F64: return spmm_typed<T, double>((T*)C, k, A, B, bdt, add, mul, accum, mask, mdt, r0, nrows, nth);
  }
  return S_UNSUP;
}

}  // namespace

extern "C" {

int orc_num_threads(void) { return omp_get_max_threads(); }

// Inner-constructor validation of SparseGPUMatrixCSR, src/storage/CSR.jl:48-63
// (CSC uses the same order, src/storage/CSC.jl:31-46).
int orc_csr_check(int64_t m, int64_t n, const int64_t* rowptr, int64_t rowptr_len,
                  const int64_t* colval, int64_t colval_len, int64_t nzval_len) {
  if (rowptr_len != m + 1) return S_ARG;                 // :48-50
  if (colval_len != nzval_len) return S_ARG;             // :51-53
  if (colval_len > 0) {                                  // :54-56
    int64_t mx = colval[0], mn = colval[0];
    for (int64_t i = 1; i < colval_len; ++i) { mx = std::max(mx, colval[i]); mn = std::min(mn, colval[i]); }
    if (mx > n || mn < 1) return S_ARG;
  }
  if (rowptr[0] != 1 || rowptr[m] != colval_len + 1) return S_ARG;  // :57-63
  return S_OK;
}

// SELL builder from a CSR, src/storage/SELL.jl:99-181 (reference layout:
// identity perm, slices padded to their longest row, column-major inside a
// slice, pad col -1 / val 0).  Phase 1 returns sizes.
int orc_sell_size(int64_t m, const int32_t* rowptr, int64_t slice_size, int64_t* eff_slice,
                  int64_t* nslices, int64_t* n_stored) {
  if (m == 0) { *eff_slice = slice_size; *nslices = 0; *n_stored = 0; return S_OK; }
  int64_t C = std::min(slice_size, m);                           // :108
  int64_t ns = (m + C - 1) / C;                                  // :109
  int64_t st = 0;
  for (int64_t s = 1; s <= ns; ++s) {                            // :118-124
    int64_t r0 = (s - 1) * C + 1, r1 = std::min(s * C, m), mx = 0;
    for (int64_t r = r0; r <= r1; ++r) mx = std::max<int64_t>(mx, rowptr[r] - rowptr[r - 1]);
    st += mx * C;
  }
  *eff_slice = C; *nslices = ns; *n_stored = st;
  return S_OK;
}

int orc_sell_fill(int64_t m, const int32_t* rowptr, const int32_t* colval, const void* nzval,
                  int val_dtype, int64_t C, int64_t nslices, int32_t* slice_ptr,
                  int32_t* colval_out, void* nzval_out) {
  int esz = val_dtype == D_BOOL ? 1 : (val_dtype == D_I32 || val_dtype == D_F32) ? 4 : 8;
  slice_ptr[0] = 1;                                              // :127
  for (int64_t s = 1; s <= nslices; ++s) {                       // :128-130
    int64_t r0 = (s - 1) * C + 1, r1 = std::min(s * C, m), mx = 0;
    for (int64_t r = r0; r <= r1; ++r) mx = std::max<int64_t>(mx, rowptr[r] - rowptr[r - 1]);
    slice_ptr[s] = (int32_t)(slice_ptr[s - 1] + mx * C);
  }
  for (int64_t s = 1; s <= nslices; ++s) {                       // :133-165
    int64_t base = slice_ptr[s - 1] - 1;
    int64_t width = (slice_ptr[s] - slice_ptr[s - 1]) / C;
    for (int64_t k = 0; k < width; ++k)
      for (int64_t r = 0; r < C; ++r) {
        int64_t pos = base + k * C + r;                          // column-major flatten :158-161
        int64_t row = (s - 1) * C + r + 1;
        int64_t len = row <= m ? rowptr[row] - rowptr[row - 1] : 0;
        if (k < len) {
          int64_t src = rowptr[row - 1] - 1 + k;
          colval_out[pos] = colval[src];
          memcpy((char*)nzval_out + pos * esz, (const char*)nzval + src * esz, esz);
        } else {
          colval_out[pos] = -1;                                  // pad :139-141
          memset((char*)nzval_out + pos * esz, 0, esz);
        }
      }
  }
  return S_OK;
}

// gpu_spmv! for CSR / SELL with its argument checks, src/spmv/CSR.jl:2-29,
// :358-408 and src/spmv/SELL.jl:2-29, :329-379.  range_len < 0 = no range.
int orc_spmv(void* c, int cdt, int64_t c_len, const orc_matrix* A, const void* b, int bdt,
             int64_t b_len, int add, int mul, int accum, const void* mask, int mdt,
             int64_t mask_len, int64_t range_start, int64_t range_len, int nthreads) {
  if (A->n != b_len) return S_DIM;                               // CSR.jl:374-376
  if (c_len != A->m) return S_DIM;                               // :377-379
  if (A->format == F_CSC) {
    // csc_spmv_kernel! src/spmv/CSC.jl:14-20: for each column, c[row] +=
    // mul(b[col], a) atomically (add / accum ignored).  Sequential here.
    if (mask || range_len >= 0) return S_ARG;
    if (cdt != D_F32 && cdt != D_F64 && cdt != D_I32) return S_UNSUP;
    for (int64_t col = 1; col <= A->n; ++col)
      for (int64_t i = A->ptr[col - 1]; i <= A->ptr[col] - 1; ++i) {
        int64_t row = A->idx[i - 1];
        if (cdt == D_F32) ((float*)c)[row - 1] += load_as<float>(b, bdt, col - 1) * load_as<float>(A->val, A->val_dtype, i - 1);
        else if (cdt == D_F64) ((double*)c)[row - 1] += load_as<double>(b, bdt, col - 1) * load_as<double>(A->val, A->val_dtype, i - 1);
        else ((int32_t*)c)[row - 1] = jadd(((int32_t*)c)[row - 1], jmul(load_as<int32_t>(b, bdt, col - 1), load_as<int32_t>(A->val, A->val_dtype, i - 1)));
      }
    return S_OK;
  }
  if (mask && mask_len != A->m) return S_DIM;                    // :400-402
  int64_t r0 = 1, nrows = A->m;
  if (range_len >= 0) {
    if (range_start < 1 || range_start + range_len - 1 > A->m) return S_BOUNDS;
    r0 = range_start; nrows = range_len;
  }
  return spmv_any(c, cdt, *A, b, bdt, add, mul, accum, mask, mdt, r0, nrows, nthreads,
                  range_len < 0);
}

// gpu_spmm! src/spmm/CSR.jl:111-204, src/spmm/SELL.jl:84-124 (col-major B, C).
int orc_spmm(void* C, int cdt, int64_t c_rows, int64_t c_cols, const orc_matrix* A,
             const void* B, int bdt, int64_t b_rows, int64_t b_cols, int add, int mul,
             int accum, const void* mask, int mdt, int64_t mask_len, int64_t range_start,
             int64_t range_len, int nthreads) {
  if (A->n != b_rows || c_rows != A->m || c_cols != b_cols) return S_ASSERT;  // :131-133
  if (A->format == F_SELL && (mask || range_len >= 0)) return S_ARG;
  int64_t r0 = 1, nrows = A->m;
  if (range_len >= 0) {
    if (mask) return S_ASSERT;                                   // :160
    if (range_start < 1 || range_start + range_len - 1 > A->m) return S_ASSERT;  // :161
    r0 = range_start; nrows = range_len;
  }
  if (mask && mask_len != A->m) return S_ASSERT;                 // :183
  int tc = promote(A->val_dtype, bdt);
  if (tc != cdt) return S_ARG;
  switch (tc) {
    case D_BOOL: return spmm_dispatch_a<jbool>(C, c_cols, *A, B, bdt, add, mul, accum, mask, mdt, r0, nrows, nthreads);
    case D_I32: return spmm_dispatch_a<int32_t>(C, c_cols, *A, B, bdt, add, mul, accum, mask, mdt, r0, nrows, nthreads);
    case D_I64: return spmm_dispatch_a<int64_t>(C, c_cols, *A, B, bdt, add, mul, accum, mask, mdt, r0, nrows, nthreads);
    case D_F32: return spmm_dispatch_a<float>(C, c_cols, *A, B, bdt, add, mul, accum, mask, mdt, r0, nrows, nthreads);
    case D_F64: return spmm_dispatch_a<double>(C, c_cols, *A, B, bdt, add, mul, accum, mask, mdt, r0, nrows, nthreads);
  }
  return S_UNSUP;
}

}  // extern "C"

namespace {

template <class T> inline bool any_nonzero(const std::vector<T>& v) {
  for (const T& x : v) if (x != T(0)) return true;
  return false;
}
template <class T> inline bool any_or(const std::vector<T>& v) {  // reduce(|, v) != 0
  T acc = T(0);
  for (const T& x : v) acc = T(acc | x);
  return acc != T(0);
}

// bfs_distances / bfs_distances! / no_mask_bfs_distances!,
// src/algorithms/bfs.jl:1-111.  Tv = value type of A_T; Ti = Int32.
template <class Tv>
int bfs_dist_typed(const orc_matrix& A, int64_t source, int32_t* dist, int use_mask, int nth,
                   int64_t* levels) {
  const int64_t m = A.m;
  std::vector<Tv> curr(m, Tv(0)), next(m, Tv(0));                       // :9-10
  std::vector<int32_t> to_explore(m, 1);                                 // :11
  for (int64_t i = 0; i < m; ++i) dist[i] = std::numeric_limits<int32_t>::max();  // :18
  curr[source - 1] = Tv(1);                                              // :45
  dist[source - 1] = 0;                                                  // :46
  to_explore[source - 1] = 0;                                            // :47
  int32_t iter = 0;
  const int dt = dcode<Tv>();
  if (use_mask) {
    while (true) {                                                       // :50
      iter += 1;
      std::fill(next.begin(), next.end(), Tv(0));                        // :49 / :70
      int s = orc_spmv(next.data(), dt, m, &A, curr.data(), dt, m, OP_BOR, OP_BAND, OP_SECOND,
                       to_explore.data(), D_I32, m, 0, -1, nth);         // :52-59
      if (s) return s;
      if (!any_or(next)) break;                                          // :60-62
      for (int64_t v = 0; v < m; ++v) {
        if (next[v] == Tv(1)) dist[v] = iter;                            // :64
        to_explore[v] = to_explore[v] & (next[v] == Tv(0) ? 1 : 0);      // :67
      }
      curr = next;                                                       // :69
    }
  } else {
    while (true) {                                                       // :95
      iter += 1;
      std::fill(next.begin(), next.end(), Tv(0));
      int s = orc_spmv(next.data(), dt, m, &A, curr.data(), dt, m, OP_BOR, OP_BAND, OP_SECOND,
                       nullptr, 0, 0, 0, -1, nth);                       // :97
      if (s) return s;
      for (int64_t v = 0; v < m; ++v) curr[v] = Tv(next[v] & Tv(to_explore[v]));  // :98
      if (!any_or(curr)) break;                                          // :99-101
      for (int64_t v = 0; v < m; ++v) {
        if (curr[v] == Tv(1)) dist[v] = iter;                            // :103
        to_explore[v] = to_explore[v] & (curr[v] == Tv(0) ? 1 : 0);      // :106
      }
    }
  }
  if (levels) *levels = iter - 1;
  return S_OK;
}

// bfs_parents / bfs_parents!, src/algorithms/bfs.jl:113-183.
template <class Tv>
int bfs_parents_typed(const orc_matrix& A, int64_t source, int32_t* parents, int64_t batch,
                      int nth) {
  const int64_t m = A.m;
  std::vector<Tv> curr(m, Tv(0)), to_explore(m, Tv(1));                  // :121,123
  std::vector<int32_t> next(m, 0);                                       // :122
  for (int64_t i = 0; i < m; ++i) parents[i] = std::numeric_limits<int32_t>::max();  // :129
  curr[source - 1] = Tv(1);                                              // :150
  parents[source - 1] = 0;                                               // :151
  to_explore[source - 1] = Tv(0);                                        // :152
  int64_t iter = 0;
  const int dt = dcode<Tv>();
  while (true) {                                                         // :156
    iter += 1;
    int s = orc_spmv(next.data(), D_I32, m, &A, curr.data(), dt, m, OP_ANY, OP_SECONDI,
                     OP_SECOND, to_explore.data(), dt, m, 0, -1, nth);   // :158-165
    if (s) return s;
    for (int64_t v = 0; v < m; ++v) curr[v] = Tv(next[v] != 0 ? 1 : 0);  // :167
    if (iter % batch == 0 && !any_nonzero(curr)) break;                  // :169-174
    for (int64_t v = 0; v < m; ++v) {
      if (curr[v] != Tv(0)) parents[v] = next[v];                        // :176
      to_explore[v] = Tv(to_explore[v] & Tv(curr[v] == Tv(0) ? 1 : 0));  // :179
      next[v] = 0;                                                       // :180
    }
  }
  return S_OK;
}

// shortest_path (one source), src/algorithms/shortest_path.jl:12-83.
template <class Td>
int sssp_typed(const orc_matrix& A, int64_t source, Td* dist, int nth, int64_t max_iter,
               int64_t* iters) {
  const int64_t m = A.m;
  std::vector<Td> next(m, J<Td>::tmax());                                // :20,28
  for (int64_t i = 0; i < m; ++i) dist[i] = J<Td>::tmax();
  dist[source - 1] = Td(0);                                              // :51
  next[source - 1] = Td(0);                                              // :52
  int64_t iter = 0;
  const int dt = dcode<Td>();
  while (true) {                                                         // :57
    iter += 1;
    int s = orc_spmv(next.data(), dt, m, &A, dist, dt, m, OP_MIN, OP_SAFE_ADD, OP_MIN,
                     nullptr, 0, 0, 0, -1, nth);                         // :61-69
    if (s) return s;
    if (iter % 8 == 0) {                                                 // :71-76
      bool progress = false;
      for (int64_t v = 0; v < m && !progress; ++v) progress = next[v] < dist[v];
      if (!progress) break;
    }
    for (int64_t v = 0; v < m; ++v) dist[v] = next[v];                   // :79
    if (max_iter > 0 && iter >= max_iter) { if (iters) *iters = iter; return S_NOCONV; }
  }
  if (iters) *iters = iter;
  return S_OK;
}

// shortest_path (k sources), src/algorithms/shortest_path.jl:88-161; dist is
// m x k column-major.
template <class Td>
int mssp_typed(const orc_matrix& A, const int32_t* sources, int64_t k, Td* dist, int nth,
               int64_t max_iter, int64_t* iters) {
  const int64_t m = A.m;
  std::vector<Td> next((size_t)m * k, J<Td>::tmax());                   // :97,105
  for (int64_t i = 0; i < m * k; ++i) dist[i] = J<Td>::tmax();
  for (int64_t i = 0; i < k; ++i) {                                      // :130-135
    dist[(sources[i] - 1) + i * m] = Td(0);
    next[(sources[i] - 1) + i * m] = Td(0);
  }
  int64_t iter = 0;
  const int dt = dcode<Td>();
  while (true) {
    iter += 1;
    int s = orc_spmm(next.data(), dt, m, k, &A, dist, dt, m, k, OP_MIN, OP_SAFE_ADD, OP_MIN,
                     nullptr, 0, 0, 0, -1, nth);                         // :142-150
    if (s) return s;
    bool progress = false;                                               // :152-155
    for (int64_t v = 0; v < m * k && !progress; ++v) progress = next[v] < dist[v];
    if (!progress) break;
    for (int64_t v = 0; v < m * k; ++v) dist[v] = next[v];               // :157
    if (max_iter > 0 && iter >= max_iter) { if (iters) *iters = iter; return S_NOCONV; }
  }
  if (iters) *iters = iter;
  return S_OK;
}

}  // namespace

extern "C" {

int orc_bfs_distances(const orc_matrix* A, int64_t source, int32_t* dist, int use_mask,
                      int nthreads, int64_t* levels) {
  if (source < 1 || source > A->m) return S_BOUNDS;
  switch (A->val_dtype) {
    case D_BOOL: return bfs_dist_typed<jbool>(*A, source, dist, use_mask, nthreads, levels);
    case D_I32: return bfs_dist_typed<int32_t>(*A, source, dist, use_mask, nthreads, levels);
    case D_I64: return bfs_dist_typed<int64_t>(*A, source, dist, use_mask, nthreads, levels);
  }
  return S_UNSUP;
}

int orc_bfs_parents(const orc_matrix* A, int64_t source, int32_t* parents, int64_t batch,
                    int nthreads) {
  if (source < 1 || source > A->m) return S_BOUNDS;
  if (batch < 1) return S_ARG;
  switch (A->val_dtype) {
    case D_BOOL: return bfs_parents_typed<jbool>(*A, source, parents, batch, nthreads);
    case D_I32: return bfs_parents_typed<int32_t>(*A, source, parents, batch, nthreads);
  }
  return S_UNSUP;
}

// dist_dtype: Int32 for Bool matrices, else the value type
// (src/algorithms/shortest_path.jl:18).
int orc_shortest_path(const orc_matrix* A, int64_t source, void* dist, int dist_dtype,
                      int nthreads, int64_t max_iter, int64_t* iters) {
  if (source < 1 || source > A->m) return S_BOUNDS;
  switch (dist_dtype) {
    case D_I32: return sssp_typed<int32_t>(*A, source, (int32_t*)dist, nthreads, max_iter, iters);
    case D_I64: return sssp_typed<int64_t>(*A, source, (int64_t*)dist, nthreads, max_iter, iters);
    case D_F32: return sssp_typed<float>(*A, source, (float*)dist, nthreads, max_iter, iters);
    case D_F64: return sssp_typed<double>(*A, source, (double*)dist, nthreads, max_iter, iters);
  }
  return S_UNSUP;
}

int orc_shortest_path_multi(const orc_matrix* A, const int32_t* sources, int64_t k, void* dist,
                            int dist_dtype, int nthreads, int64_t max_iter, int64_t* iters) {
  for (int64_t i = 0; i < k; ++i)
    if (sources[i] < 1 || sources[i] > A->m) return S_BOUNDS;
  switch (dist_dtype) {
    case D_I32: return mssp_typed<int32_t>(*A, sources, k, (int32_t*)dist, nthreads, max_iter, iters);
    case D_I64: return mssp_typed<int64_t>(*A, sources, k, (int64_t*)dist, nthreads, max_iter, iters);
    case D_F32: return mssp_typed<float>(*A, sources, k, (float*)dist, nthreads, max_iter, iters);
    case D_F64: return mssp_typed<double>(*A, sources, k, (double*)dist, nthreads, max_iter, iters);
  }
  return S_UNSUP;
}

// Masked SpMSpV, the product src/spmspv.jl:2-54 intends (its body does not
// compile, so the semantics are this build's): for each stored b_j in
// ascending j and each stored A[i,j] of column j (CSC), c_i = add(c_i,
// mul(A[i,j], b_j, col=j)) starting from the monoid neutral; rows whose mask
// is zero (non-zero when mask_complement) take no contribution.  Output:
// indices that received at least one contribution, ascending.
int orc_spmspv(int32_t* c_nzind, void* c_nzval, int64_t* c_nnz, int cdt, int64_t c_len,
               const orc_matrix* A, const int32_t* b_nzind, const void* b_nzval, int64_t b_nnz,
               int bdt, int64_t b_len, int add, int mul, const void* mask, int mdt,
               int64_t mask_len, int mask_complement) {
  if (A->format != F_CSC) return S_ARG;
  if (A->n != b_len || A->m != c_len) return S_DIM;
  if (mask && mask_len != A->m) return S_DIM;
  int tc = promote(A->val_dtype, bdt);
  if (mul == OP_SECONDI && cdt == D_I32 && tc != D_F32 && tc != D_F64) tc = D_I32;
  if (tc != cdt) return S_ARG;
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    Monoid<T> mo;
    if (!make_monoid<T>(add, mo) || !mul_supported<T>(mul)) return S_UNSUP;
    std::vector<T> acc(A->m, mo.neutral);
    std::vector<uint8_t> hit(A->m, 0);
    for (int64_t q = 0; q < b_nnz; ++q) {
      int64_t j = b_nzind[q];
      T bj = load_as<T>(b_nzval, bdt, q);
      for (int64_t i = A->ptr[j - 1]; i <= A->ptr[j] - 1; ++i) {
        int64_t row = A->idx[i - 1];
        bool on = mask_on(mask, mdt, row - 1);
        if (mask && mask_complement) on = !on;
        if (!on) continue;
        acc[row - 1] = madd<T>(add, acc[row - 1], bmul<T>(mul, load_as<T>(A->val, A->val_dtype, i - 1), bj, j));
        hit[row - 1] = 1;
      }
    }
    int64_t cnt = 0;
    for (int64_t r = 0; r < A->m; ++r)
      if (hit[r]) { c_nzind[cnt] = (int32_t)(r + 1); ((T*)c_nzval)[cnt] = acc[r]; ++cnt; }
    *c_nnz = cnt;
    return S_OK;
  };
  switch (tc) {
    case D_BOOL: return run(jbool(0));
    case D_I32: return run(int32_t(0));
    case D_I64: return run(int64_t(0));
    case D_F32: return run(float(0));
    case D_F64: return run(double(0));
  }
  return S_UNSUP;
}


// ------------------------------------------------------ full-size checkers
// Textbook queue BFS, the independent check of BFS levels at any size (it is
// not the reference's pull loop, bfs.jl:31-73, but computes the same unique
// levels).  A_T is CSR of the transposed adjacency: row v lists the
// in-neighbours u of v, so u -> v is an edge when A_T[v, u] is stored and
// band(A_T[v, u], one) != 0 (the or-and product of bfs.jl:52-59 with the
// frontier's ones: Bool true, or an odd integer).  One serial BFS per source,
// sources spread over the threads; dist is m x k column-major Int32 with
// typemax for unreached vertices (bfs.jl:18).
int orc_bfs_queue_multi(const orc_matrix* A, const int32_t* sources, int64_t k, int32_t* dist,
                        int nthreads) {
  if (A->format != F_CSR || A->m != A->n) return S_ARG;
  if (A->val_dtype != D_BOOL && A->val_dtype != D_I32 && A->val_dtype != D_I64) return S_UNSUP;
  const int64_t m = A->m;
  for (int64_t i = 0; i < k; ++i)
    if (sources[i] < 1 || sources[i] > m) return S_BOUNDS;
  auto active = [&](int64_t e) -> bool {
    switch (A->val_dtype) {
      case D_BOOL: return ((const uint8_t*)A->val)[e] & 1;
      case D_I32: return ((const int32_t*)A->val)[e] & 1;
      default: return ((const int64_t*)A->val)[e] & 1;
    }
  };
  // out-neighbour lists: the transpose of A_T (counting sort by column)
  std::vector<int64_t> optr(m + 1, 0);
  const int64_t nnz = (int64_t)A->ptr[m] - 1;
  for (int64_t e = 0; e < nnz; ++e)
    if (active(e)) optr[A->idx[e]] += 1;
  for (int64_t u = 0; u < m; ++u) optr[u + 1] += optr[u];
  std::vector<int32_t> oidx(optr[m] > 0 ? optr[m] : 1);
  {
    std::vector<int64_t> at(optr.begin(), optr.end() - 1);
    for (int64_t v = 0; v < m; ++v)
      for (int64_t e = A->ptr[v] - 1; e < A->ptr[v + 1] - 1; ++e)
        if (active(e)) oidx[at[A->idx[e] - 1]++] = (int32_t)v;
  }
  const int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nth)
  {
    std::vector<int32_t> queue(m > 0 ? m : 1);
#pragma omp for schedule(dynamic, 1)
    for (int64_t q = 0; q < k; ++q) {
      int32_t* d = dist + q * m;
      for (int64_t v = 0; v < m; ++v) d[v] = std::numeric_limits<int32_t>::max();
      int64_t head = 0, tail = 0;
      const int32_t s = sources[q] - 1;
      d[s] = 0;
      queue[tail++] = s;
      while (head < tail) {
        const int32_t u = queue[head++];
        for (int64_t e = optr[u]; e < optr[u + 1]; ++e) {
          const int32_t v = oidx[e];
          if (d[v] == std::numeric_limits<int32_t>::max()) {
            d[v] = d[u] + 1;
            queue[tail++] = v;
          }
        }
      }
    }
  }
  return S_OK;
}

// Shortest-path certificate for positive edge weights (what the mathematics
// fixes, independent of the Bellman-Ford loop of shortest_path.jl:57-81):
// d is the exact distance vector from s iff d[s] = 0, no edge lowers any
// distance (d[v] <= d[u] + w(u -> v), so d <= the true distance), and every
// other vertex with a finite d has a tight in-edge (d[v] == d[u] + w for some
// u with a finite d[u]; following tight edges back strictly decreases d and
// ends at s, so d is the length of a real path).  Unreached = typemax
// (Inf for floats).  Sums are formed in double, exact for integer-valued
// weights and distances below 2^53.  dist is m x k column-major of dist_dtype;
// bad[q] receives the number of vertices of column q that violate a condition.
int orc_sssp_certificate(const orc_matrix* A, const int32_t* sources, int64_t k,
                         const void* dist, int dist_dtype, int nthreads, int64_t* bad) {
  if (A->format != F_CSR || A->m != A->n) return S_ARG;
  const int64_t m = A->m;
  for (int64_t i = 0; i < k; ++i)
    if (sources[i] < 1 || sources[i] > m) return S_BOUNDS;
  auto dval = [&](int64_t i) -> double {
    switch (dist_dtype) {
      case D_I32: { int32_t x = ((const int32_t*)dist)[i];
        return x == std::numeric_limits<int32_t>::max() ? INFINITY : (double)x; }
      case D_I64: { int64_t x = ((const int64_t*)dist)[i];
        return x == std::numeric_limits<int64_t>::max() ? INFINITY : (double)x; }
      case D_F32: return (double)((const float*)dist)[i];
      default: return ((const double*)dist)[i];
    }
  };
  auto wval = [&](int64_t e) -> double { return load_as<double>(A->val, A->val_dtype, e); };
  if (dist_dtype != D_I32 && dist_dtype != D_I64 && dist_dtype != D_F32 && dist_dtype != D_F64)
    return S_ARG;
  const int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
  for (int64_t q = 0; q < k; ++q) {
    const int64_t off = q * m;
    const int64_t s = sources[q] - 1;
    int64_t nbad = 0;
#pragma omp parallel for num_threads(nth) reduction(+ : nbad) schedule(static, 4096)
    for (int64_t v = 0; v < m; ++v) {
      const double dv = dval(off + v);
      double best = INFINITY;
      for (int64_t e = A->ptr[v] - 1; e < A->ptr[v + 1] - 1; ++e) {
        const double du = dval(off + A->idx[e] - 1);
        const double c = du + wval(e);
        if (c < best) best = c;
      }
      bool ok = dv <= best;
      if (v == s) ok = ok && dv == 0.0;
      else if (std::isfinite(dv)) ok = ok && dv == best;
      if (!ok) nbad += 1;
    }
    bad[q] = nbad;
  }
  return S_OK;
}

}  // extern "C"
