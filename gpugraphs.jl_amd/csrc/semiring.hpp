// semiring.hpp -- device-side operators of src/algebra.jl with Julia semantics.
//
// Operators are evaluated in the compute type TC = promote_type(Tv, Tin) that
// the reference's kernels use for their accumulator (src/spmv/CSR.jl:79-80).
// Julia integer arithmetic wraps, min/max propagate NaN and order -0.0 below
// 0.0, safe_add saturates at typemax (algebra.jl:47-48); the helpers below do
// the same so integer and min-plus results are bit-exact.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "../../include/gpugraphs.h"

namespace gg {

typedef uint8_t jbool;

template <class T> struct Lim;
template <> struct Lim<jbool> {
  __host__ __device__ static constexpr jbool max() { return 1; }
  __host__ __device__ static constexpr jbool min() { return 0; }
};
template <> struct Lim<int32_t> {
  __host__ __device__ static constexpr int32_t max() { return 2147483647; }
  __host__ __device__ static constexpr int32_t min() { return -2147483647 - 1; }
};
template <> struct Lim<int64_t> {
  __host__ __device__ static constexpr int64_t max() { return 9223372036854775807LL; }
  __host__ __device__ static constexpr int64_t min() { return -9223372036854775807LL - 1; }
};
template <> struct Lim<float> {
  __host__ __device__ static constexpr float max() { return __builtin_huge_valf(); }
  __host__ __device__ static constexpr float min() { return -__builtin_huge_valf(); }
};
template <> struct Lim<double> {
  __host__ __device__ static constexpr double max() { return __builtin_huge_val(); }
  __host__ __device__ static constexpr double min() { return -__builtin_huge_val(); }
};

template <class T> __device__ __forceinline__ bool is_nan(T x) {
  if constexpr (std::is_floating_point<T>::value) return x != x; else return false;
}
template <class T> __device__ __forceinline__ T jmin(T x, T y) {
  if constexpr (std::is_floating_point<T>::value) {
    if (is_nan(x)) return x;
    if (is_nan(y)) return y;
    if (x < y) return x;
    if (y < x) return y;
    return __builtin_signbit(x) ? x : y;
  } else {
    return x < y ? x : y;
  }
}
template <class T> __device__ __forceinline__ T jmax(T x, T y) {
  if constexpr (std::is_floating_point<T>::value) {
    if (is_nan(x)) return x;
    if (is_nan(y)) return y;
    if (x > y) return x;
    if (y > x) return y;
    return __builtin_signbit(x) ? y : x;
  } else {
    return x > y ? x : y;
  }
}
template <class T> __device__ __forceinline__ T jadd(T x, T y) {
  if constexpr (std::is_same<T, int32_t>::value) return (int32_t)((uint32_t)x + (uint32_t)y);
  else if constexpr (std::is_same<T, int64_t>::value) return (int64_t)((uint64_t)x + (uint64_t)y);
  else return x + y;
}
template <class T> __device__ __forceinline__ T jmul(T x, T y) {
  if constexpr (std::is_same<T, int32_t>::value) return (int32_t)((uint32_t)x * (uint32_t)y);
  else if constexpr (std::is_same<T, int64_t>::value) return (int64_t)((uint64_t)x * (uint64_t)y);
  else if constexpr (std::is_same<T, jbool>::value) return (jbool)(x & y);
  else return x * y;
}

// ------------------------------------------------------------------ monoids
// Accumulator type: plus over float32 accumulates in float64 (rounded once
// when stored); every other monoid is exact and keeps the compute type.
template <int ADD, class T> struct AccOf { typedef T type; };
template <> struct AccOf<GG_OP_ADD, float> { typedef double type; };

// monoid_neutral / monoid_absorb, src/algebra.jl:36-37,44-45,66-70,78-82,96-100.
template <int ADD, class A> struct Monoid;
template <class A> struct Monoid<GG_OP_ADD, A> {
  static constexpr bool has_absorb = true;
  __host__ __device__ static A neutral() { return A(0); }
  __host__ __device__ static A absorb() { return Lim<A>::max(); }
  __device__ static A op(A x, A y) { return jadd(x, y); }
};
template <class A> struct Monoid<GG_OP_MIN, A> {
  static constexpr bool has_absorb = true;
  __host__ __device__ static A neutral() { return Lim<A>::max(); }
  __host__ __device__ static A absorb() { return Lim<A>::min(); }
  __device__ static A op(A x, A y) { return jmin(x, y); }
};
template <class A> struct Monoid<GG_OP_MAX, A> {
  static constexpr bool has_absorb = true;
  __host__ __device__ static A neutral() { return Lim<A>::min(); }
  __host__ __device__ static A absorb() { return Lim<A>::max(); }
  __device__ static A op(A x, A y) { return jmax(x, y); }
};
template <class A> struct Monoid<GG_OP_ANY, A> {  // max, never absorbs (absorb = Any)
  static constexpr bool has_absorb = false;
  __host__ __device__ static A neutral() { return A(0); }
  __host__ __device__ static A absorb() { return A(0); }
  __device__ static A op(A x, A y) { return jmax(x, y); }
};
template <class A> struct Monoid<GG_OP_BOR, A> {
  static constexpr bool has_absorb = true;
  __host__ __device__ static A neutral() { return A(0); }
  __host__ __device__ static A absorb() { return std::is_same<A, jbool>::value ? A(1) : A(~A(0)); }
  __device__ static A op(A x, A y) { return A(x | y); }
};
template <class A> struct Monoid<GG_OP_BAND, A> {
  static constexpr bool has_absorb = true;
  __host__ __device__ static A neutral() { return std::is_same<A, jbool>::value ? A(1) : A(~A(0)); }
  __host__ __device__ static A absorb() { return A(0); }
  __device__ static A op(A x, A y) { return A(x & y); }
};
template <class A> struct Monoid<GG_OP_LOR, A> {
  static constexpr bool has_absorb = true;
  __host__ __device__ static A neutral() { return A(0); }
  __host__ __device__ static A absorb() { return A(1); }
  __device__ static A op(A x, A y) { return A(x | y); }
};
template <class A> struct Monoid<GG_OP_LAND, A> {
  static constexpr bool has_absorb = true;
  __host__ __device__ static A neutral() { return A(1); }
  __host__ __device__ static A absorb() { return A(0); }
  __device__ static A op(A x, A y) { return A(x & y); }
};

// ------------------------------------------------------------ multiplies
// Binary operators of src/algebra.jl in the compute type T; col is B_i.
template <int MUL, class T> struct Mul;
template <class T> struct Mul<GG_OP_MUL, T> {
  __device__ static T ap(T a, T x, int32_t) { return jmul(a, x); }
};
template <class T> struct Mul<GG_OP_ADD, T> {
  __device__ static T ap(T a, T x, int32_t) { return jadd(a, x); }
};
template <class T> struct Mul<GG_OP_SAFE_ADD, T> {
  __device__ static T ap(T a, T x, int32_t) {
    const T sat = Lim<T>::max();
    return jmax(a, x) == sat ? sat : jadd(a, x);
  }
};
template <class T> struct Mul<GG_OP_BAND, T> {
  __device__ static T ap(T a, T x, int32_t) { return T(a & x); }
};
template <class T> struct Mul<GG_OP_LAND, T> {
  __device__ static T ap(T a, T x, int32_t) { return T(a & x); }
};
template <class T> struct Mul<GG_OP_SECONDI, T> {
  __device__ static T ap(T, T x, int32_t col) { return jmul(T(col), x); }
};

// --------------------------------------------------------------- semirings
// The (add, mul) pairs with kernels.  TC is the compute type (the type of
// the result array); TA / TX the stored value and input vector types.
template <int ADD_, int MUL_> struct SR {
  static constexpr int ADD = ADD_;
  static constexpr int MUL = MUL_;
};

template <class T, class U> struct Promote;
template <class T> struct Promote<T, T> { typedef T type; };
template <> struct Promote<jbool, int32_t> { typedef int32_t type; };
template <> struct Promote<int32_t, jbool> { typedef int32_t type; };

template <class TA, class TX> constexpr bool pair_ok() {
  return (std::is_same<TA, TX>::value) ||
         (std::is_same<TA, jbool>::value && std::is_same<TX, int32_t>::value) ||
         (std::is_same<TA, int32_t>::value && std::is_same<TX, jbool>::value);
}

template <class S, class TA, class TX> struct OutT {
  typedef typename Promote<TA, TX>::type type;
};
template <class TA, class TX> struct OutT<SR<GG_OP_ANY, GG_OP_SECONDI>, TA, TX> {
  typedef int32_t type;
};

template <class T> constexpr bool is_int() {
  return std::is_same<T, int32_t>::value || std::is_same<T, int64_t>::value;
}
template <class T> constexpr bool is_fp() { return std::is_floating_point<T>::value; }

template <class S, class TA, class TX> constexpr bool semiring_ok() {
  if constexpr (!pair_ok<TA, TX>()) {
    return false;
  } else {
    typedef typename Promote<TA, TX>::type TC;
    constexpr int A = S::ADD, M = S::MUL;
    if (A == GG_OP_ADD && M == GG_OP_MUL) return is_fp<TC>() || is_int<TC>();
    if (A == GG_OP_BOR && M == GG_OP_BAND) return is_int<TC>() || std::is_same<TC, jbool>::value;
    if (A == GG_OP_LOR && M == GG_OP_LAND) return std::is_same<TC, jbool>::value;
    if (A == GG_OP_ANY && M == GG_OP_SECONDI)
      return (std::is_same<TA, jbool>::value || std::is_same<TA, int32_t>::value) &&
             (std::is_same<TX, jbool>::value || std::is_same<TX, int32_t>::value);
    if (A == GG_OP_MIN && (M == GG_OP_SAFE_ADD || M == GG_OP_ADD))
      return is_fp<TC>() || is_int<TC>();
    return false;
  }
}

// Compute type used inside the loop: the promoted type, except secondi
// whose product is the Int32 column index times y.
template <class S, class TA, class TX> struct CompT {
  typedef typename Promote<TA, TX>::type type;
};
template <class TA, class TX> struct CompT<SR<GG_OP_ANY, GG_OP_SECONDI>, TA, TX> {
  typedef int32_t type;
};

// accum(c, acc) at the store, src/spmv/CSR.jl:53 (codes of include/gpugraphs.h).
template <class T> __device__ __forceinline__ T accum_apply(int op, T c, T acc) {
  switch (op) {
    case GG_OP_FIRST: return c;
    case GG_OP_SECOND: return acc;
    case GG_OP_ADD: return jadd(c, acc);
    case GG_OP_MUL: return jmul(c, acc);
    case GG_OP_MIN: return jmin(c, acc);
    case GG_OP_MAX: return jmax(c, acc);
    default:
      if constexpr (std::is_floating_point<T>::value) {
        return acc;
      } else {
        if (op == GG_OP_BOR || op == GG_OP_LOR) return T(c | acc);
        if (op == GG_OP_BAND || op == GG_OP_LAND) return T(c & acc);
        return acc;
      }
  }
}
__host__ inline bool accum_ok(int op, bool floating) {
  switch (op) {
    case GG_OP_FIRST: case GG_OP_SECOND: case GG_OP_ADD: case GG_OP_MUL: case GG_OP_MIN:
    case GG_OP_MAX:
      return true;
    case GG_OP_BOR: case GG_OP_BAND: case GG_OP_LOR: case GG_OP_LAND:
      return !floating;
  }
  return false;
}
__device__ __forceinline__ bool accum_reads_c(int op) { return op != GG_OP_SECOND; }

// mask[r] != 0 for the mask element types the reference allows (Integer).
__device__ __forceinline__ bool mask_on(const void* mask, int mdt, int64_t r) {
  if (!mask) return true;
  switch (mdt) {
    case GG_BOOL: return ((const uint8_t*)mask)[r] != 0;
    case GG_INT32: return ((const int32_t*)mask)[r] != 0;
    case GG_INT64: return ((const int64_t*)mask)[r] != 0;
    case GG_FLOAT32: return ((const float*)mask)[r] != 0.0f;
    case GG_FLOAT64: return ((const double*)mask)[r] != 0.0;
  }
  return true;
}

// ------------------------------------------------------ wavefront reductions
template <class T> __device__ __forceinline__ T shfl_xor(T v, int lane_mask) {
  if constexpr (sizeof(T) == 8) {
    int2 p = *reinterpret_cast<int2*>(&v);
    p.x = __shfl_xor(p.x, lane_mask, 64);
    p.y = __shfl_xor(p.y, lane_mask, 64);
    return *reinterpret_cast<T*>(&p);
  } else if constexpr (sizeof(T) == 4) {
    int q = __shfl_xor(*reinterpret_cast<int*>(&v), lane_mask, 64);
    return *reinterpret_cast<T*>(&q);
  } else {
    int q = __shfl_xor((int)v, lane_mask, 64);
    return (T)q;
  }
}

// Reduce across aligned groups of L lanes (L a power of two <= 64).
template <class MO, class A> __device__ __forceinline__ A group_reduce(A v, int L) {
  for (int o = L >> 1; o > 0; o >>= 1) v = MO::op(v, shfl_xor(v, o));
  return v;
}

}  // namespace gg

// ----------------------------------------------------- dispatch machinery
// Calls f(SR<add,mul>{}, TA{}, TX{}) for the compiled combinations; returns
// GG_ERR_UNSUPPORTED otherwise.  Used by every kernel family.
namespace gg {

template <class T> struct Tag { typedef T type; };

template <class F> int with_type(int code, F&& f) {
  switch (code) {
    case GG_BOOL: return f(Tag<jbool>{});
    case GG_INT32: return f(Tag<int32_t>{});
    case GG_INT64: return f(Tag<int64_t>{});
    case GG_FLOAT32: return f(Tag<float>{});
    case GG_FLOAT64: return f(Tag<double>{});
  }
  return GG_ERR_UNSUPPORTED;
}

template <class F> int with_semiring(int add, int mul, F&& f) {
  if (add == GG_OP_ADD && mul == GG_OP_MUL) return f(SR<GG_OP_ADD, GG_OP_MUL>{});
  if (add == GG_OP_BOR && mul == GG_OP_BAND) return f(SR<GG_OP_BOR, GG_OP_BAND>{});
  if (add == GG_OP_LOR && mul == GG_OP_LAND) return f(SR<GG_OP_LOR, GG_OP_LAND>{});
  if (add == GG_OP_ANY && mul == GG_OP_SECONDI) return f(SR<GG_OP_ANY, GG_OP_SECONDI>{});
  if (add == GG_OP_MIN && mul == GG_OP_SAFE_ADD) return f(SR<GG_OP_MIN, GG_OP_SAFE_ADD>{});
  if (add == GG_OP_MIN && mul == GG_OP_ADD) return f(SR<GG_OP_MIN, GG_OP_ADD>{});
  return GG_ERR_UNSUPPORTED;
}

// f(S, TA, TX) for a supported (semiring, TA, TX); TC follows.
template <class F> int dispatch(int add, int mul, int ta, int tx, F&& f) {
  return with_semiring(add, mul, [&](auto s) {
    return with_type(ta, [&](auto a) {
      return with_type(tx, [&](auto x) {
        typedef decltype(s) S;
        typedef typename decltype(a)::type TA;
        typedef typename decltype(x)::type TX;
        if constexpr (semiring_ok<S, TA, TX>()) {
          return f(s, a, x);
        } else {
          return (int)GG_ERR_UNSUPPORTED;
        }
      });
    });
  });
}

template <class T> constexpr int code_of();
template <> constexpr int code_of<jbool>() { return GG_BOOL; }
template <> constexpr int code_of<int32_t>() { return GG_INT32; }
template <> constexpr int code_of<int64_t>() { return GG_INT64; }
template <> constexpr int code_of<float>() { return GG_FLOAT32; }
template <> constexpr int code_of<double>() { return GG_FLOAT64; }

}  // namespace gg
