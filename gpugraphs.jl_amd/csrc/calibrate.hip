// calibrate.hip -- streaming kernels that measure what this device's HBM
// delivers to a kernel of this library's kind (libgg_calibrate.so; used by
// bench.py, not by the product path).
//
//  - ggc_stream_read: every lane reads 16-byte words with non-temporal loads
//    (the index / value streams of the sparse kernels are read that way) in a
//    grid-stride loop, 4 words in flight per lane, and folds them into one
//    word per block so the loads are not dead;
//  - ggc_stream_write: 16-byte non-temporal stores of a constant;
//  - ggc_stream_copy: both;
//  - ggc_gather / ggc_gather8: the gathers of a sparse product alone (an
//    index stream and the 4- or 8-byte values it names), the part of an SpMV
//    or of a multi-source BFS level that does not stream.
// Each runs `reps` launches on the given stream between two events and
// returns the mean milliseconds per launch.  The grid is 8 blocks of 256
// threads per CU, enough waves to cover the memory latency of every CU.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

typedef uint32_t u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float fold(float a, float b) { return a + b; }
__device__ __forceinline__ uint64_t fold(uint64_t a, uint64_t b) { return a | b; }

__global__ __launch_bounds__(256) void read_kernel(const u4* __restrict__ src, int64_t n16,
                                                   u4* __restrict__ sink) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  u4 acc = {0u, 0u, 0u, 0u};
  for (; i + 3 * stride < n16; i += 4 * stride) {
    const u4 a = __builtin_nontemporal_load(&src[i]);
    const u4 b = __builtin_nontemporal_load(&src[i + stride]);
    const u4 c = __builtin_nontemporal_load(&src[i + 2 * stride]);
    const u4 d = __builtin_nontemporal_load(&src[i + 3 * stride]);
    acc ^= a ^ b ^ c ^ d;
  }
  for (; i < n16; i += stride) acc ^= __builtin_nontemporal_load(&src[i]);
  __shared__ u4 red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    u4 t = red[0];
    for (int j = 1; j < 256; ++j) t ^= red[j];
    sink[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(256) void write_kernel(u4* __restrict__ dst, int64_t n16, uint32_t v) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  const u4 w = {v, v, v, v};
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride)
    __builtin_nontemporal_store(w, &dst[i]);
}

__global__ __launch_bounds__(256) void copy_kernel(const u4* __restrict__ src, u4* __restrict__ dst,
                                                   int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (; i + stride < n16; i += 2 * stride) {
    const u4 a = __builtin_nontemporal_load(&src[i]);
    const u4 b = __builtin_nontemporal_load(&src[i + stride]);
    __builtin_nontemporal_store(a, &dst[i]);
    __builtin_nontemporal_store(b, &dst[i + stride]);
  }
  for (; i < n16; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}

// Gathers x[idx[i] - 1] for every positive 1-based index of idx, read in
// storage order (lane i of a wavefront takes entry i of its run of 64, as the
// SELL SpMV's lanes take entry k of 64 rows stored contiguously), 4 entries
// in flight per lane; the sum (or OR, for 8-byte masks) of each block goes to
// sink so nothing is dead.
template <class T>
__global__ __launch_bounds__(256) void gather_kernel(const int32_t* __restrict__ idx, int64_t n,
                                                     const T* __restrict__ x,
                                                     T* __restrict__ sink) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  T acc = T(0);
  for (; i + 3 * stride < n; i += 4 * stride) {
    int32_t c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = __builtin_nontemporal_load(&idx[i + u * stride]);
    T v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = c[u] > 0 ? x[c[u] - 1] : T(0);
#pragma unroll
    for (int u = 0; u < 4; ++u) acc = fold(acc, v[u]);
  }
  for (; i < n; i += stride) {
    const int32_t c = idx[i];
    if (c > 0) acc = fold(acc, x[c - 1]);
  }
  __shared__ T red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = T(0);
    for (int j = 0; j < 256; ++j) t = fold(t, red[j]);
    sink[blockIdx.x] = t;
  }
}

unsigned grid_blocks() {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) cus = p.multiProcessorCount;
  }
  return (unsigned)cus * 8u;
}

template <class F> int timed(hipStream_t st, int reps, F&& launch, double* ms) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return 1;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return 1;
  }
  launch();  // warm-up
  int rc = hipEventRecord(e0, st) != hipSuccess;
  for (int r = 0; r < reps && !rc; ++r) launch();
  rc |= hipGetLastError() != hipSuccess;
  rc |= hipEventRecord(e1, st) != hipSuccess;
  rc |= hipEventSynchronize(e1) != hipSuccess;
  float t = 0.f;
  rc |= hipEventElapsedTime(&t, e0, e1) != hipSuccess;
  *ms = reps > 0 ? (double)t / reps : 0.0;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

}  // namespace

extern "C" {

// bytes must be a multiple of 16 and sink must hold 16 bytes per block
// (ggc_sink_bytes); returns 0 on success.
size_t ggc_sink_bytes(void) { return (size_t)grid_blocks() * 16; }

int ggc_stream_read(void* stream, const void* src, size_t bytes, void* sink, int reps, double* ms) {
  if (!src || !sink || !ms || (bytes & 15)) return 2;
  const unsigned nb = grid_blocks();
  hipStream_t st = (hipStream_t)stream;
  return timed(st, reps, [&] {
    hipLaunchKernelGGL(read_kernel, dim3(nb), dim3(256), 0, st, (const u4*)src,
                       (int64_t)(bytes / 16), (u4*)sink);
  }, ms);
}

int ggc_stream_write(void* stream, void* dst, size_t bytes, int reps, double* ms) {
  if (!dst || !ms || (bytes & 15)) return 2;
  const unsigned nb = grid_blocks();
  hipStream_t st = (hipStream_t)stream;
  return timed(st, reps, [&] {
    hipLaunchKernelGGL(write_kernel, dim3(nb), dim3(256), 0, st, (u4*)dst, (int64_t)(bytes / 16),
                       0x3f800000u);
  }, ms);
}

// idx: n 1-based column indices (<= 0 skipped), each at most nx; sink holds
// ggc_sink_bytes() bytes.
int ggc_gather(void* stream, const int32_t* idx, int64_t n, const float* x, void* sink, int reps,
               double* ms) {
  if (!idx || !x || !sink || !ms || n < 0) return 2;
  const unsigned nb = grid_blocks();
  hipStream_t st = (hipStream_t)stream;
  return timed(st, reps, [&] {
    hipLaunchKernelGGL(gather_kernel<float>, dim3(nb), dim3(256), 0, st, idx, n, x, (float*)sink);
  }, ms);
}

// The same with 8-byte words (the source masks of a multi-source BFS level).
int ggc_gather8(void* stream, const int32_t* idx, int64_t n, const uint64_t* x, void* sink,
                int reps, double* ms) {
  if (!idx || !x || !sink || !ms || n < 0) return 2;
  const unsigned nb = grid_blocks();
  hipStream_t st = (hipStream_t)stream;
  return timed(st, reps, [&] {
    hipLaunchKernelGGL(gather_kernel<uint64_t>, dim3(nb), dim3(256), 0, st, idx, n, x,
                       (uint64_t*)sink);
  }, ms);
}

int ggc_stream_copy(void* stream, const void* src, void* dst, size_t bytes, int reps, double* ms) {
  if (!src || !dst || !ms || (bytes & 15)) return 2;
  const unsigned nb = grid_blocks();
  hipStream_t st = (hipStream_t)stream;
  return timed(st, reps, [&] {
    hipLaunchKernelGGL(copy_kernel, dim3(nb), dim3(256), 0, st, (const u4*)src, (u4*)dst,
                       (int64_t)(bytes / 16));
  }, ms);
}

}  // extern "C"
