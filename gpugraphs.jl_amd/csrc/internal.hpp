// internal.hpp -- context, workspace and error plumbing shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <functional>
#include <map>
#include <string>
#include <vector>

#include "../../include/gpugraphs.h"
#include "semiring.hpp"

namespace gg {

// Thread-local message of the last failure (gg_last_error).
int set_error(int status, const char* fmt, ...);
void clear_error();

#define GG_HIP(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return ::gg::set_error(GG_ERR_HIP, "%s failed: %s (%s:%d)", #call,           \
                             hipGetErrorString(e_), __FILE__, __LINE__);           \
  } while (0)

#define GG_TRY(expr)            \
  do {                          \
    int s_ = (expr);            \
    if (s_ != GG_OK) return s_; \
  } while (0)

// Named workspace slots owned by a context.  A slot grows on demand (the
// stream is synchronised before a reallocation) and is never shared by two
// live uses within one entry point.
enum WsSlot {
  WS_PARTIAL = 0,   // heavy-row segment partials
  WS_COUNTER,       // per-heavy-row arrival counters (kept zero between launches)
  WS_FLAG,          // small device flags / counters
  WS_BFS0, WS_BFS1, WS_BFS2, WS_BFS3, WS_BFS4, WS_BFS5,
  WS_SSSP0, WS_SSSP1, WS_SSSP2,
  WS_SPMSPV0, WS_SPMSPV1, WS_SPMSPV2,
  WS_MS0, WS_MS1, WS_MS2, WS_MS3, WS_MS4,
  WS_XPERM,         // input vector / matrix gathered into internal order
  WS_MSDIST,        // row-major distances of a multi-source traversal
  WS_MSHEAVY,       // per-heavy-row segment accumulators (kept zero between levels)
  WS_MSPUSH,        // top-down step buffers of multi-source BFS (accumulators kept zero)
  WS_BROWS,         // a column-major dense operand made row-major
  WS_NSLOTS
};

}  // namespace gg

struct gg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* ws[gg::WS_NSLOTS] = {};
  size_t ws_bytes[gg::WS_NSLOTS] = {};
  void* host_pinned = nullptr;  // 4 KiB pinned staging for flags / counts
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around a whole entry point
  hipEvent_t ev2 = nullptr, ev3 = nullptr;  // around one hot-kernel launch
  double last_ms = 0.0;
  double hot_ms = 0.0;                       // summed hot-kernel launch durations
  int64_t hot_launches = 0;
  int num_cus = 256;
  // captured loop steps (gg_ctx_set_graph_steps): steps per graph launch, the
  // stream graphs are captured on, and the instantiated graphs by signature
  int graph_steps = 0;
  int csr_log_lanes = -1;  // lanes per CSR row as log2 (gg_ctx_set_csr_lanes), -1 = by row length
  double push_alpha = 16.0;  // top-down when frontier out-edges * alpha < entries
  double sssp_alpha = 16.0;  // sparse k-source Bellman-Ford iteration: fallen rows'
                             // out-edges * alpha < entries
  hipStream_t cap_stream = nullptr;
  std::map<std::vector<uint64_t>, hipGraphExec_t> graphs;
};

namespace gg {

// Device pointer of at least `bytes` in `slot`; zero-filled when `zero`
// and freshly (re)allocated.
int ws_get(gg_ctx* ctx, int slot, size_t bytes, void** out, bool zero_new = false);

// Time the launches between timer_begin / timer_end when the context asks.
void timer_begin(gg_ctx* ctx);
void timer_end(gg_ctx* ctx);

// Times one launch of a hot kernel (the traversal / product kernel of an
// entry point) with events on the context stream when timing is enabled.
struct HotTimer {
  gg_ctx* c;
  bool on;
  bool count;  // false: the time joins the hot kernel's without counting a launch
  explicit HotTimer(gg_ctx* ctx, bool active = true, bool counted = true)
      : c(ctx), on(active && ctx->timing), count(counted) {
    if (on) (void)hipEventRecord(c->ev2, c->stream);
  }
  ~HotTimer() {
    if (!on) return;
    (void)hipEventRecord(c->ev3, c->stream);
    if (hipEventSynchronize(c->ev3) == hipSuccess) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, c->ev2, c->ev3) == hipSuccess) {
        c->hot_ms += ms;
        if (count) c->hot_launches += 1;
      }
    }
  }
};

inline size_t dtype_size(int code) {
  switch (code) {
    case GG_BOOL: return 1;
    case GG_INT32: case GG_FLOAT32: return 4;
    case GG_INT64: case GG_FLOAT64: case GG_UINT64: return 8;
  }
  return 0;
}

inline unsigned blocks_for(int64_t work, int per_block) {
  int64_t b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Validation shared by gpu_spmv!-style entry points (src/spmv/CSR.jl:358-408).
int check_matrix(const gg_matrix* A);

// Number of SELL slots a kernel walks: every slot of every slice when a
// permutation is given (empty slots hold perm 0), else the identity layout's
// rows, which the slices must cover.  Zero slices means zero slots whatever
// perm is (a hybrid whose rows are all in the heavy partition).
inline int64_t sell_slots(const gg_matrix* A) {
  const int64_t all = A->nslices * A->slice_size;
  return A->perm ? all : (A->m < all ? A->m : all);
}

// Lanes per row of the CSR kernels (log2): the context's setting, else the
// next power of two >= the average row length, at most a wavefront.
inline int csr_log_lanes(const gg_ctx* ctx, const gg_matrix* A) {
  if (ctx->csr_log_lanes >= 0) return ctx->csr_log_lanes;
  if (A->m == 0) return 0;
  const double avg = (double)A->nnz / (double)A->m;
  int lg = 0;
  while (lg < 6 && (double)(1 << lg) < avg) ++lg;
  return lg;
}

// Appends the bytes of a value to a graph signature.
template <class T> inline void sig_add(std::vector<uint64_t>& key, const T& v) {
  const size_t n = (sizeof(T) + 7) / 8;
  const size_t at = key.size();
  key.resize(at + n, 0);
  memcpy(key.data() + at, &v, sizeof(T));
}

// Runs a loop as a captured graph: `enqueue(stream)` enqueues U steps of the
// loop on `stream` (captured once per signature `key` and cached in the
// context); the graph is launched on the context stream until the device
// flag *done becomes nonzero, at most max_launches times.  The steps must be
// kernels only, and each must do nothing once *done is set.
// With readback, each check copies readback[0 .. readback_words) (which holds
// done) into ctx->host_pinned, so the caller finds those counters there after
// the loop without another round trip.
int graph_loop(gg_ctx* ctx, const std::vector<uint64_t>& key, const int32_t* done,
               int64_t max_launches, const std::function<int(hipStream_t)>& enqueue,
               int64_t* launches, const int32_t* readback = nullptr, int readback_words = 0);

// Internal launchers (used by the algorithms; no argument checks).
// rows (CSR only): a 1-based row list of nrows entries instead of r0 + q.
int launch_spmv(gg_ctx* ctx, void* C, int c_dtype, const gg_matrix* A, const void* B,
                int b_dtype, int add, int mul, int accum, const void* mask, int mask_dtype,
                int64_t r0, int64_t nrows, bool has_range, const int32_t* rows = nullptr);

}  // namespace gg
