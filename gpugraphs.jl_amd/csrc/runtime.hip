// runtime.hip -- contexts, device memory, workspace, errors, timing.
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "internal.hpp"

namespace gg {

static thread_local char g_err[512];

int set_error(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return status;
}

void clear_error() { g_err[0] = 0; }

int ws_get(gg_ctx* ctx, int slot, size_t bytes, void** out, bool zero_new) {
  if (bytes < 256) bytes = 256;  // whole words for byte-wide atomics, never zero
  if (ctx->ws_bytes[slot] < bytes) {
    if (ctx->ws[slot]) {
      GG_HIP(hipStreamSynchronize(ctx->stream));
      GG_HIP(hipFree(ctx->ws[slot]));
      ctx->ws[slot] = nullptr;
      ctx->ws_bytes[slot] = 0;
    }
    size_t want = bytes + bytes / 4;  // headroom so small growth does not reallocate
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess)
      return set_error(GG_ERR_NOMEM, "workspace slot %d: hipMalloc(%zu) failed", slot, want);
    if (zero_new) GG_HIP(hipMemsetAsync(p, 0, want, ctx->stream));
    ctx->ws[slot] = p;
    ctx->ws_bytes[slot] = want;
  }
  *out = ctx->ws[slot];
  return GG_OK;
}

void timer_begin(gg_ctx* ctx) {
  if (ctx->timing) (void)hipEventRecord(ctx->ev0, ctx->stream);
}
void timer_end(gg_ctx* ctx) {
  if (!ctx->timing) return;
  (void)hipEventRecord(ctx->ev1, ctx->stream);
  if (hipEventSynchronize(ctx->ev1) == hipSuccess) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_ms = ms;
  }
}

int graph_loop(gg_ctx* ctx, const std::vector<uint64_t>& key, const int32_t* done,
               int64_t max_launches, const std::function<int(hipStream_t)>& enqueue,
               int64_t* launches, const int32_t* readback, int readback_words) {
  int64_t di = 0;  // done's index in the readback
  if (readback) {
    di = done - readback;
    if (di < 0 || di >= readback_words || readback_words * 4 > 1024)
      return set_error(GG_ERR_ARGUMENT, "readback range must hold the done flag");
  } else {
    readback = done;
    readback_words = 1;
  }
  hipGraphExec_t exec = nullptr;
  auto it = ctx->graphs.find(key);
  if (it != ctx->graphs.end()) {
    exec = it->second;
  } else {
    if (!ctx->cap_stream) GG_HIP(hipStreamCreateWithFlags(&ctx->cap_stream, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    GG_HIP(hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeRelaxed));
    const int st = enqueue(ctx->cap_stream);
    const hipError_t e = hipStreamEndCapture(ctx->cap_stream, &graph);
    if (st != GG_OK) {
      if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
      return st;
    }
    GG_HIP(e);
    const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    GG_HIP(e2);
    if (ctx->graphs.size() >= 64) {  // bound the cache
      for (auto& kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
      ctx->graphs.clear();
    }
    ctx->graphs[key] = exec;
  }
  for (int64_t l = 1;; ++l) {
    GG_HIP(hipGraphLaunch(exec, ctx->stream));
    GG_HIP(hipMemcpyAsync(ctx->host_pinned, readback, sizeof(int32_t) * readback_words,
                          hipMemcpyDeviceToHost, ctx->stream));
    GG_HIP(hipStreamSynchronize(ctx->stream));
    if (((const int32_t*)ctx->host_pinned)[di]) {
      if (launches) *launches = l;
      return GG_OK;
    }
    if (l >= max_launches)
      return set_error(GG_ERR_NOT_CONVERGED, "captured loop did not finish in %lld launches",
                       (long long)l);
  }
}

int check_matrix(const gg_matrix* A) {
  if (!A) return set_error(GG_ERR_ARGUMENT, "matrix descriptor is NULL");
  if (A->m < 0 || A->n < 0 || A->nnz < 0)
    return set_error(GG_ERR_ARGUMENT, "negative matrix dimension");
  if (A->m > 2147483647LL || A->n > 2147483647LL)
    return set_error(GG_ERR_ARGUMENT, "dimensions beyond Int32 indices");
  if (dtype_size(A->val_dtype) == 0 || A->val_dtype == GG_UINT64)
    return set_error(GG_ERR_ARGUMENT, "unknown matrix value type %d", A->val_dtype);
  switch (A->format) {
    case GG_CSR: case GG_CSC:
      if (A->flags & GG_FLAG_INTERNAL_IDS)
        return set_error(GG_ERR_ARGUMENT, "internal ids exist only for a relabelled SELL");
      return GG_OK;
    case GG_SELL:
      if (A->slice_size < 1 && A->m > 0)
        return set_error(GG_ERR_ARGUMENT, "SELL slice_size must be >= 1");
      if (A->nslices < 0 || A->n_heavy < 0 || A->n_heavy > A->m)
        return set_error(GG_ERR_ARGUMENT, "SELL slice / heavy counts out of range");
      if (A->nslices > 0 && (!A->ptr || (A->n_stored > 0 && (!A->idx || !A->val))))
        return set_error(GG_ERR_ARGUMENT, "SELL slices without arrays");
      if (!A->perm) {
        // identity layout: the slices hold every row and nothing is heavy
        if (A->n_heavy > 0 && A->nslices > 0)
          return set_error(GG_ERR_ARGUMENT, "SELL heavy rows need a slot permutation");
        if (A->n_heavy == 0 && A->nslices * A->slice_size < A->m)
          return set_error(GG_ERR_ARGUMENT, "SELL slices do not cover the %lld rows",
                           (long long)A->m);
        if (A->slot_len)
          return set_error(GG_ERR_ARGUMENT, "SELL slot lengths need a slot permutation");
      }
      if (A->n_heavy > 0 && (!A->heavy_row || !A->heavy_ptr || !A->heavy_seg ||
                             !A->heavy_col || A->heavy_seg_len < 1 || A->heavy_nseg < A->n_heavy))
        return set_error(GG_ERR_ARGUMENT, "SELL heavy partition incomplete");
      if ((A->flags & GG_FLAG_INTERNAL_IDS) &&
          (!A->vinv || A->m != A->n || (A->nslices > 0 && !A->perm)))
        return set_error(GG_ERR_ARGUMENT, "relabelled SELL needs vinv, a permutation and m == n");
      return GG_OK;
  }
  return set_error(GG_ERR_ARGUMENT, "unknown matrix format %d", A->format);
}

}  // namespace gg

using namespace gg;

extern "C" {

int gg_abi_version(void) { return GG_ABI_VERSION; }

int gg_matrix_check(const gg_matrix* A) {
  clear_error();
  return check_matrix(A);
}

int gg_matrix_slots(const gg_matrix* A, int64_t* n_slots) {
  if (!A || !n_slots) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  GG_TRY(check_matrix(A));
  *n_slots = A->format == GG_SELL ? sell_slots(A) : 0;
  return GG_OK;
}

const char* gg_status_string(int status) {
  switch (status) {
    case GG_OK: return "OK";
    case GG_ERR_DIMENSION: return "DimensionMismatch";
    case GG_ERR_ARGUMENT: return "ArgumentError";
    case GG_ERR_ASSERTION: return "AssertionError";
    case GG_ERR_BOUNDS: return "BoundsError";
    case GG_ERR_UNSUPPORTED: return "Unsupported";
    case GG_ERR_HIP: return "HIPError";
    case GG_ERR_NOMEM: return "OutOfMemory";
    case GG_ERR_NOT_CONVERGED: return "NotConverged";
    case GG_ERR_BUDGET: return "BudgetExceeded";
  }
  return "Unknown";
}

const char* gg_last_error(void) { return g_err; }

int gg_ctx_create(int device, void* stream, gg_ctx** out) {
  if (!out) return set_error(GG_ERR_ARGUMENT, "out is NULL");
  int n = 0;
  GG_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return set_error(GG_ERR_ARGUMENT, "device %d not present (%d devices)", device, n);
  GG_HIP(hipSetDevice(device));
  gg_ctx* c = new gg_ctx();
  c->device = device;
  c->stream = (hipStream_t)stream;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
  if (hipHostMalloc(&c->host_pinned, 4096, hipHostMallocDefault) != hipSuccess) {
    delete c;
    return set_error(GG_ERR_HIP, "hipHostMalloc failed");
  }
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess) {
    (void)hipHostFree(c->host_pinned);
    delete c;
    return set_error(GG_ERR_HIP, "hipEventCreate failed");
  }
  *out = c;
  return GG_OK;
}

int gg_ctx_destroy(gg_ctx* ctx) {
  if (!ctx) return GG_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
  if (ctx->cap_stream) (void)hipStreamDestroy(ctx->cap_stream);
  for (int i = 0; i < WS_NSLOTS; ++i)
    if (ctx->ws[i]) (void)hipFree(ctx->ws[i]);
  if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->ev2) (void)hipEventDestroy(ctx->ev2);
  if (ctx->ev3) (void)hipEventDestroy(ctx->ev3);
  delete ctx;
  return GG_OK;
}

int gg_ctx_synchronize(gg_ctx* ctx) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  GG_HIP(hipStreamSynchronize(ctx->stream));
  return GG_OK;
}

int gg_ctx_set_stream(gg_ctx* ctx, void* stream) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  GG_HIP(hipStreamSynchronize(ctx->stream));
  ctx->stream = (hipStream_t)stream;
  return GG_OK;
}

int gg_malloc(gg_ctx* ctx, size_t bytes, void** out) {
  if (!ctx || !out) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  GG_HIP(hipSetDevice(ctx->device));
  if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess)
    return set_error(GG_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
  return GG_OK;
}

int gg_free(gg_ctx* ctx, void* ptr) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  if (ptr) GG_HIP(hipFree(ptr));
  return GG_OK;
}

int gg_memcpy_htod(gg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  if (bytes) {
    GG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    GG_HIP(hipStreamSynchronize(ctx->stream));
  }
  return GG_OK;
}

int gg_memcpy_dtoh(gg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  if (bytes) {
    GG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    GG_HIP(hipStreamSynchronize(ctx->stream));
  }
  return GG_OK;
}

int gg_ctx_enable_timing(gg_ctx* ctx, int on) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  ctx->timing = on != 0;
  return GG_OK;
}

int gg_ctx_set_graph_steps(gg_ctx* ctx, int steps) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  if (steps < 0 || steps > 256) return set_error(GG_ERR_ARGUMENT, "graph steps must be in 0:256");
  ctx->graph_steps = steps;
  return GG_OK;
}

int gg_ctx_set_csr_lanes(gg_ctx* ctx, int lanes) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  if (lanes == 0) {
    ctx->csr_log_lanes = -1;
    return GG_OK;
  }
  if (lanes < 1 || lanes > 64 || (lanes & (lanes - 1)))
    return set_error(GG_ERR_ARGUMENT, "lanes per CSR row must be 0 or a power of two in 1:64");
  int l = 0;
  while ((1 << l) < lanes) ++l;
  ctx->csr_log_lanes = l;
  return GG_OK;
}

int gg_ctx_set_push_threshold(gg_ctx* ctx, double alpha) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  if (!(alpha >= 0.0)) return set_error(GG_ERR_ARGUMENT, "alpha must be >= 0");
  ctx->push_alpha = alpha;
  return GG_OK;
}

int gg_ctx_set_sssp_threshold(gg_ctx* ctx, double alpha) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  if (!(alpha >= 0.0)) return set_error(GG_ERR_ARGUMENT, "alpha must be >= 0");
  ctx->sssp_alpha = alpha;
  return GG_OK;
}

int gg_ctx_last_kernel_ms(gg_ctx* ctx, double* ms) {
  if (!ctx || !ms) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  *ms = ctx->last_ms;
  return GG_OK;
}

int gg_ctx_hot_kernel_stats(gg_ctx* ctx, double* total_ms, int64_t* launches, int reset) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "ctx is NULL");
  if (total_ms) *total_ms = ctx->hot_ms;
  if (launches) *launches = ctx->hot_launches;
  if (reset) {
    ctx->hot_ms = 0.0;
    ctx->hot_launches = 0;
  }
  return GG_OK;
}

}  // extern "C"
