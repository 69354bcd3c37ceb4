// tune.hip -- the autotuner behind the C ABI (gg_tune_*): it builds every
// candidate layout of one graph on the device from its CSR (gg_sell_plan_device
// / gg_sell_fill_device), times the workload the caller names on each with HIP
// events on the context stream, and returns the fastest; a second entry point
// times the multi-source BFS direction switch.  Nothing here is a kernel of its
// own: the timed calls are the library's entry points, so a tuned choice
// computes exactly what any other layout computes.
#include <stdlib.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "internal.hpp"

using namespace gg;

namespace {

// Device arrays of one candidate SELL layout; freed by the destructor.
struct DevSell {
  gg_matrix d{};
  std::vector<void*> bufs;
  ~DevSell() {
    for (void* p : bufs) (void)hipFree(p);
  }
  int alloc(size_t bytes, void** out) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess)
      return set_error(GG_ERR_NOMEM, "tuner: hipMalloc(%zu) failed", bytes);
    bufs.push_back(p);
    *out = p;
    return GG_OK;
  }
};

int build_sell(gg_ctx* ctx, const gg_matrix* A, const void* vals, int vdt, const gg_tune_candidate& c,
               DevSell& out) {
  gg_sell_plan* plan = nullptr;
  gg_sell_sizes z{};
  GG_TRY(gg_sell_plan_device(ctx, A->m, A->n, A->ptr, A->idx, vals, vdt, c.slice_size, c.sigma,
                             c.heavy_threshold, c.heavy_seg_len, c.sell_flags, &plan, &z));
  gg_matrix& d = out.d;
  const size_t es = dtype_size(vdt);
  int st = GG_OK;
  auto get = [&](size_t bytes, auto** p) {
    void* q = nullptr;
    if (st == GG_OK) st = out.alloc(bytes, &q);
    *p = (std::remove_reference_t<decltype(*p)>)q;
  };
  int32_t *sp, *cv, *perm = nullptr, *sl = nullptr, *hr = nullptr, *hs = nullptr, *hc = nullptr,
          *vinv = nullptr;
  int64_t* hp = nullptr;
  void *nz, *hv = nullptr;
  get(sizeof(int32_t) * (z.nslices + 1), &sp);
  get(sizeof(int32_t) * z.n_stored, &cv);
  get(es * z.n_stored, &nz);
  if (z.has_perm) {
    get(sizeof(int32_t) * z.n_slots, &perm);
    get(sizeof(int32_t) * z.n_slots, &sl);
  }
  if (z.n_heavy) {
    get(sizeof(int32_t) * z.n_heavy, &hr);
    get(sizeof(int64_t) * (z.n_heavy + 1), &hp);
    get(sizeof(int32_t) * (z.n_heavy + 1), &hs);
    get(sizeof(int32_t) * z.heavy_nnz, &hc);
    get(es * z.heavy_nnz, &hv);
  }
  if (z.has_vinv) get(sizeof(int32_t) * A->m, &vinv);
  if (st == GG_OK) {
    d.ptr = sp;
    d.idx = cv;
    d.val = nz;
    d.perm = perm;
    d.slot_len = sl;
    d.heavy_row = hr;
    d.heavy_ptr = hp;
    d.heavy_seg = hs;
    d.heavy_col = hc;
    d.heavy_val = hv;
    d.vinv = vinv;
    st = gg_sell_fill_device(plan, &d);
  }
  gg_sell_plan_free(plan);
  GG_TRY(st);
  // the pattern promises of the source matrix carry over
  d.flags |= A->flags & (GG_FLAG_UNIT_PATTERN | GG_FLAG_SYMMETRIC);
  if (vals != A->val) d.flags &= ~(int64_t)GG_FLAG_UNIT_PATTERN;
  return GG_OK;
}

__global__ void ones_f32_kernel(float* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 1.0f;
}

// Buffers the workloads write, sized once for the largest of them.
struct Outputs {
  void* x = nullptr;   // SpMV input (n fp32)
  void* y = nullptr;   // SpMV result / one-source distances / BFS levels
  void* D = nullptr;   // k-source distances (m x k)
  ~Outputs() {
    if (x) (void)hipFree(x);
    if (y) (void)hipFree(y);
    if (D) (void)hipFree(D);
  }
};

int dist_code(int vdt) { return vdt == GG_BOOL ? GG_INT32 : vdt; }

int run_workload(gg_ctx* ctx, const gg_matrix* M, int workload, const int32_t* roots,
                 int64_t n_roots, const Outputs& o) {
  switch (workload) {
    case GG_TUNE_SPMV:
      return gg_spmv(ctx, o.y, GG_FLOAT32, M->m, M, o.x, GG_FLOAT32, M->n, GG_OP_ADD, GG_OP_MUL,
                     GG_OP_SECOND, nullptr, 0, 0, 0, -1);
    case GG_TUNE_BFS:
      for (int64_t i = 0; i < n_roots; ++i)
        GG_TRY(gg_bfs_distances(ctx, M, roots[i], (int32_t*)o.y, 1));
      return GG_OK;
    case GG_TUNE_MSBFS:
      return gg_msbfs_distances(ctx, M, roots, std::min<int64_t>(n_roots, 64), (int32_t*)o.D,
                                nullptr);
    case GG_TUNE_SSSP:
      return gg_shortest_path(ctx, M, roots[0], o.y, dist_code(M->val_dtype), nullptr);
    case GG_TUNE_MSSSP:
      return gg_shortest_path_multi(ctx, M, roots, std::min<int64_t>(n_roots, 32), o.D,
                                    dist_code(M->val_dtype), nullptr);
  }
  return set_error(GG_ERR_ARGUMENT, "unknown tuning workload %d", workload);
}

// Best of `reps` timed runs after one untimed run, in ms (events on the
// context stream around each run; the entry points return synchronised or
// are followed by the event wait).
struct Events {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~Events() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  int init() {
    GG_HIP(hipEventCreate(&e0));
    GG_HIP(hipEventCreate(&e1));
    return GG_OK;
  }
};

int time_workload(gg_ctx* ctx, const gg_matrix* M, int workload, const int32_t* roots,
                  int64_t n_roots, const Outputs& o, int reps, const Events& ev, double* ms) {
  GG_TRY(run_workload(ctx, M, workload, roots, n_roots, o));
  GG_HIP(hipStreamSynchronize(ctx->stream));
  double best = 1e300;
  for (int r = 0; r < std::max(reps, 1); ++r) {
    GG_HIP(hipEventRecord(ev.e0, ctx->stream));
    GG_TRY(run_workload(ctx, M, workload, roots, n_roots, o));
    GG_HIP(hipEventRecord(ev.e1, ctx->stream));
    GG_HIP(hipEventSynchronize(ev.e1));
    float t = 0.f;
    GG_HIP(hipEventElapsedTime(&t, ev.e0, ev.e1));
    best = std::min(best, (double)t);
  }
  *ms = best;
  return GG_OK;
}

gg_tune_candidate csr_layout(int lanes) {
  gg_tune_candidate c{};
  c.format = GG_CSR;
  c.csr_lanes = lanes;
  return c;
}
gg_tune_candidate sell_layout(int64_t C, int64_t sigma, int64_t heavy, int64_t seg, int flags) {
  gg_tune_candidate c{};
  c.format = GG_SELL;
  c.slice_size = C;
  c.sigma = sigma;
  c.heavy_threshold = heavy;
  c.heavy_seg_len = seg;
  c.sell_flags = flags;
  return c;
}

}  // namespace

// Times `workload` (its first `cols` roots at once) under each value of the
// context threshold *knob, leaves the fastest on the context and in *best.
static int tune_threshold(gg_ctx* ctx, double* knob, int workload, int64_t cols,
                          const gg_matrix* A_T, const int32_t* roots, int64_t k,
                          const double* alphas, double* ms, int64_t n_alphas, int reps,
                          double* best_alpha) {
  if (!A_T || !roots || !alphas || !best_alpha || n_alphas < 1 || k < 1)
    return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  for (int64_t i = 0; i < n_alphas; ++i)
    if (!(alphas[i] >= 0.0)) return set_error(GG_ERR_ARGUMENT, "alpha %g must be >= 0", alphas[i]);
  Outputs o;
  Events ev;
  GG_TRY(ev.init());
  const size_t esz = workload == GG_TUNE_MSSSP ? 8 : 4;  // room for 64-bit distances
  if (hipMalloc(&o.D, esz * std::max<int64_t>(A_T->m * cols, 1)) != hipSuccess)
    return set_error(GG_ERR_NOMEM, "tuner: output buffer");
  const double saved = *knob;
  int64_t bi = -1;
  double bms = 0.0;
  for (int64_t i = 0; i < n_alphas; ++i) {
    *knob = alphas[i];
    double t = 0.0;
    const int st = time_workload(ctx, A_T, workload, roots, k, o, reps, ev, &t);
    if (st != GG_OK) {
      *knob = saved;
      return st;
    }
    if (ms) ms[i] = t;
    if (bi < 0 || t < bms) {
      bi = i;
      bms = t;
    }
  }
  *knob = alphas[bi];  // the context keeps the best switch
  *best_alpha = alphas[bi];
  return GG_OK;
}

extern "C" {

int gg_tune_default_candidates(const gg_matrix* A, int64_t max_degree, gg_tune_candidate* out,
                               int64_t capacity, int64_t* n_out) {
  if (!A || !n_out || (capacity > 0 && !out)) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  std::vector<gg_tune_candidate> c;
  for (int L : {1, 4, 16, 64}) c.push_back(csr_layout(L));
  const bool square = A->m == A->n;
  if (square) {
    for (int64_t heavy : {256, 1024, 4096})
      if (heavy < max_degree)
        for (int64_t seg : {4096, 65536}) c.push_back(sell_layout(64, 1, heavy, seg, GG_SELL_RELABEL));
    c.push_back(sell_layout(64, 1, 0, 4096, GG_SELL_RELABEL));
    if (1024 < max_degree) {
      c.push_back(sell_layout(64, 1, 1024, 4096, GG_SELL_RELABEL | GG_SELL_SORT_ENTRIES));
      c.push_back(sell_layout(64, 1, 1024, 65536, GG_SELL_RELABEL | GG_SELL_SORT_ENTRIES));
    }
    c.push_back(sell_layout(32, 1, 1024 < max_degree ? 1024 : 0, 4096, GG_SELL_RELABEL));
  }
  c.push_back(sell_layout(64, 4096, 256 < max_degree ? 256 : 0, 4096, 0));
  *n_out = (int64_t)c.size();
  if ((int64_t)c.size() > capacity)
    return set_error(GG_ERR_ARGUMENT, "%lld candidates need room for %lld", (long long)c.size(),
                     (long long)c.size());
  std::copy(c.begin(), c.end(), out);
  return GG_OK;
}

int gg_tune_layout(gg_ctx* ctx, const gg_matrix* A, int workload, const int32_t* roots,
                   int64_t n_roots, gg_tune_candidate* candidates, int64_t n_candidates, int reps,
                   gg_tune_candidate* best) {
  if (!ctx || !A || !best || !candidates || n_candidates < 1)
    return set_error(GG_ERR_ARGUMENT, "NULL argument or no candidates");
  clear_error();
  GG_TRY(check_matrix(A));
  if (A->format != GG_CSR) return set_error(GG_ERR_ARGUMENT, "the tuner builds from a CSR matrix");
  if (workload < GG_TUNE_SPMV || workload > GG_TUNE_MSSSP)
    return set_error(GG_ERR_ARGUMENT, "unknown tuning workload %d", workload);
  if (workload != GG_TUNE_SPMV) {
    if (n_roots < 1 || !roots) return set_error(GG_ERR_ARGUMENT, "graph workloads need roots");
    if (A->m != A->n) return set_error(GG_ERR_DIMENSION, "A_T must be square");
    for (int64_t i = 0; i < n_roots; ++i)
      if (roots[i] < 1 || roots[i] > A->m)
        return set_error(GG_ERR_BOUNDS, "root %d outside 1:%lld", roots[i], (long long)A->m);
  }
  GG_HIP(hipSetDevice(ctx->device));
  // SpMV times plus-times fp32: other value types are timed on fp32 ones
  const void* vals = A->val;
  int vdt = A->val_dtype;
  void* ones = nullptr;
  if (workload == GG_TUNE_SPMV && vdt != GG_FLOAT32) {
    if (hipMalloc(&ones, sizeof(float) * std::max<int64_t>(A->nnz, 1)) != hipSuccess)
      return set_error(GG_ERR_NOMEM, "tuner: fp32 values");
    hipLaunchKernelGGL(ones_f32_kernel, dim3(blocks_for(A->nnz, 256)), dim3(256), 0, ctx->stream,
                       (float*)ones, A->nnz);
    vals = ones;
    vdt = GG_FLOAT32;
  }
  gg_matrix Acsr = *A;
  Acsr.val = const_cast<void*>(vals);
  Acsr.val_dtype = vdt;
  if (vals != A->val) Acsr.flags &= ~(int64_t)GG_FLAG_UNIT_PATTERN;
  Outputs o;
  Events ev;
  GG_TRY(ev.init());
  const int64_t kk = workload == GG_TUNE_MSBFS ? std::min<int64_t>(n_roots, 64)
                     : workload == GG_TUNE_MSSSP ? std::min<int64_t>(n_roots, 32) : 0;
  int st = GG_OK;
  if (hipMalloc(&o.x, sizeof(float) * std::max<int64_t>(A->n, 1)) != hipSuccess ||
      hipMalloc(&o.y, 8 * std::max<int64_t>(A->m, 1)) != hipSuccess ||
      (kk > 0 && hipMalloc(&o.D, 8 * std::max<int64_t>(A->m * kk, 1)) != hipSuccess))
    st = set_error(GG_ERR_NOMEM, "tuner: output buffers");
  if (st == GG_OK && workload == GG_TUNE_SPMV)
    hipLaunchKernelGGL(ones_f32_kernel, dim3(blocks_for(A->n, 256)), dim3(256), 0, ctx->stream,
                       (float*)o.x, A->n);
  const int saved_lanes = ctx->csr_log_lanes;
  int64_t best_i = -1;
  for (int64_t i = 0; i < n_candidates && st == GG_OK; ++i) {
    gg_tune_candidate& c = candidates[i];
    c.ms = -1.0;
    if (c.format == GG_CSR) {
      st = gg_ctx_set_csr_lanes(ctx, c.csr_lanes);
      if (st == GG_OK) st = time_workload(ctx, &Acsr, workload, roots, n_roots, o, reps, ev, &c.ms);
      ctx->csr_log_lanes = saved_lanes;
    } else if (c.format == GG_SELL) {
      if ((c.sell_flags & GG_SELL_RELABEL) && A->m != A->n) continue;  // needs a square matrix
      DevSell M;
      st = build_sell(ctx, &Acsr, vals, vdt, c, M);
      if (st == GG_OK) st = time_workload(ctx, &M.d, workload, roots, n_roots, o, reps, ev, &c.ms);
      (void)hipStreamSynchronize(ctx->stream);  // the candidate's arrays are freed next
    } else {
      st = set_error(GG_ERR_ARGUMENT, "candidate %lld: format must be GG_CSR or GG_SELL",
                     (long long)i);
    }
    if (st == GG_OK && c.ms >= 0 && (best_i < 0 || c.ms < candidates[best_i].ms)) best_i = i;
  }
  if (ones) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ones);
  }
  GG_TRY(st);
  if (best_i < 0) return set_error(GG_ERR_ARGUMENT, "no candidate applies to this matrix");
  *best = candidates[best_i];
  return GG_OK;
}

int gg_tune_push_threshold(gg_ctx* ctx, const gg_matrix* A_T, const int32_t* roots, int64_t k,
                           const double* alphas, double* ms, int64_t n_alphas, int reps,
                           double* best_alpha) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  return tune_threshold(ctx, &ctx->push_alpha, GG_TUNE_MSBFS, std::min<int64_t>(k, 64), A_T, roots,
                        k, alphas, ms, n_alphas, reps, best_alpha);
}

int gg_tune_sssp_threshold(gg_ctx* ctx, const gg_matrix* A_T, const int32_t* roots, int64_t k,
                           const double* alphas, double* ms, int64_t n_alphas, int reps,
                           double* best_alpha) {
  if (!ctx) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  return tune_threshold(ctx, &ctx->sssp_alpha, GG_TUNE_MSSSP, std::min<int64_t>(k, 32), A_T, roots,
                        k, alphas, ms, n_alphas, reps, best_alpha);
}

}  // extern "C"
