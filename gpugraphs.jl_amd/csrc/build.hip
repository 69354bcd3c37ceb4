// build.hip -- the SELL-C-sigma builder on the device (gg_sell_plan_device /
// gg_sell_fill_device): the layouts of gg_sell_build, array for array, from
// a CSR already in HBM.
//
// The reference builds SELL on the host with a serial loop and a dense
// temporary per slice (src/storage/SELL.jl:99-181), then uploads.  Here the
// row order comes from stable radix sorts (decreasing length, ties in row
// order; per sigma window with the window index in the high key bits), the
// heavy / light split from two stable selections, slice widths and pointers
// from a max per slice and a scan, and every slot writes its own column of
// the slice, so a graph never leaves the device.  Two calls so that the
// caller allocates the output arrays (Julia's GC, torch): the plan computes
// the order and every size, the fill writes the caller's arrays.
#include <hipcub/hipcub.hpp>

#include "internal.hpp"

namespace gg {

constexpr int kB = 256;

__global__ void lens_kernel(const int32_t* __restrict__ rowptr, int64_t m, int32_t* len) {
  const int64_t r = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (r < m) len[r] = rowptr[r + 1] - rowptr[r];
}

__global__ void iota_kernel(int32_t* p, int64_t n) {  // 1-based rows
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) p[i] = (int32_t)(i + 1);
}

// Decreasing length as an increasing key.
__global__ void desc_key_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ len,
                                int64_t n, uint32_t* key) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) key[i] = 0xffffffffu - (uint32_t)len[rows[i] - 1];
}

// (sigma window, decreasing length) of the light rows in their current order.
__global__ void window_key_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ len,
                                  int64_t n, int64_t sigma, uint64_t* key) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n)
    key[i] = ((uint64_t)(i / sigma) << 32) | (uint64_t)(0xffffffffu - (uint32_t)len[rows[i] - 1]);
}

__global__ void heavy_flag_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ len,
                                  int64_t n, int64_t thr, uint8_t* heavy, uint8_t* light) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const bool h = len[rows[i] - 1] > thr;
  heavy[i] = h;
  light[i] = !h;
}

// Width of slice s times C, at w[s + 1] (w[0] = 0 for the scan).
__global__ void width_kernel(const int32_t* __restrict__ slot_rows, const int32_t* __restrict__ len,
                             int64_t nl, int64_t ns, int64_t C, int64_t* w) {
  const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (s == 0) w[0] = 0;
  if (s >= ns) return;
  int32_t mx = 0;
  for (int64_t q = s * C; q < min(nl, (s + 1) * C); ++q) mx = max(mx, len[slot_rows[q] - 1]);
  w[s + 1] = (int64_t)mx * C;
}

// Entries and segments of the heavy rows, at e[i + 1] and g[i + 1].
__global__ void heavy_count_kernel(const int32_t* __restrict__ heavy, const int32_t* __restrict__ len,
                                   int64_t nh, int64_t seg_len, int64_t* e, int64_t* g) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i == 0) {
    e[0] = 0;
    g[0] = 0;
  }
  if (i >= nh) return;
  const int64_t l = len[heavy[i] - 1];
  e[i + 1] = l;
  g[i + 1] = (l + seg_len - 1) / seg_len;
}

__global__ void vinv_kernel(const int32_t* __restrict__ slot_rows, int64_t nl,
                            const int32_t* __restrict__ heavy, int64_t nh, int64_t n_slots,
                            int32_t* vinv) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < nl) vinv[slot_rows[i] - 1] = (int32_t)(i + 1);
  else if (i < nl + nh) vinv[heavy[i - nl] - 1] = (int32_t)(n_slots + (i - nl) + 1);
}

// Sort keys of the entries (the stored column) and their positions, for
// GG_SELL_SORT_ENTRIES.
__global__ void entry_keys_kernel(const int32_t* __restrict__ colval, int64_t nnz,
                                  const int32_t* __restrict__ vinv, uint32_t* key, int32_t* pos) {
  const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (e >= nnz) return;
  const int32_t c = colval[e];
  key[e] = (uint32_t)(vinv ? vinv[c - 1] : c);
  pos[e] = (int32_t)e;
}
__global__ void zero_based_kernel(const int32_t* __restrict__ rowptr, int64_t m, int32_t* off) {
  const int64_t r = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (r <= m) off[r] = rowptr[r] - 1;
}

// Entry t of a row starting at `start`: its stored column and its position
// in the input (through the sorted order when sorted_col is set).
struct EntryMap {
  const int32_t* colval;
  const int32_t* vinv;
  const uint32_t* sorted_col;  // mapped columns, rows sorted; NULL: input order
  const int32_t* sorted_pos;
  __device__ __forceinline__ void get(int64_t e, int32_t& col, int64_t& src) const {
    if (sorted_col) {
      col = (int32_t)sorted_col[e];
      src = sorted_pos[e];
    } else {
      const int32_t c = colval[e];
      col = vinv ? vinv[c - 1] : c;
      src = e;
    }
  }
};

template <class V>
__global__ void slots_kernel(const int32_t* __restrict__ rowptr, EntryMap em,
                             const V* __restrict__ val, const int32_t* __restrict__ slot_rows,
                             int64_t nl, int64_t n_slots, int64_t C,
                             const int64_t* __restrict__ wscan,
                             int32_t* slice_ptr, int32_t* out_col, V* out_val, int32_t* perm,
                             int32_t* slot_len) {
  const int64_t q = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (q >= n_slots) return;
  const int64_t s = q / C, r = q - s * C;
  const int64_t base = wscan[s], w = (wscan[s + 1] - base) / C;
  if (r == 0) {
    slice_ptr[s] = (int32_t)(base + 1);
    if (s == n_slots / C - 1) slice_ptr[s + 1] = (int32_t)(wscan[s + 1] + 1);
  }
  const int32_t row = q < nl ? slot_rows[q] : 0;
  const int64_t start = row ? (int64_t)rowptr[row - 1] - 1 : 0;
  const int64_t len = row ? (int64_t)rowptr[row] - 1 - start : 0;
  if (perm) {
    perm[q] = row;
    slot_len[q] = (int32_t)len;
  }
  for (int64_t k = 0; k < w; ++k) {
    const int64_t pos = base + k * C + r;  // column-major inside the slice
    if (k < len) {
      int32_t c;
      int64_t src;
      em.get(start + k, c, src);
      out_col[pos] = c;
      out_val[pos] = val[src];
    } else {
      out_col[pos] = -1;
      out_val[pos] = V(0);
    }
  }
}

template <class V>
__global__ void heavy_kernel(const int32_t* __restrict__ rowptr, EntryMap em,
                             const V* __restrict__ val, const int32_t* __restrict__ heavy,
                             int64_t nh, const int64_t* __restrict__ escan,
                             const int64_t* __restrict__ gscan,
                             int32_t* h_row, int64_t* h_ptr, int32_t* h_seg, int32_t* h_col,
                             V* h_val) {
  const int64_t i = blockIdx.x;
  const int32_t row = heavy[i];
  const int64_t start = (int64_t)rowptr[row - 1] - 1, o = escan[i];
  const int64_t len = escan[i + 1] - o;
  if (threadIdx.x == 0) {
    h_row[i] = row;
    h_ptr[i] = o;
    h_seg[i] = (int32_t)gscan[i];
    if (i == nh - 1) {
      h_ptr[nh] = escan[nh];
      h_seg[nh] = (int32_t)gscan[nh];
    }
  }
  for (int64_t t = threadIdx.x; t < len; t += kB) {
    int32_t c;
    int64_t src;
    em.get(start + t, c, src);
    h_col[o + t] = c;
    h_val[o + t] = val[src];
  }
}

}  // namespace gg

using namespace gg;

struct gg_sell_plan {
  gg_ctx* ctx = nullptr;
  int64_t m = 0, n = 0, nnz = 0;
  int val_dtype = 0;
  const int32_t* rowptr = nullptr;
  const int32_t* colval = nullptr;
  const void* nzval = nullptr;
  int64_t C = 0, seg_len = 0, nl = 0, nh = 0, ns = 0;
  bool relabel = false, reference = false;
  int32_t* len = nullptr;        // [m]
  int32_t* slot_rows = nullptr;  // [m] light rows in slot order (first nl)
  int32_t* heavy = nullptr;      // [m] heavy rows (first nh)
  int64_t* wscan = nullptr;      // [ns+1] slice offsets (entries)
  int64_t* escan = nullptr;      // [nh+1] heavy entry offsets
  int64_t* gscan = nullptr;      // [nh+1] heavy segment offsets
  int32_t* vinv = nullptr;       // [m] (relabel)
  bool sort_entries = false;
  uint32_t* sorted_col = nullptr;  // [nnz] rows by increasing stored column
  int32_t* sorted_pos = nullptr;   // [nnz] their input positions
  gg_sell_sizes sizes{};
  std::vector<void*> allocs;
};

namespace {

int plan_alloc(gg_sell_plan* p, size_t bytes, void** out) {
  if (hipMalloc(out, bytes ? bytes : 16) != hipSuccess)
    return set_error(GG_ERR_NOMEM, "device allocation of %zu bytes failed", bytes);
  p->allocs.push_back(*out);
  return GG_OK;
}

template <class T> int plan_array(gg_sell_plan* p, int64_t n, T** out) {
  void* q;
  GG_TRY(plan_alloc(p, sizeof(T) * (size_t)std::max<int64_t>(n, 1), &q));
  *out = (T*)q;
  return GG_OK;
}

// hipCUB call with its temporary storage (allocated per call, freed by the plan).
template <class F> int cub_call(gg_sell_plan* p, F&& f) {
  size_t tb = 0;
  GG_HIP(f(nullptr, tb));
  void* tmp;
  GG_TRY(plan_alloc(p, tb, &tmp));
  GG_HIP(f(tmp, tb));
  return GG_OK;
}

int read_i64(gg_ctx* ctx, const int64_t* d, int64_t* h) {
  GG_HIP(hipMemcpyAsync(ctx->host_pinned, d, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  GG_HIP(hipStreamSynchronize(ctx->stream));
  *h = *(const int64_t*)ctx->host_pinned;
  return GG_OK;
}

int plan_build(gg_sell_plan* p, int64_t slice_size, int64_t sigma, int64_t thr) {
  gg_ctx* ctx = p->ctx;
  hipStream_t st = ctx->stream;
  const int64_t m = p->m;
  GG_TRY(plan_array(p, m, &p->len));
  GG_TRY(plan_array(p, m, &p->slot_rows));
  GG_TRY(plan_array(p, m, &p->heavy));
  if (m > 0) {
    hipLaunchKernelGGL(lens_kernel, dim3(blocks_for(m, kB)), dim3(kB), 0, st, p->rowptr, m, p->len);
    hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(m, kB)), dim3(kB), 0, st, p->slot_rows, m);
  }
  p->nl = m;
  p->C = p->reference ? (m > 0 ? std::min<int64_t>(slice_size, m) : slice_size) : slice_size;
  if (!p->reference && m > 0) {
    int32_t* order = p->slot_rows;
    if (p->relabel) {  // every row by decreasing length, ties in row order
      uint32_t *k0, *k1;
      int32_t* o1;
      GG_TRY(plan_array(p, m, &k0));
      GG_TRY(plan_array(p, m, &k1));
      GG_TRY(plan_array(p, m, &o1));
      hipLaunchKernelGGL(desc_key_kernel, dim3(blocks_for(m, kB)), dim3(kB), 0, st, p->slot_rows,
                         p->len, m, k0);
      GG_TRY(cub_call(p, [&](void* t, size_t& tb) {
        return hipcub::DeviceRadixSort::SortPairs(t, tb, k0, k1, p->slot_rows, o1, (int)m, 0, 32,
                                                  st);
      }));
      order = o1;
    }
    if (thr > 0) {  // stable split into light and heavy rows
      uint8_t *fh, *fl;
      int64_t* cnt;
      GG_TRY(plan_array(p, m, &fh));
      GG_TRY(plan_array(p, m, &fl));
      GG_TRY(plan_array(p, 2, &cnt));
      hipLaunchKernelGGL(heavy_flag_kernel, dim3(blocks_for(m, kB)), dim3(kB), 0, st, order, p->len,
                         m, thr, fh, fl);
      int32_t* light_out = order == p->slot_rows ? nullptr : p->slot_rows;
      int32_t* tmp_light = nullptr;
      if (!light_out) {
        GG_TRY(plan_array(p, m, &tmp_light));
        light_out = tmp_light;
      }
      GG_TRY(cub_call(p, [&](void* t, size_t& tb) {
        return hipcub::DeviceSelect::Flagged(t, tb, order, fl, light_out, cnt, (int)m, st);
      }));
      GG_TRY(cub_call(p, [&](void* t, size_t& tb) {
        return hipcub::DeviceSelect::Flagged(t, tb, order, fh, p->heavy, cnt + 1, (int)m, st);
      }));
      int64_t hc[2];
      GG_HIP(hipMemcpyAsync(ctx->host_pinned, cnt, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
      memcpy(hc, ctx->host_pinned, sizeof hc);
      p->nl = hc[0];
      p->nh = hc[1];
      p->slot_rows = light_out;
    } else if (order != p->slot_rows) {
      p->slot_rows = order;
    }
    if (sigma > 1 && !p->relabel && p->nl > 0) {  // windows of sigma rows by decreasing length
      uint64_t *k0, *k1;
      int32_t* o1;
      const int64_t nl = p->nl;
      GG_TRY(plan_array(p, nl, &k0));
      GG_TRY(plan_array(p, nl, &k1));
      GG_TRY(plan_array(p, nl, &o1));
      hipLaunchKernelGGL(window_key_kernel, dim3(blocks_for(nl, kB)), dim3(kB), 0, st,
                         p->slot_rows, p->len, nl, sigma, k0);
      GG_TRY(cub_call(p, [&](void* t, size_t& tb) {
        return hipcub::DeviceRadixSort::SortPairs(t, tb, k0, k1, p->slot_rows, o1, (int)nl, 0, 64,
                                                  st);
      }));
      p->slot_rows = o1;
    }
  }
  const int64_t C = p->C;
  p->ns = p->nl > 0 ? (p->nl + C - 1) / C : 0;
  GG_TRY(plan_array(p, p->ns + 1, &p->wscan));
  hipLaunchKernelGGL(width_kernel, dim3(blocks_for(std::max<int64_t>(p->ns, 1), kB)), dim3(kB), 0,
                     st, p->slot_rows, p->len, p->nl, p->ns, C, p->wscan);
  if (p->ns > 0)
    GG_TRY(cub_call(p, [&](void* t, size_t& tb) {
      return hipcub::DeviceScan::InclusiveSum(t, tb, p->wscan + 1, p->wscan + 1, (int)p->ns, st);
    }));
  int64_t n_stored = 0;
  GG_TRY(read_i64(ctx, p->wscan + p->ns, &n_stored));
  if (n_stored + 1 > 2147483647LL)
    return set_error(GG_ERR_ARGUMENT,
                     "SELL storage exceeds Int32 slice_ptr; raise sigma or use a heavy threshold");
  int64_t heavy_nnz = 0, n_seg = 0;
  GG_TRY(plan_array(p, p->nh + 1, &p->escan));
  GG_TRY(plan_array(p, p->nh + 1, &p->gscan));
  hipLaunchKernelGGL(heavy_count_kernel, dim3(blocks_for(std::max<int64_t>(p->nh, 1), kB)),
                     dim3(kB), 0, st, p->heavy, p->len, p->nh, std::max<int64_t>(p->seg_len, 1),
                     p->escan, p->gscan);
  if (p->nh > 0) {
    GG_TRY(cub_call(p, [&](void* t, size_t& tb) {
      return hipcub::DeviceScan::InclusiveSum(t, tb, p->escan + 1, p->escan + 1, (int)p->nh, st);
    }));
    GG_TRY(cub_call(p, [&](void* t, size_t& tb) {
      return hipcub::DeviceScan::InclusiveSum(t, tb, p->gscan + 1, p->gscan + 1, (int)p->nh, st);
    }));
    GG_TRY(read_i64(ctx, p->escan + p->nh, &heavy_nnz));
    GG_TRY(read_i64(ctx, p->gscan + p->nh, &n_seg));
    if (n_seg > 2147483647LL) return set_error(GG_ERR_ARGUMENT, "too many heavy segments");
  }
  if (p->relabel) {
    if (p->ns * C + p->nh > 2147483647LL)
      return set_error(GG_ERR_ARGUMENT, "internal ids exceed Int32");
    GG_TRY(plan_array(p, m, &p->vinv));
    if (p->nl + p->nh > 0)
      hipLaunchKernelGGL(vinv_kernel, dim3(blocks_for(p->nl + p->nh, kB)), dim3(kB), 0, st,
                         p->slot_rows, p->nl, p->heavy, p->nh, p->ns * C, p->vinv);
  }
  if (p->sort_entries && p->nnz > 0) {  // each row's entries by stored column, ties in input order
    if (p->nnz > 2147483647LL) return set_error(GG_ERR_ARGUMENT, "too many entries to sort");
    uint32_t* k0;
    int32_t *p0, *off;
    GG_TRY(plan_array(p, p->nnz, &k0));
    GG_TRY(plan_array(p, p->nnz, &p0));
    GG_TRY(plan_array(p, p->nnz, &p->sorted_col));
    GG_TRY(plan_array(p, p->nnz, &p->sorted_pos));
    GG_TRY(plan_array(p, m + 1, &off));
    hipLaunchKernelGGL(entry_keys_kernel, dim3(blocks_for(p->nnz, kB)), dim3(kB), 0, st, p->colval,
                       p->nnz, p->relabel ? p->vinv : nullptr, k0, p0);
    hipLaunchKernelGGL(zero_based_kernel, dim3(blocks_for(m + 1, kB)), dim3(kB), 0, st, p->rowptr,
                       m, off);
    GG_TRY(cub_call(p, [&](void* t, size_t& tb) {
      return hipcub::DeviceSegmentedRadixSort::SortPairs(t, tb, k0, p->sorted_col, p0,
                                                         p->sorted_pos, (int)p->nnz, (int)m, off,
                                                         off + 1, 0, 32, st);
    }));
  }
  GG_HIP(hipGetLastError());
  gg_sell_sizes& z = p->sizes;
  z.slice_size = C;
  z.nslices = p->ns;
  z.n_stored = n_stored;
  z.n_slots = p->ns * C;
  z.n_heavy = p->nh;
  z.heavy_nnz = heavy_nnz;
  z.n_seg = n_seg;
  z.n_internal = p->relabel ? p->ns * C + p->nh : m;
  z.has_perm = p->reference ? 0 : 1;
  z.has_vinv = p->relabel ? 1 : 0;
  return GG_OK;
}

}  // namespace

extern "C" {

int gg_sell_plan_device(gg_ctx* ctx, int64_t m, int64_t n, const int32_t* rowptr,
                        const int32_t* colval, const void* nzval, int val_dtype,
                        int64_t slice_size, int64_t sigma, int64_t heavy_threshold,
                        int64_t heavy_seg_len, int flags, gg_sell_plan** plan,
                        gg_sell_sizes* sizes) {
  if (!ctx || !plan || !sizes || (m > 0 && !rowptr) || m < 0 || n < 0)
    return set_error(GG_ERR_ARGUMENT, "NULL or negative argument");
  clear_error();
  const size_t esz = dtype_size(val_dtype);
  if (!esz || val_dtype == GG_UINT64) return set_error(GG_ERR_ARGUMENT, "bad value type");
  if (slice_size < 1) return set_error(GG_ERR_ARGUMENT, "slice_size must be >= 1");
  const bool relabel = (flags & GG_SELL_RELABEL) != 0;
  if (relabel && m != n) return set_error(GG_ERR_ARGUMENT, "relabelling needs a square matrix");
  if (flags & GG_SELL_RCM)
    return set_error(GG_ERR_ARGUMENT, "reverse Cuthill-McKee ties are built on the host (gg_sell_build)");
  if (heavy_threshold > 0 && heavy_seg_len < 1)
    return set_error(GG_ERR_ARGUMENT, "heavy_seg_len must be >= 1");
  if (m > 2147483646LL) return set_error(GG_ERR_ARGUMENT, "rows exceed Int32");
  GG_HIP(hipSetDevice(ctx->device));
  int32_t rp[2] = {1, 1};
  if (m > 0) {
    GG_HIP(hipMemcpyAsync(ctx->host_pinned, rowptr, sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    GG_HIP(hipMemcpyAsync((char*)ctx->host_pinned + 4, rowptr + m, sizeof(int32_t),
                          hipMemcpyDeviceToHost, ctx->stream));
    GG_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(rp, ctx->host_pinned, sizeof rp);
  }
  if (rp[0] != 1) return set_error(GG_ERR_ARGUMENT, "rowptr[1] must be 1");
  gg_sell_plan* p = new gg_sell_plan();
  p->ctx = ctx;
  p->m = m;
  p->n = n;
  p->nnz = (int64_t)rp[1] - 1;
  p->val_dtype = val_dtype;
  p->rowptr = rowptr;
  p->colval = colval;
  p->nzval = nzval;
  p->seg_len = heavy_seg_len;
  p->relabel = relabel;
  p->sort_entries = (flags & GG_SELL_SORT_ENTRIES) != 0;
  p->reference = sigma <= 1 && heavy_threshold <= 0 && !relabel;
  const int st = plan_build(p, slice_size, sigma, heavy_threshold);
  if (st != GG_OK) {
    gg_sell_plan_free(p);
    return st;
  }
  *sizes = p->sizes;
  *plan = p;
  return GG_OK;
}

int gg_sell_fill_device(gg_sell_plan* p, gg_matrix* out) {
  if (!p || !out) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  clear_error();
  const gg_sell_sizes& z = p->sizes;
  if (!out->ptr || (z.n_stored > 0 && (!out->idx || !out->val)) ||
      (z.has_perm && z.n_slots > 0 && (!out->perm || !out->slot_len)) ||
      (z.n_heavy > 0 && (!out->heavy_row || !out->heavy_ptr || !out->heavy_seg ||
                         (z.heavy_nnz > 0 && (!out->heavy_col || !out->heavy_val)))) ||
      (z.has_vinv && p->m > 0 && !out->vinv))
    return set_error(GG_ERR_ARGUMENT, "an output array of the plan's sizes is missing");
  gg_ctx* ctx = p->ctx;
  GG_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const EntryMap em{p->colval, p->relabel ? p->vinv : nullptr, p->sorted_col, p->sorted_pos};
  if (z.n_slots == 0) {
    const int32_t one = 1;
    GG_HIP(hipMemcpyAsync((int32_t*)out->ptr, &one, sizeof one, hipMemcpyHostToDevice, st));
    GG_HIP(hipStreamSynchronize(st));  // `one` lives on this stack frame
  }
  int s = with_type(p->val_dtype, [&](auto t) -> int {
    typedef typename decltype(t)::type T;
    if (z.n_slots > 0)
      hipLaunchKernelGGL(slots_kernel<T>, dim3(blocks_for(z.n_slots, kB)), dim3(kB), 0, st,
                         p->rowptr, em, (const T*)p->nzval, p->slot_rows, p->nl, z.n_slots,
                         p->C, p->wscan, (int32_t*)out->ptr, (int32_t*)out->idx,
                         (T*)out->val, z.has_perm ? (int32_t*)out->perm : nullptr,
                         z.has_perm ? (int32_t*)out->slot_len : nullptr);
    if (z.n_heavy > 0)
      hipLaunchKernelGGL(heavy_kernel<T>, dim3((unsigned)z.n_heavy), dim3(kB), 0, st, p->rowptr,
                         em, (const T*)p->nzval, p->heavy, z.n_heavy, p->escan, p->gscan,
                         (int32_t*)out->heavy_row, (int64_t*)out->heavy_ptr,
                         (int32_t*)out->heavy_seg, (int32_t*)out->heavy_col,
                         (T*)out->heavy_val);
    GG_HIP(hipGetLastError());
    return GG_OK;
  });
  GG_TRY(s);
  if (z.has_vinv && p->m > 0)
    GG_HIP(hipMemcpyAsync((int32_t*)out->vinv, p->vinv, sizeof(int32_t) * p->m,
                          hipMemcpyDeviceToDevice, st));
  GG_HIP(hipStreamSynchronize(st));
  out->format = GG_SELL;
  out->val_dtype = p->val_dtype;
  out->m = p->m;
  out->n = p->n;
  out->nnz = p->nnz;
  out->slice_size = z.slice_size;
  out->nslices = z.nslices;
  out->n_stored = z.n_stored;
  if (!z.has_perm) {
    out->perm = nullptr;
    out->slot_len = nullptr;
  }
  out->n_heavy = z.n_heavy;
  out->heavy_seg_len = z.n_heavy > 0 ? p->seg_len : 0;
  out->heavy_nseg = z.n_seg;
  out->flags = p->relabel ? GG_FLAG_INTERNAL_IDS : 0;
  if (!z.has_vinv) out->vinv = nullptr;
  return GG_OK;
}

void gg_sell_plan_free(gg_sell_plan* p) {
  if (!p) return;
  if (p->ctx) (void)hipStreamSynchronize(p->ctx->stream);
  for (void* a : p->allocs) (void)hipFree(a);
  delete p;
}

}  // extern "C"
