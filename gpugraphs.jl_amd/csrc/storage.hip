// storage.hip -- host-side validation and the SELL-C-sigma builder.
//
// gg_sell_build lays out slices as src/storage/SELL.jl:99-181 does (slice of
// C rows padded to its longest row, entry k of every row contiguous, pad col
// -1 / val 0).  With sigma > 1 rows are first sorted by length (longest first,
// ties in row order) inside windows of sigma rows, the ordering the
// reference leaves commented out at src/storage/SELL.jl:113-116; with a heavy
// threshold the rows longer than it are kept in a CSR partition processed in
// fixed-size segments, so one hub row does not widen a whole slice.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <numeric>
#include <thread>
#include <vector>

#include "internal.hpp"

using namespace gg;

namespace {

template <class F> void parallel_for(int64_t n, int64_t grain, F&& f) {
  unsigned nt = std::max(1u, std::thread::hardware_concurrency());
  if (n <= grain || nt == 1) {
    f(0, n);
    return;
  }
  std::atomic<int64_t> next{0};
  auto worker = [&]() {
    while (true) {
      int64_t b = next.fetch_add(grain);
      if (b >= n) break;
      f(b, std::min(n, b + grain));
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) th.emplace_back(worker);
  worker();
  for (auto& t : th) t.join();
}

}  // namespace

extern "C" {

int gg_csr_check(int64_t m, int64_t n, const int64_t* rowptr, int64_t rowptr_len,
                 const int64_t* colval, int64_t colval_len, int64_t nzval_len) {
  // src/storage/CSR.jl:48-63, same order
  if (rowptr_len != m + 1) return set_error(GG_ERR_ARGUMENT, "length(rowptr) must be equal to m + 1");
  if (colval_len != nzval_len)
    return set_error(GG_ERR_ARGUMENT, "length(colval) must be equal to length(nzval)");
  if (colval_len > 0) {
    int64_t mx = colval[0], mn = colval[0];
    for (int64_t i = 1; i < colval_len; ++i) {
      mx = std::max(mx, colval[i]);
      mn = std::min(mn, colval[i]);
    }
    if (mx > n || mn < 1) return set_error(GG_ERR_ARGUMENT, "colval contains an index out of bounds");
  }
  if (rowptr[0] != 1 || rowptr[m] != colval_len + 1)
    return set_error(GG_ERR_ARGUMENT,
                     "rowptr[1] must be equal to 1 and rowptr[m+1] must be equal to length(colval) + 1");
  // Beyond the reference's checks: the kernels index colval through rowptr,
  // so a decreasing pointer would send them outside the arrays.
  for (int64_t i = 0; i < m; ++i)
    if (rowptr[i + 1] < rowptr[i]) return set_error(GG_ERR_ARGUMENT, "rowptr must be non-decreasing");
  return GG_OK;
}

int gg_sell_check(int64_t n, int64_t nslices, int64_t n_stored, int64_t slice_ptr_len,
                  const int64_t* colval, int64_t colval_len, int64_t nzval_len) {
  // src/storage/SELL.jl:42-52, same order
  if (slice_ptr_len != nslices + 1)
    return set_error(GG_ERR_ARGUMENT, "length(slice_ptr) must be equal to nslices+1");
  if (colval_len != nzval_len || colval_len != n_stored)
    return set_error(GG_ERR_ARGUMENT, "length(colval) and length(nzval) must be equal to n_stored");
  if (colval_len > 0) {
    int64_t mx = colval[0], mn = colval[0];
    for (int64_t i = 1; i < colval_len; ++i) {
      mx = std::max(mx, colval[i]);
      mn = std::min(mn, colval[i]);
    }
    if (mx > n || mn < -1) return set_error(GG_ERR_ARGUMENT, "colval contains an index out of bounds");
  }
  return GG_OK;
}

void gg_sell_host_free(gg_sell_host* h) {
  if (!h) return;
  free(h->slice_ptr);
  free(h->colval);
  free(h->nzval);
  free(h->perm);
  free(h->slot_len);
  free(h->heavy_row);
  free(h->heavy_ptr);
  free(h->heavy_seg);
  free(h->heavy_col);
  free(h->heavy_val);
  free(h->vinv);
  delete h;
}

}  // extern "C"

namespace gg {

// Reverse Cuthill-McKee order (George and Liu, "Computer Solution of Large
// Sparse Positive Definite Systems", 1981): components are taken from their
// unvisited vertex of lowest degree (ties by id), each is walked breadth
// first with every vertex's unvisited neighbours appended by increasing
// degree (ties by id), and the whole sequence is reversed.  Neighbours are
// the row's columns (for a symmetric pattern, the vertex's neighbours).
// out[i] = 1-based vertex at position i.
static void rcm_order(int64_t m, const int32_t* rowptr, const int32_t* colval, int32_t* out) {
  std::vector<int32_t> deg(m);
  int32_t maxd = 0;
  for (int64_t v = 0; v < m; ++v) {
    deg[v] = rowptr[v + 1] - rowptr[v];
    maxd = std::max(maxd, deg[v]);
  }
  // vertices by increasing degree, ties by id (counting sort)
  std::vector<int64_t> cnt((size_t)maxd + 2, 0);
  for (int64_t v = 0; v < m; ++v) ++cnt[(size_t)deg[v] + 1];
  for (size_t d = 1; d < cnt.size(); ++d) cnt[d] += cnt[d - 1];
  std::vector<int32_t> by_deg(m);
  for (int64_t v = 0; v < m; ++v) by_deg[cnt[deg[v]]++] = (int32_t)v;
  std::vector<uint8_t> seen(m, 0);
  std::vector<int32_t> nb;
  int64_t head = 0, tail = 0;  // out[0 .. m) is the breadth-first queue
  for (int64_t i = 0; i < m; ++i) {
    const int32_t s0 = by_deg[i];
    if (seen[s0]) continue;
    seen[s0] = 1;
    out[tail++] = s0;
    while (head < tail) {
      const int32_t u = out[head++];
      nb.clear();
      for (int64_t e = rowptr[u] - 1; e < rowptr[u + 1] - 1; ++e) {
        const int32_t c = colval[e] - 1;
        if (c >= 0 && c < m && !seen[c]) {
          seen[c] = 1;
          nb.push_back(c);
        }
      }
      std::sort(nb.begin(), nb.end(), [&](int32_t x, int32_t y) {
        return deg[x] != deg[y] ? deg[x] < deg[y] : x < y;
      });
      for (int32_t c : nb) out[tail++] = c;
    }
  }
  std::reverse(out, out + m);
  for (int64_t i = 0; i < m; ++i) out[i] += 1;
}

}  // namespace gg

extern "C" {

int gg_rcm_order(int64_t m, const int32_t* rowptr, const int32_t* colval, int32_t* order) {
  if (m < 0 || (m > 0 && (!rowptr || !order || (!colval && rowptr[m] > 1))))
    return set_error(GG_ERR_ARGUMENT, "NULL or negative argument");
  clear_error();
  if (m > 2147483646LL) return set_error(GG_ERR_ARGUMENT, "rows exceed Int32");
  if (m > 0 && rowptr[0] != 1) return set_error(GG_ERR_ARGUMENT, "rowptr[1] must be 1");
  rcm_order(m, rowptr, colval, order);
  return GG_OK;
}

int gg_sell_build(int64_t m, int64_t n, const int32_t* rowptr, const int32_t* colval,
                  const void* nzval, int val_dtype, int64_t slice_size, int64_t sigma,
                  int64_t heavy_threshold, int64_t heavy_seg_len, int flags,
                  gg_sell_host** out) {
  if (!out || (m > 0 && !rowptr)) return set_error(GG_ERR_ARGUMENT, "NULL argument");
  const size_t esz = dtype_size(val_dtype);
  if (!esz || val_dtype == GG_UINT64) return set_error(GG_ERR_ARGUMENT, "bad value type");
  if (slice_size < 1) return set_error(GG_ERR_ARGUMENT, "slice_size must be >= 1");
  if (rowptr && rowptr[0] != 1) return set_error(GG_ERR_ARGUMENT, "rowptr[1] must be 1");
  const bool relabel = (flags & GG_SELL_RELABEL) != 0;
  if (relabel && m != n) return set_error(GG_ERR_ARGUMENT, "relabelling needs a square matrix");
  if ((flags & GG_SELL_RCM) && !relabel)
    return set_error(GG_ERR_ARGUMENT, "GG_SELL_RCM orders the ties of GG_SELL_RELABEL");
  const bool reference_layout = sigma <= 1 && heavy_threshold <= 0 && !relabel;
  if (heavy_threshold > 0 && heavy_seg_len < 1)
    return set_error(GG_ERR_ARGUMENT, "heavy_seg_len must be >= 1");

  gg_sell_host* h = new gg_sell_host();
  memset(h, 0, sizeof *h);
  h->m = m;
  h->n = n;
  h->nnz = m > 0 ? (int64_t)rowptr[m] - 1 : 0;
  h->sigma = sigma;
  h->val_dtype = val_dtype;
  h->heavy_seg_len = heavy_seg_len;
  auto fail = [&](int st, const char* msg) {
    gg_sell_host_free(h);
    return set_error(st, "%s", msg);
  };

  // Rows that go into slices, in slot order.
  std::vector<int32_t> slot_rows;  // 1-based row of each slot (light rows only)
  std::vector<int32_t> heavy;
  int64_t C = slice_size;
  if (reference_layout) {
    C = m > 0 ? std::min<int64_t>(slice_size, m) : slice_size;  // src/storage/SELL.jl:108
    slot_rows.resize(m);
    std::iota(slot_rows.begin(), slot_rows.end(), 1);
  } else {
    slot_rows.reserve(m);
    std::vector<int32_t> order(m);
    if (flags & GG_SELL_RCM) rcm_order(m, rowptr, colval, order.data());
    else std::iota(order.begin(), order.end(), 1);
    if (relabel)  // every row by decreasing length, ties in row (or RCM) order
      std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return (rowptr[x] - rowptr[x - 1]) > (rowptr[y] - rowptr[y - 1]);
      });
    for (int32_t r : order) {
      int64_t len = rowptr[r] - rowptr[r - 1];
      if (heavy_threshold > 0 && len > heavy_threshold) heavy.push_back((int32_t)r);
      else slot_rows.push_back((int32_t)r);
    }
    if (sigma > 1 && !relabel) {
      const int64_t nl = (int64_t)slot_rows.size();
      parallel_for((nl + sigma - 1) / sigma, 64, [&](int64_t w0, int64_t w1) {
        for (int64_t w = w0; w < w1; ++w) {
          auto b = slot_rows.begin() + w * sigma;
          auto e = slot_rows.begin() + std::min(nl, (w + 1) * sigma);
          std::stable_sort(b, e, [&](int32_t x, int32_t y) {
            return (rowptr[x] - rowptr[x - 1]) > (rowptr[y] - rowptr[y - 1]);
          });
        }
      });
    }
  }
  const int64_t nl = (int64_t)slot_rows.size();
  const int64_t ns = nl > 0 ? (nl + C - 1) / C : 0;
  h->slice_size = C;
  h->nslices = ns;
  h->n_slots = ns * C;

  // Relabelled ids: a light row's internal id is its slot (1-based), a heavy
  // row's is n_slots + its heavy index; stored columns use internal ids, so
  // the hubs' entries of x sit together at the front of the slot space.
  std::vector<int32_t> vinv;
  const int64_t nh_all = (int64_t)heavy.size();
  if (relabel) {
    if (ns * C + nh_all > 2147483647LL) return fail(GG_ERR_ARGUMENT, "internal ids exceed Int32");
    vinv.assign(m, 0);
    for (int64_t q = 0; q < nl; ++q) vinv[slot_rows[q] - 1] = (int32_t)(q + 1);
    for (int64_t i = 0; i < nh_all; ++i) vinv[heavy[i] - 1] = (int32_t)(ns * C + i + 1);
    h->vinv = (int32_t*)malloc(sizeof(int32_t) * std::max<int64_t>(1, m));
    if (!h->vinv) return fail(GG_ERR_NOMEM, "host allocation failed");
    memcpy(h->vinv, vinv.data(), sizeof(int32_t) * m);
    h->n_internal = ns * C + nh_all;
  } else {
    h->n_internal = m;
  }
  // entries of a row in storage order: (stored column, source position); with
  // relabelled ids the columns are mapped but keep the row's input order, so
  // every fold visits a row's terms in the order the reference's SELL does
  const bool sort_entries = (flags & GG_SELL_SORT_ENTRIES) != 0;
  auto row_entries = [&](int32_t row, std::vector<std::pair<int32_t, int64_t>>& ent) {
    const int64_t start = rowptr[row - 1] - 1, len = rowptr[row] - rowptr[row - 1];
    ent.resize(len);
    for (int64_t t = 0; t < len; ++t) {
      const int32_t c = colval[start + t];
      ent[t] = {relabel ? vinv[c - 1] : c, start + t};
    }
    if (sort_entries) std::stable_sort(ent.begin(), ent.end());
  };

  // slice widths and pointers (src/storage/SELL.jl:117-131)
  std::vector<int64_t> width(ns, 0);
  parallel_for(ns, 1024, [&](int64_t s0, int64_t s1) {
    for (int64_t s = s0; s < s1; ++s) {
      int64_t mx = 0;
      for (int64_t q = s * C; q < std::min(nl, (s + 1) * C); ++q) {
        int32_t r = slot_rows[q];
        mx = std::max<int64_t>(mx, rowptr[r] - rowptr[r - 1]);
      }
      width[s] = mx;
    }
  });
  h->slice_ptr = (int32_t*)malloc(sizeof(int32_t) * (ns + 1));
  if (!h->slice_ptr) return fail(GG_ERR_NOMEM, "host allocation failed");
  int64_t acc = 1;
  h->slice_ptr[0] = 1;
  for (int64_t s = 0; s < ns; ++s) {
    acc += width[s] * C;
    if (acc > 2147483647LL)
      return fail(GG_ERR_ARGUMENT, "SELL storage exceeds Int32 slice_ptr; raise sigma or use a heavy threshold");
    h->slice_ptr[s + 1] = (int32_t)acc;
  }
  h->n_stored = acc - 1;
  h->colval = (int32_t*)malloc(sizeof(int32_t) * std::max<int64_t>(1, h->n_stored));
  h->nzval = malloc(esz * std::max<int64_t>(1, h->n_stored));
  if (!h->colval || !h->nzval) return fail(GG_ERR_NOMEM, "host allocation failed");
  if (!reference_layout) {
    h->perm = (int32_t*)calloc(std::max<int64_t>(1, h->n_slots), sizeof(int32_t));
    h->slot_len = (int32_t*)calloc(std::max<int64_t>(1, h->n_slots), sizeof(int32_t));
    if (!h->perm || !h->slot_len) return fail(GG_ERR_NOMEM, "host allocation failed");
  }
  const char* src_val = (const char*)nzval;
  char* dst_val = (char*)h->nzval;
  parallel_for(ns, 256, [&](int64_t s0, int64_t s1) {
    std::vector<std::pair<int32_t, int64_t>> ent;
    for (int64_t s = s0; s < s1; ++s) {
      const int64_t base = h->slice_ptr[s] - 1;
      for (int64_t r = 0; r < C; ++r) {
        const int64_t q = s * C + r;
        int64_t len = 0;
        ent.clear();
        if (q < nl) {
          int32_t row = slot_rows[q];
          row_entries(row, ent);
          len = (int64_t)ent.size();
          if (h->perm) {
            h->perm[q] = row;
            h->slot_len[q] = (int32_t)len;
          }
        }
        for (int64_t k = 0; k < width[s]; ++k) {
          const int64_t pos = base + k * C + r;  // column-major inside the slice
          if (k < len) {
            h->colval[pos] = ent[k].first;
            memcpy(dst_val + pos * esz, src_val + ent[k].second * esz, esz);
          } else {
            h->colval[pos] = -1;
            memset(dst_val + pos * esz, 0, esz);
          }
        }
      }
    }
  });

  // heavy partition
  const int64_t nh = (int64_t)heavy.size();
  h->n_heavy = nh;
  if (nh > 0) {
    h->heavy_row = (int32_t*)malloc(sizeof(int32_t) * nh);
    h->heavy_ptr = (int64_t*)malloc(sizeof(int64_t) * (nh + 1));
    h->heavy_seg = (int32_t*)malloc(sizeof(int32_t) * (nh + 1));
    if (!h->heavy_row || !h->heavy_ptr || !h->heavy_seg) return fail(GG_ERR_NOMEM, "host allocation failed");
    h->heavy_ptr[0] = 0;
    h->heavy_seg[0] = 0;
    for (int64_t i = 0; i < nh; ++i) {
      int32_t r = heavy[i];
      int64_t len = rowptr[r] - rowptr[r - 1];
      h->heavy_row[i] = r;
      h->heavy_ptr[i + 1] = h->heavy_ptr[i] + len;
      int64_t segs = (len + heavy_seg_len - 1) / heavy_seg_len;
      if ((int64_t)h->heavy_seg[i] + segs > 2147483647LL) return fail(GG_ERR_ARGUMENT, "too many heavy segments");
      h->heavy_seg[i + 1] = (int32_t)(h->heavy_seg[i] + segs);
    }
    h->heavy_nnz = h->heavy_ptr[nh];
    h->n_seg = h->heavy_seg[nh];
    h->heavy_col = (int32_t*)malloc(sizeof(int32_t) * std::max<int64_t>(1, h->heavy_nnz));
    h->heavy_val = malloc(esz * std::max<int64_t>(1, h->heavy_nnz));
    if (!h->heavy_col || !h->heavy_val) return fail(GG_ERR_NOMEM, "host allocation failed");
    parallel_for(nh, 64, [&](int64_t i0, int64_t i1) {
      std::vector<std::pair<int32_t, int64_t>> ent;
      for (int64_t i = i0; i < i1; ++i) {
        row_entries(heavy[i], ent);
        const int64_t o = h->heavy_ptr[i];
        for (size_t t = 0; t < ent.size(); ++t) {
          h->heavy_col[o + t] = ent[t].first;
          memcpy((char*)h->heavy_val + (o + t) * esz, src_val + ent[t].second * esz, esz);
        }
      }
    });
  }
  *out = h;
  return GG_OK;
}

}  // extern "C"
