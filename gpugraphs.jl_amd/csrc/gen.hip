// gen.hip -- seeded synthetic graph generators (host).
//
// Every random quantity is a hash of (seed, ids), so a graph is a pure
// function of its parameters whatever the thread count.  Outputs are CSR of
// A_T in the reference's convention (row v lists the in-neighbours u of v,
// 1-based Int32, sorted, no duplicates; test/bfs.jl:24 builds A_T the same
// way), with values of the requested type.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "internal.hpp"

using namespace gg;

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
inline uint64_t h2(uint64_t seed, uint64_t a) { return mix64(seed ^ mix64(a)); }
inline uint64_t h3(uint64_t seed, uint64_t a, uint64_t b) {
  return mix64(h2(seed, a) ^ (b * 0xD1342543DE82EF95ULL));
}
inline float u01f(uint64_t r) { return (float)(r >> 40) * (1.0f / 16777216.0f); }  // 24 bits: exact fp32 in [0,1)
inline double u01d(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

const uint64_t SALT_PERM = 0x5045524d5045524dULL, SALT_VAL = 0x56414c5545534545ULL,
               SALT_LR = 0x4c4f4e4752414e47ULL;

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

// Value of the stored entry for the directed edge u -> v (0-based ids).
struct ValueRule {
  int dtype;
  int kind;  // 0: unit (Bool true / 1), 1: U[0,1), 2: integer weight in [wmin, wmax]
  int32_t wmin, wmax;
  uint64_t seed;
  void put(void* out, int64_t pos, int64_t u, int64_t v) const {
    double x;
    if (kind == 0) x = 1.0;
    else if (kind == 1) x = u01f(h3(seed ^ SALT_VAL, (uint64_t)u, (uint64_t)v));
    else x = (double)(wmin + (int64_t)(h3(seed ^ SALT_VAL, (uint64_t)u, (uint64_t)v) %
                                       (uint64_t)(wmax - wmin + 1)));
    switch (dtype) {
      case GG_BOOL: ((uint8_t*)out)[pos] = x != 0.0; break;
      case GG_INT32: ((int32_t*)out)[pos] = (int32_t)x; break;
      case GG_INT64: ((int64_t*)out)[pos] = (int64_t)x; break;
      case GG_FLOAT32: ((float*)out)[pos] = (float)x; break;
      case GG_FLOAT64: ((double*)out)[pos] = kind == 1 ? (double)(float)x : x; break;
    }
  }
};

// Directed edge list (0-based src -> dst) -> CSR of A_T: row dst holds src.
// Drops self-loops when asked, sorts rows, removes duplicates.
int edges_to_csr(int64_t nv, const std::vector<int32_t>& src, const std::vector<int32_t>& dst,
                 bool drop_self, const ValueRule& vr, gg_csr_host** out) {
  const int64_t E = (int64_t)src.size();
  std::vector<std::atomic<int32_t>> cnt(nv);
  for (int64_t i = 0; i < nv; ++i) cnt[i].store(0, std::memory_order_relaxed);
  parallel_for(E, 1 << 16, [&](int64_t a, int64_t b) {
    for (int64_t e = a; e < b; ++e)
      if (!(drop_self && src[e] == dst[e])) cnt[dst[e]].fetch_add(1, std::memory_order_relaxed);
  });
  std::vector<int64_t> off(nv + 1, 0);
  for (int64_t i = 0; i < nv; ++i) off[i + 1] = off[i] + cnt[i].load(std::memory_order_relaxed);
  std::vector<int32_t> tmp(std::max<int64_t>(1, off[nv]));
  std::vector<std::atomic<int64_t>> cur(nv);
  for (int64_t i = 0; i < nv; ++i) cur[i].store(off[i], std::memory_order_relaxed);
  parallel_for(E, 1 << 16, [&](int64_t a, int64_t b) {
    for (int64_t e = a; e < b; ++e)
      if (!(drop_self && src[e] == dst[e]))
        tmp[cur[dst[e]].fetch_add(1, std::memory_order_relaxed)] = src[e];
  });
  std::vector<int64_t> len(nv);
  parallel_for(nv, 4096, [&](int64_t a, int64_t b) {
    for (int64_t v = a; v < b; ++v) {
      auto s = tmp.begin() + off[v], e = tmp.begin() + off[v + 1];
      std::sort(s, e);
      len[v] = std::unique(s, e) - s;
    }
  });
  gg_csr_host* h = new gg_csr_host();
  h->m = h->n = nv;
  h->val_dtype = vr.dtype;
  h->rowptr = (int32_t*)malloc(sizeof(int32_t) * (nv + 1));
  int64_t acc = 1;
  h->rowptr[0] = 1;
  for (int64_t v = 0; v < nv; ++v) {
    acc += len[v];
    if (acc > 2147483647LL) {
      free(h->rowptr);
      delete h;
      return set_error(GG_ERR_ARGUMENT, "graph exceeds Int32 row pointers");
    }
    h->rowptr[v + 1] = (int32_t)acc;
  }
  h->nnz = acc - 1;
  h->colval = (int32_t*)malloc(sizeof(int32_t) * std::max<int64_t>(1, h->nnz));
  h->nzval = malloc(dtype_size(vr.dtype) * std::max<int64_t>(1, h->nnz));
  if (!h->colval || !h->nzval) {
    gg_csr_host_free(h);
    return set_error(GG_ERR_NOMEM, "host allocation failed");
  }
  parallel_for(nv, 4096, [&](int64_t a, int64_t b) {
    for (int64_t v = a; v < b; ++v) {
      int64_t o = h->rowptr[v] - 1;
      for (int64_t k = 0; k < len[v]; ++k) {
        int32_t u = tmp[off[v] + k];
        h->colval[o + k] = u + 1;
        vr.put(h->nzval, o + k, u, v);
      }
    }
  });
  *out = h;
  return GG_OK;
}

}  // namespace

extern "C" {

void gg_csr_host_free(gg_csr_host* h) {
  if (!h) return;
  free(h->rowptr);
  free(h->colval);
  free(h->nzval);
  delete h;
}

int gg_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                int permute, int val_dtype, int values, gg_csr_host** out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0 || !dtype_size(val_dtype))
    return set_error(GG_ERR_ARGUMENT, "bad R-MAT parameters");
  const int64_t nv = 1LL << scale;
  const int64_t E = (int64_t)edge_factor << scale;
  std::vector<int32_t> perm(nv);
  for (int64_t i = 0; i < nv; ++i) perm[i] = (int32_t)i;
  if (permute)
    for (int64_t i = nv - 1; i > 0; --i)  // Fisher-Yates, hashed draws
      std::swap(perm[i], perm[h2(seed ^ SALT_PERM, (uint64_t)i) % (uint64_t)(i + 1)]);
  // symmetrised: both directions of every drawn edge
  std::vector<int32_t> src(2 * E), dst(2 * E);
  const double ab = a + b, abc = a + b + c;
  parallel_for(E, 1 << 14, [&](int64_t e0, int64_t e1) {
    for (int64_t e = e0; e < e1; ++e) {
      int64_t u = 0, v = 0;
      for (int l = 0; l < scale; ++l) {
        double p = u01d(h3(seed, (uint64_t)e, (uint64_t)l));
        int bu = p >= ab, bv = (p >= a && p < ab) || p >= abc;
        u = (u << 1) | bu;
        v = (v << 1) | bv;
      }
      int32_t pu = perm[u], pv = perm[v];
      src[2 * e] = pu; dst[2 * e] = pv;
      src[2 * e + 1] = pv; dst[2 * e + 1] = pu;
    }
  });
  ValueRule vr{val_dtype, values, 1, 1, seed};
  return edges_to_csr(nv, src, dst, true, vr, out);
}

int gg_gen_grid(int64_t rows, int64_t cols, double long_range_frac, int32_t wmin, int32_t wmax,
                uint64_t seed, int val_dtype, gg_csr_host** out) {
  if (!out || rows < 1 || cols < 1 || rows * cols > 2147483647LL || wmax < wmin ||
      long_range_frac < 0 || !dtype_size(val_dtype))
    return set_error(GG_ERR_ARGUMENT, "bad grid parameters");
  const int64_t nv = rows * cols;
  const int64_t und = rows * (cols - 1) + cols * (rows - 1);
  const int64_t nlr = (int64_t)(long_range_frac * (double)und + 0.5);
  std::vector<int32_t> src, dst;
  src.reserve(2 * (und + nlr));
  dst.reserve(2 * (und + nlr));
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t cc = 0; cc < cols; ++cc) {
      int32_t v = (int32_t)(r * cols + cc);
      if (cc + 1 < cols) { src.push_back(v); dst.push_back(v + 1); src.push_back(v + 1); dst.push_back(v); }
      if (r + 1 < rows) {
        int32_t w = (int32_t)(v + cols);
        src.push_back(v); dst.push_back(w); src.push_back(w); dst.push_back(v);
      }
    }
  for (int64_t i = 0; i < nlr; ++i) {  // long-range pairs, u != v, both directions
    uint64_t t = 0;
    int64_t u, v;
    do {
      u = (int64_t)(h3(seed ^ SALT_LR, (uint64_t)i, 2 * t) % (uint64_t)nv);
      v = (int64_t)(h3(seed ^ SALT_LR, (uint64_t)i, 2 * t + 1) % (uint64_t)nv);
      ++t;
    } while (u == v);
    src.push_back((int32_t)u); dst.push_back((int32_t)v);
    src.push_back((int32_t)v); dst.push_back((int32_t)u);
  }
  ValueRule vr{val_dtype, 2, wmin, wmax, seed};
  return edges_to_csr(nv, src, dst, true, vr, out);
}

int gg_gen_uniform_rows(int64_t n, int64_t per_row, uint64_t seed, int val_dtype,
                        gg_csr_host** out) {
  if (!out || n < 1 || per_row < 0 || per_row > n || n * per_row > 2147483646LL ||
      !dtype_size(val_dtype))
    return set_error(GG_ERR_ARGUMENT, "bad uniform-rows parameters");
  gg_csr_host* h = new gg_csr_host();
  h->m = h->n = n;
  h->nnz = n * per_row;
  h->val_dtype = val_dtype;
  h->rowptr = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
  h->colval = (int32_t*)malloc(sizeof(int32_t) * std::max<int64_t>(1, h->nnz));
  h->nzval = malloc(dtype_size(val_dtype) * std::max<int64_t>(1, h->nnz));
  if (!h->rowptr || !h->colval || !h->nzval) {
    gg_csr_host_free(h);
    return set_error(GG_ERR_NOMEM, "host allocation failed");
  }
  for (int64_t r = 0; r <= n; ++r) h->rowptr[r] = (int32_t)(1 + r * per_row);
  ValueRule vr{val_dtype, 1, 1, 1, seed};
  parallel_for(n, 1024, [&](int64_t a, int64_t b) {
    std::vector<int32_t> cs;
    for (int64_t r = a; r < b; ++r) {
      cs.clear();
      uint64_t t = 0;
      while ((int64_t)cs.size() < per_row) {  // distinct columns, drawn without replacement
        int32_t c = (int32_t)(h3(seed, (uint64_t)r, t++) % (uint64_t)n);
        if (std::find(cs.begin(), cs.end(), c) == cs.end()) cs.push_back(c);
      }
      std::sort(cs.begin(), cs.end());
      for (int64_t k = 0; k < per_row; ++k) {
        h->colval[r * per_row + k] = cs[k] + 1;
        vr.put(h->nzval, r * per_row + k, cs[k], r);
      }
    }
  });
  *out = h;
  return GG_OK;
}

int gg_gen_random_graph(int64_t n, int64_t degree, uint64_t seed, int val_dtype, int values,
                        int32_t wmin, int32_t wmax, gg_csr_host** out) {
  if (!out || n < 2 || degree < 1 || !dtype_size(val_dtype))
    return set_error(GG_ERR_ARGUMENT, "bad random-graph parameters");
  const int64_t ne = n * degree / 2;  // undirected draws; symmetrised below
  std::vector<int32_t> src(2 * ne), dst(2 * ne);
  parallel_for(ne, 1 << 14, [&](int64_t a, int64_t b) {
    for (int64_t e = a; e < b; ++e) {
      int32_t u = (int32_t)(h3(seed, (uint64_t)e, 0) % (uint64_t)n);
      int32_t v = (int32_t)(h3(seed, (uint64_t)e, 1) % (uint64_t)n);
      src[2 * e] = u; dst[2 * e] = v; src[2 * e + 1] = v; dst[2 * e + 1] = u;
    }
  });
  ValueRule vr{val_dtype, values, wmin, wmax, seed};
  return edges_to_csr(n, src, dst, true, vr, out);
}

}  // extern "C"
