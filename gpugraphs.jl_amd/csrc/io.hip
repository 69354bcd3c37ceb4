// io.hip -- graph and matrix file readers (host): MatrixMarket coordinate
// files and plain edge lists, into the 1-based Int32 CSR of gg_csr_host.
//
// The reference reads its real-world inputs as Rutherford-Boeing archives
// fetched from the SuiteSparse collection (benchmark/benchmarks.jl:227-231,
// 262-263), which needs the network; these readers take the same matrices
// from local files instead.
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace gg;

namespace {

struct Entry {
  int64_t row, col;  // 0-based
  double dv;         // real / pattern value
  int64_t iv;        // integer value
};

// Whole file in memory, NUL-terminated.
int slurp(const char* path, std::string& buf) {
  FILE* f = fopen(path, "rb");
  if (!f) return set_error(GG_ERR_ARGUMENT, "cannot open %s: %s", path, strerror(errno));
  if (fseek(f, 0, SEEK_END) != 0) {
    fclose(f);
    return set_error(GG_ERR_ARGUMENT, "cannot seek %s", path);
  }
  const long len = ftell(f);
  rewind(f);
  buf.resize(len > 0 ? (size_t)len : 0);
  const size_t got = len > 0 ? fread(&buf[0], 1, (size_t)len, f) : 0;
  fclose(f);
  if ((long)got != len) return set_error(GG_ERR_ARGUMENT, "short read of %s", path);
  return GG_OK;
}

struct Cursor {
  const char* p;
  const char* end;
  int64_t line = 1;
  void skip_blank() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  }
  bool at_eol() {
    skip_blank();
    return p >= end || *p == '\n';
  }
  void next_line() {
    while (p < end && *p != '\n') ++p;
    if (p < end) {
      ++p;
      ++line;
    }
  }
  // skips empty lines and lines starting with one of `comment`
  bool next_data_line(const char* comment) {
    while (p < end) {
      skip_blank();
      if (p < end && (*p == '\n' || strchr(comment, *p))) {
        next_line();
        continue;
      }
      return p < end;
    }
    return false;
  }
  bool read_i64(int64_t& v) {
    skip_blank();
    char* e;
    errno = 0;
    v = strtoll(p, &e, 10);
    if (e == p || errno) return false;
    p = e;
    return true;
  }
  bool read_f64(double& v) {
    skip_blank();
    char* e;
    v = strtod(p, &e);
    if (e == p) return false;
    p = e;
    return true;
  }
};

// Entries -> CSR with rows sorted by column; duplicates combined by `comb`.
// Values are written in `dtype` from the double or the integer value.
template <class Comb>
int build_csr(int64_t m, int64_t n, std::vector<Entry>& ents, int dtype, bool integer, Comb comb,
              gg_csr_host** out) {
  std::sort(ents.begin(), ents.end(), [](const Entry& a, const Entry& b) {
    return a.row != b.row ? a.row < b.row : a.col < b.col;
  });
  size_t w = 0;
  for (size_t i = 0; i < ents.size(); ++i) {
    if (w > 0 && ents[w - 1].row == ents[i].row && ents[w - 1].col == ents[i].col)
      comb(ents[w - 1], ents[i]);
    else
      ents[w++] = ents[i];
  }
  ents.resize(w);
  if ((int64_t)ents.size() > 2147483646LL)
    return set_error(GG_ERR_ARGUMENT, "matrix exceeds Int32 row pointers");
  gg_csr_host* h = new gg_csr_host();
  h->m = m;
  h->n = n;
  h->nnz = (int64_t)ents.size();
  h->val_dtype = dtype;
  h->rowptr = (int32_t*)malloc(sizeof(int32_t) * (m + 1));
  h->colval = (int32_t*)malloc(sizeof(int32_t) * std::max<int64_t>(1, h->nnz));
  h->nzval = malloc(dtype_size(dtype) * std::max<int64_t>(1, h->nnz));
  if (!h->rowptr || !h->colval || !h->nzval) {
    gg_csr_host_free(h);
    return set_error(GG_ERR_NOMEM, "host allocation failed");
  }
  std::vector<int64_t> cnt(m + 1, 0);
  for (const Entry& e : ents) ++cnt[e.row + 1];
  h->rowptr[0] = 1;
  for (int64_t r = 0; r < m; ++r) h->rowptr[r + 1] = (int32_t)(h->rowptr[r] + cnt[r + 1]);
  for (int64_t i = 0; i < h->nnz; ++i) {
    const Entry& e = ents[i];
    h->colval[i] = (int32_t)(e.col + 1);
    switch (dtype) {
      case GG_BOOL: ((uint8_t*)h->nzval)[i] = integer ? e.iv != 0 : e.dv != 0.0; break;
      case GG_INT32: ((int32_t*)h->nzval)[i] = (int32_t)(integer ? e.iv : (int64_t)e.dv); break;
      case GG_INT64: ((int64_t*)h->nzval)[i] = integer ? e.iv : (int64_t)e.dv; break;
      case GG_FLOAT32: ((float*)h->nzval)[i] = integer ? (float)e.iv : (float)e.dv; break;
      default: ((double*)h->nzval)[i] = integer ? (double)e.iv : e.dv; break;
    }
  }
  *out = h;
  return GG_OK;
}

// Integer values must fit the requested type; reals must be integral to be
// read as integers.
int check_values(const std::vector<Entry>& ents, int dtype, bool integer) {
  for (const Entry& e : ents) {
    if (dtype == GG_INT32 || dtype == GG_INT64) {
      if (!integer && e.dv != std::floor(e.dv))
        return set_error(GG_ERR_ARGUMENT, "non-integral value %g for an integer type", e.dv);
      const double v = integer ? (double)e.iv : e.dv;
      if (dtype == GG_INT32 && (v > 2147483647.0 || v < -2147483648.0))
        return set_error(GG_ERR_ARGUMENT, "value %g outside Int32", v);
    }
  }
  return GG_OK;
}

}  // namespace

extern "C" {

int gg_read_matrix_market(const char* path, int val_dtype, int transpose, gg_csr_host** out,
                          int* symmetric) {
  if (!path || !out) return set_error(GG_ERR_ARGUMENT, "path and out are required");
  clear_error();
  if (val_dtype != 0 && (val_dtype < GG_BOOL || val_dtype > GG_FLOAT64))
    return set_error(GG_ERR_ARGUMENT, "value type %d (Bool, Int32, Int64, Float32, Float64)",
                     val_dtype);
  std::string buf;
  GG_TRY(slurp(path, buf));
  Cursor c{buf.data(), buf.data() + buf.size()};
  // banner: %%MatrixMarket matrix coordinate <field> <symmetry>
  char obj[64] = {0}, fmt[64] = {0}, field[64] = {0}, sym[64] = {0};
  if (sscanf(c.p, "%%%%MatrixMarket %63s %63s %63s %63s", obj, fmt, field, sym) != 4)
    return set_error(GG_ERR_ARGUMENT, "%s: missing %%%%MatrixMarket banner", path);
  for (char* s : {obj, fmt, field, sym})
    for (char* q = s; *q; ++q) *q = (char)tolower(*q);
  if (strcmp(obj, "matrix") || strcmp(fmt, "coordinate"))
    return set_error(GG_ERR_UNSUPPORTED, "%s: only 'matrix coordinate' files are read", path);
  const bool pattern = !strcmp(field, "pattern"), integer = !strcmp(field, "integer");
  if (!pattern && !integer && strcmp(field, "real"))
    return set_error(GG_ERR_UNSUPPORTED, "%s: field '%s' (real, integer, pattern)", path, field);
  const bool general = !strcmp(sym, "general"), symm = !strcmp(sym, "symmetric"),
             skew = !strcmp(sym, "skew-symmetric");
  if (!general && !symm && !skew)
    return set_error(GG_ERR_UNSUPPORTED, "%s: symmetry '%s'", path, sym);
  const int dtype = val_dtype ? val_dtype : pattern ? GG_BOOL : integer ? GG_INT64 : GG_FLOAT64;
  if (skew && dtype == GG_BOOL)
    return set_error(GG_ERR_ARGUMENT, "%s: skew-symmetric values cannot be Bool", path);
  c.next_line();
  if (!c.next_data_line("%"))
    return set_error(GG_ERR_ARGUMENT, "%s: missing size line", path);
  int64_t m, n, nz;
  if (!c.read_i64(m) || !c.read_i64(n) || !c.read_i64(nz) || m < 0 || n < 0 || nz < 0)
    return set_error(GG_ERR_ARGUMENT, "%s:%lld: bad size line", path, (long long)c.line);
  if ((symm || skew) && m != n)
    return set_error(GG_ERR_ARGUMENT, "%s: symmetric file of a non-square matrix", path);
  if (m > 2147483647LL || n > 2147483647LL)
    return set_error(GG_ERR_ARGUMENT, "%s: dimensions exceed Int32 indices", path);
  c.next_line();
  std::vector<Entry> ents;
  ents.reserve((size_t)nz * ((symm || skew) ? 2 : 1));
  for (int64_t k = 0; k < nz; ++k) {
    if (!c.next_data_line("%"))
      return set_error(GG_ERR_ARGUMENT, "%s: %lld of %lld entries", path, (long long)k,
                       (long long)nz);
    Entry e{};
    if (!c.read_i64(e.row) || !c.read_i64(e.col))
      return set_error(GG_ERR_ARGUMENT, "%s:%lld: bad entry", path, (long long)c.line);
    if (e.row < 1 || e.row > m || e.col < 1 || e.col > n)
      return set_error(GG_ERR_BOUNDS, "%s:%lld: index (%lld, %lld) outside %lld x %lld", path,
                       (long long)c.line, (long long)e.row, (long long)e.col, (long long)m,
                       (long long)n);
    e.row -= 1;
    e.col -= 1;
    if (pattern) {
      e.dv = 1.0;
      e.iv = 1;
    } else if (integer) {
      if (!c.read_i64(e.iv))
        return set_error(GG_ERR_ARGUMENT, "%s:%lld: bad value", path, (long long)c.line);
      e.dv = (double)e.iv;
    } else if (!c.read_f64(e.dv)) {
      return set_error(GG_ERR_ARGUMENT, "%s:%lld: bad value", path, (long long)c.line);
    }
    if (!c.at_eol())
      return set_error(GG_ERR_ARGUMENT, "%s:%lld: extra fields", path, (long long)c.line);
    c.next_line();
    if (transpose) std::swap(e.row, e.col);
    ents.push_back(e);
    if ((symm || skew) && e.row != e.col) {
      Entry t = e;
      std::swap(t.row, t.col);
      if (skew) {
        t.dv = -t.dv;
        t.iv = -t.iv;
      }
      ents.push_back(t);
    }
  }
  GG_TRY(check_values(ents, dtype, integer && !pattern));
  const bool ints = integer || pattern;
  if (transpose) std::swap(m, n);
  // duplicates add up, as sparse(I, J, V) does; a pattern stays a pattern
  int st = build_csr(m, n, ents, dtype, ints,
                     [&](Entry& a, const Entry& b) {
                       if (pattern) return;
                       a.dv += b.dv;
                       a.iv += b.iv;
                     },
                     out);
  if (st == GG_OK && symmetric) *symmetric = (symm || skew) ? 1 : 0;
  return st;
}

int gg_read_edge_list(const char* path, int base, int directed, int val_dtype,
                      gg_csr_host** out) {
  if (!path || !out) return set_error(GG_ERR_ARGUMENT, "path and out are required");
  clear_error();
  if (base != 0 && base != 1) return set_error(GG_ERR_ARGUMENT, "base must be 0 or 1");
  if (val_dtype < GG_BOOL || val_dtype > GG_FLOAT64)
    return set_error(GG_ERR_ARGUMENT, "value type %d (Bool, Int32, Int64, Float32, Float64)",
                     val_dtype);
  std::string buf;
  GG_TRY(slurp(path, buf));
  Cursor c{buf.data(), buf.data() + buf.size()};
  std::vector<Entry> ents;
  int64_t nv = 0;
  int weighted = -1;  // decided by the first line
  const bool int_type = val_dtype == GG_INT32 || val_dtype == GG_INT64 || val_dtype == GG_BOOL;
  while (c.next_data_line("#%")) {
    int64_t u, v;
    if (!c.read_i64(u) || !c.read_i64(v))
      return set_error(GG_ERR_ARGUMENT, "%s:%lld: bad edge", path, (long long)c.line);
    if (u < base || v < base)
      return set_error(GG_ERR_BOUNDS, "%s:%lld: vertex below %d", path, (long long)c.line, base);
    u -= base;
    v -= base;
    if (u > 2147483646LL || v > 2147483646LL)
      return set_error(GG_ERR_ARGUMENT, "%s:%lld: vertex exceeds Int32", path, (long long)c.line);
    Entry e{};
    // row = head v, column = tail u: the CSR of A_T, row v lists in-neighbours
    e.row = v;
    e.col = u;
    e.dv = 1.0;
    e.iv = 1;
    const bool has_w = !c.at_eol();
    if (weighted < 0) weighted = has_w ? 1 : 0;
    if (has_w != (weighted == 1))
      return set_error(GG_ERR_ARGUMENT, "%s:%lld: weights on some lines only", path,
                       (long long)c.line);
    if (has_w) {
      if (int_type) {
        if (!c.read_i64(e.iv))
          return set_error(GG_ERR_ARGUMENT, "%s:%lld: bad weight", path, (long long)c.line);
        e.dv = (double)e.iv;
      } else if (!c.read_f64(e.dv)) {
        return set_error(GG_ERR_ARGUMENT, "%s:%lld: bad weight", path, (long long)c.line);
      }
      if (!c.at_eol())
        return set_error(GG_ERR_ARGUMENT, "%s:%lld: extra fields", path, (long long)c.line);
    }
    c.next_line();
    nv = std::max(nv, std::max(u, v) + 1);
    ents.push_back(e);
    if (!directed && u != v) {
      Entry t = e;
      std::swap(t.row, t.col);
      ents.push_back(t);
    }
  }
  GG_TRY(check_values(ents, val_dtype, int_type));
  // a repeated edge keeps its smallest weight (the one a shortest path uses)
  return build_csr(nv, nv, ents, val_dtype, int_type,
                   [&](Entry& a, const Entry& b) {
                     a.dv = std::min(a.dv, b.dv);
                     a.iv = std::min(a.iv, b.iv);
                   },
                   out);
}

}  // extern "C"
