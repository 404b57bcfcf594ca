// Text graph formats -> host CSR, for running the solver on the public graphs the reference evaluated
// (its graph lists name SuiteSparse matrices and DIMACS graphs, converted to Galois .gr outside the
// reference): MatrixMarket coordinate files and DIMACS shortest-path files.  The weight conventions
// are stated at the declarations in include/adds.h.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/adds.h"
#include "host_util.h"

namespace {

struct EdgeList {
  uint32_t n = 0;
  std::vector<uint32_t> src, dst, w;
};

// Stable counting sort by source: edges keep their file order within a row.  Arrays are malloc'd so
// that adds_host_csr_free releases them.
adds_status to_csr(const EdgeList& el, adds_wtype wt, adds_host_csr* out) {
  const uint64_t m = el.src.size();
  if (m > 0xFFFFFFFFull) {
    adds::set_error_detail("more than 2^32-1 edges");
    return ADDS_EOVERFLOW;
  }
  uint32_t* row = (uint32_t*)calloc((size_t)el.n + 1, sizeof(uint32_t));
  uint32_t* col = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
  uint32_t* w = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
  if (!row || !col || !w) {
    free(row);
    free(col);
    free(w);
    return ADDS_ENOMEM;
  }
  for (uint64_t e = 0; e < m; e++) row[el.src[e] + 1]++;
  for (uint32_t v = 0; v < el.n; v++) row[v + 1] += row[v];
  std::vector<uint32_t> next(row, row + el.n);
  for (uint64_t e = 0; e < m; e++) {
    const uint32_t p = next[el.src[e]]++;
    col[p] = el.dst[e];
    w[p] = el.w[e];
  }
  out->n = el.n;
  out->m = m;
  out->row_ptr = row;
  out->col = col;
  out->w = w;
  out->wtype = wt;
  return ADDS_OK;
}

// |v| as a weight word of type wt; false when it cannot be one (NaN, or above 2^31-1 for integers).
bool weight_word(double v, adds_wtype wt, uint32_t* out) {
  if (std::isnan(v)) return false;
  const double a = std::fabs(v);
  if (wt == ADDS_W_F32) {
    const float f = (float)a;  // round to nearest
    if (std::isinf(f)) return false;
    memcpy(out, &f, 4);
    return true;
  }
  const double r = std::nearbyint(a);
  if (r > 2147483647.0) return false;
  *out = (uint32_t)r;
  return true;
}

struct LineReader {
  FILE* f = nullptr;
  char* buf = nullptr;
  size_t cap = 0;
  uint64_t lineno = 0;
  ~LineReader() {
    free(buf);
    if (f) fclose(f);
  }
  const char* next() {
    const ssize_t r = getline(&buf, &cap, f);
    if (r < 0) return nullptr;
    lineno++;
    return buf;
  }
};

adds_status fail(const LineReader& lr, const char* path, const char* what, adds_status s = ADDS_EIO) {
  adds::set_error_detail(std::string(path) + ":" + std::to_string(lr.lineno) + ": " + what);
  return s;
}

bool blank(const char* p) {
  while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') p++;
  return *p == 0;
}

adds_status read_mtx(const char* path, adds_wtype wt, adds_host_csr* out) {
  if (!path || !out || (wt != ADDS_W_U32 && wt != ADDS_W_F32)) return ADDS_EINVAL;
  memset(out, 0, sizeof(*out));
  LineReader lr;
  lr.f = fopen(path, "r");
  if (!lr.f) {
    adds::set_error_detail(std::string(path) + ": " + strerror(errno));
    return ADDS_EIO;
  }
  const char* line = lr.next();
  char obj[64], fmt[64], field[64], sym[64];
  if (!line || sscanf(line, "%%%%MatrixMarket %63s %63s %63s %63s", obj, fmt, field, sym) != 4)
    return fail(lr, path, "missing %%MatrixMarket header");
  for (char* s : {obj, fmt, field, sym})
    for (char* c = s; *c; c++) *c = (char)tolower((unsigned char)*c);
  if (strcmp(obj, "matrix") || strcmp(fmt, "coordinate"))
    return fail(lr, path, "only 'matrix coordinate' files are graphs");
  const bool pattern = !strcmp(field, "pattern");
  if (!pattern && strcmp(field, "real") && strcmp(field, "integer") && strcmp(field, "double"))
    return fail(lr, path, "field must be real, integer or pattern");
  const bool mirror = strcmp(sym, "general") != 0;  // symmetric, skew-symmetric, hermitian
  if (mirror && strcmp(sym, "symmetric") && strcmp(sym, "skew-symmetric") && strcmp(sym, "hermitian"))
    return fail(lr, path, "unknown symmetry");
  unsigned long long rows = 0, cols = 0, nnz = 0;
  while ((line = lr.next())) {
    if (line[0] == '%' || blank(line)) continue;
    if (sscanf(line, "%llu %llu %llu", &rows, &cols, &nnz) != 3) return fail(lr, path, "bad size line");
    break;
  }
  if (!line) return fail(lr, path, "missing size line");
  const unsigned long long n = rows > cols ? rows : cols;
  if (n > 0xFFFFFFFEull) return fail(lr, path, "more than 2^32-2 vertices", ADDS_EOVERFLOW);
  EdgeList el;
  el.n = (uint32_t)n;
  el.src.reserve(mirror ? 2 * nnz : nnz);
  el.dst.reserve(mirror ? 2 * nnz : nnz);
  el.w.reserve(mirror ? 2 * nnz : nnz);
  unsigned long long seen = 0;
  while ((line = lr.next())) {
    if (line[0] == '%' || blank(line)) continue;
    char* p = (char*)line;
    char* e;
    const unsigned long long i = strtoull(p, &e, 10);
    if (e == p) return fail(lr, path, "bad entry");
    p = e;
    const unsigned long long j = strtoull(p, &e, 10);
    if (e == p) return fail(lr, path, "bad entry");
    p = e;
    double v = 1.0;
    if (!pattern) {
      v = strtod(p, &e);  // a complex entry's real part
      if (e == p) return fail(lr, path, "missing value");
    }
    if (i < 1 || j < 1 || i > rows || j > cols) return fail(lr, path, "index out of range");
    uint32_t ww;
    if (!weight_word(v, wt, &ww)) return fail(lr, path, "weight not representable", ADDS_EOVERFLOW);
    el.src.push_back((uint32_t)(i - 1));
    el.dst.push_back((uint32_t)(j - 1));
    el.w.push_back(ww);
    if (mirror && i != j) {
      el.src.push_back((uint32_t)(j - 1));
      el.dst.push_back((uint32_t)(i - 1));
      el.w.push_back(ww);
    }
    seen++;
  }
  if (seen != nnz) return fail(lr, path, "entry count differs from the size line");
  return to_csr(el, wt, out);
}

adds_status read_dimacs(const char* path, adds_wtype wt, adds_host_csr* out) {
  if (!path || !out || (wt != ADDS_W_U32 && wt != ADDS_W_F32)) return ADDS_EINVAL;
  memset(out, 0, sizeof(*out));
  LineReader lr;
  lr.f = fopen(path, "r");
  if (!lr.f) {
    adds::set_error_detail(std::string(path) + ": " + strerror(errno));
    return ADDS_EIO;
  }
  EdgeList el;
  bool have_p = false;
  unsigned long long n = 0, m = 0;
  const char* line;
  while ((line = lr.next())) {
    if (line[0] == 'c' || blank(line)) continue;
    if (line[0] == 'p') {
      char kind[16];
      if (have_p || sscanf(line, "p %15s %llu %llu", kind, &n, &m) != 3 || strcmp(kind, "sp"))
        return fail(lr, path, "bad problem line (want 'p sp <n> <m>')");
      if (n > 0xFFFFFFFEull) return fail(lr, path, "more than 2^32-2 vertices", ADDS_EOVERFLOW);
      have_p = true;
      el.n = (uint32_t)n;
      el.src.reserve(m);
      el.dst.reserve(m);
      el.w.reserve(m);
      continue;
    }
    if (line[0] != 'a') return fail(lr, path, "unknown line type");
    if (!have_p) return fail(lr, path, "arc before the problem line");
    unsigned long long u, v;
    double x;
    if (sscanf(line, "a %llu %llu %lf", &u, &v, &x) != 3) return fail(lr, path, "bad arc line");
    if (u < 1 || v < 1 || u > n || v > n) return fail(lr, path, "vertex out of range");
    uint32_t ww;
    if (!weight_word(x, wt, &ww)) return fail(lr, path, "weight not representable", ADDS_EOVERFLOW);
    el.src.push_back((uint32_t)(u - 1));
    el.dst.push_back((uint32_t)(v - 1));
    el.w.push_back(ww);
  }
  if (!have_p) return fail(lr, path, "missing problem line");
  if (el.src.size() != m) return fail(lr, path, "arc count differs from the problem line");
  return to_csr(el, wt, out);
}

// The C entry points: no exception leaves the library (a size line can ask for any amount of memory).
template <class F>
adds_status guarded(F&& f, adds_host_csr* out) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    if (out) adds_host_csr_free(out);
    adds::set_error_detail("out of host memory");
    return ADDS_ENOMEM;
  } catch (const std::length_error&) {
    if (out) adds_host_csr_free(out);
    adds::set_error_detail("edge count too large");
    return ADDS_ENOMEM;
  }
}

}  // namespace

extern "C" {

adds_status adds_read_mtx(const char* path, adds_wtype wt, adds_host_csr* out) {
  return guarded([&] { return read_mtx(path, wt, out); }, out);
}

adds_status adds_read_dimacs(const char* path, adds_wtype wt, adds_host_csr* out) {
  return guarded([&] { return read_dimacs(path, wt, out); }, out);
}

}  // extern "C"
