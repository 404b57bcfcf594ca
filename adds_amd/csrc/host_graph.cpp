// Host-side graph handling for the ADDS solver: Galois .gr v1 I/O, per-vertex distance text, and the
// seeded synthetic generators (RMAT, grid + shortcuts, Erdos-Renyi) used for tests and benchmarks.
//
// The .gr layout follows CSRGraph::readFromGR (ads_int/csr_graph.cu:203-296): u64 version(=1),
// u64 sizeEdgeTy, u64 numNodes, u64 numEdges, u64 outIdx[numNodes] (end offsets), u32 outs[numEdges],
// one u32 pad when numEdges is odd, then numEdges 4-byte edge values.  The text format follows
// output() (ads_int/support.cu:78-96) for int and the float CPU writer (cpu_sssp_float.patch:241-248).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/adds.h"
#include "host_util.h"

namespace adds {

// ---------------------------------------------------------------------------------------------------
// Parallel helpers (std::thread; the library has no OpenMP dependency)
// ---------------------------------------------------------------------------------------------------
unsigned host_threads() {
  static unsigned t = [] {
    const char* e = getenv("ADDS_HOST_THREADS");
    unsigned v = e ? (unsigned)atoi(e) : std::thread::hardware_concurrency();
    if (v == 0) v = 1;
    if (v > 64) v = 64;
    return v;
  }();
  return t;
}

void parallel_for(uint64_t n, uint64_t grain, const std::function<void(uint64_t, uint64_t)>& fn) {
  unsigned nt = host_threads();
  if (n <= grain || nt == 1) {
    if (n) fn(0, n);
    return;
  }
  std::atomic<uint64_t> next{0};
  auto worker = [&] {
    for (;;) {
      uint64_t b = next.fetch_add(grain);
      if (b >= n) break;
      fn(b, std::min(n, b + grain));
    }
  };
  std::vector<std::thread> ts;
  for (unsigned i = 1; i < nt; i++) ts.emplace_back(worker);
  worker();
  for (auto& t : ts) t.join();
}

// ---------------------------------------------------------------------------------------------------
// Counter-based PRNG: splitmix64 finaliser over (key + (i+1)*golden).  Every random quantity is a pure
// function of (seed, stream, index), so generation is order-independent and parallel.
// ---------------------------------------------------------------------------------------------------
static inline uint64_t fmix64(uint64_t x) {
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
static inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return fmix64(seed * 0x9E3779B97F4A7C15ULL + stream * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL);
}
static inline uint64_t rnd(uint64_t key, uint64_t i) { return fmix64(key + (i + 1) * 0x9E3779B97F4A7C15ULL); }

enum : uint64_t { STREAM_EDGES = 1, STREAM_WEIGHTS = 2, STREAM_SHORTCUT = 3, STREAM_SAMPLE = 4 };

// Weight of the undirected pair {a, b}: a pure function of the unordered pair, so both directions
// carry the same value.
struct WeightFn {
  adds_wkind wk;
  double lo, hi;
  uint64_t key;
  uint32_t operator()(uint32_t a, uint32_t b) const {
    uint32_t x = std::min(a, b), y = std::max(a, b);
    uint64_t h = rnd(key, ((uint64_t)x << 32) | y);
    switch (wk) {
      case ADDS_WK_U32_UNIFORM: {
        uint64_t l = (uint64_t)lo, range = (uint64_t)hi - l + 1;
        return (uint32_t)(l + (((h >> 32) * range) >> 32));
      }
      case ADDS_WK_F32_UNIFORM: {
        double u = (double)(h >> 40) * (1.0 / 16777216.0);  // 24-bit uniform in [0,1), exact
        float f = (float)(lo + (hi - lo) * u);
        if ((double)f >= hi) f = std::nextafter((float)hi, -INFINITY);
        if ((double)f < lo) f = (float)lo;
        uint32_t bits;
        memcpy(&bits, &f, 4);
        return bits;
      }
      case ADDS_WK_F32_LOGNORMAL: {
        // Box-Muller: z = sqrt(-2 ln u1) cos(2 pi u2); w = exp(sigma z) with sigma = lo.
        double u1 = ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0);
        uint64_t h2 = fmix64(h ^ 0xA0761D6478BD642FULL);
        double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
        double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        float f = (float)std::exp(lo * z);
        uint32_t bits;
        memcpy(&bits, &f, 4);
        return bits;
      }
    }
    return 0;
  }
};

// Builds a symmetric CSR from undirected pairs: both directions, self loops dropped, duplicates
// removed, each row sorted by destination.
static adds_status build_symmetric_csr(uint32_t n, const std::vector<uint32_t>& src,
                                       const std::vector<uint32_t>& dst, const WeightFn& wf,
                                       adds_host_csr* out) {
  const uint64_t P = src.size();
  std::vector<uint32_t> deg((size_t)n + 1, 0);
  parallel_for(P, 1 << 16, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; i++) {
      uint32_t u = src[i], v = dst[i];
      if (u == v) continue;
      __atomic_fetch_add(&deg[u], 1u, __ATOMIC_RELAXED);
      __atomic_fetch_add(&deg[v], 1u, __ATOMIC_RELAXED);
    }
  });
  std::vector<uint64_t> off((size_t)n + 1, 0);
  for (uint32_t v = 0; v < n; v++) off[v + 1] = off[v] + deg[v];
  const uint64_t M2 = off[n];
  std::vector<uint32_t> tmp(M2);
  {
    std::vector<uint64_t> cursor(off.begin(), off.end() - 1);
    parallel_for(P, 1 << 16, [&](uint64_t b, uint64_t e) {
      for (uint64_t i = b; i < e; i++) {
        uint32_t u = src[i], v = dst[i];
        if (u == v) continue;
        tmp[__atomic_fetch_add(&cursor[u], 1ull, __ATOMIC_RELAXED)] = v;
        tmp[__atomic_fetch_add(&cursor[v], 1ull, __ATOMIC_RELAXED)] = u;
      }
    });
  }
  // sort + dedup each row in place; deg[v] becomes the deduplicated degree
  parallel_for(n, 4096, [&](uint64_t b, uint64_t e) {
    for (uint64_t v = b; v < e; v++) {
      uint32_t* r0 = tmp.data() + off[v];
      uint32_t* r1 = tmp.data() + off[v + 1];
      std::sort(r0, r1);
      deg[v] = (uint32_t)(std::unique(r0, r1) - r0);
    }
  });
  uint64_t m = 0;
  for (uint32_t v = 0; v < n; v++) m += deg[v];
  if (m > 0xFFFFFFFFull) return ADDS_EINVAL;
  out->n = n;
  out->m = m;
  out->row_ptr = (uint32_t*)malloc(((size_t)n + 1) * 4);
  out->col = (uint32_t*)malloc(std::max<uint64_t>(m, 1) * 4);
  out->w = (uint32_t*)malloc(std::max<uint64_t>(m, 1) * 4);
  out->wtype = (wf.wk == ADDS_WK_U32_UNIFORM) ? ADDS_W_U32 : ADDS_W_F32;
  if (!out->row_ptr || !out->col || !out->w) {
    adds_host_csr_free(out);
    return ADDS_ENOMEM;
  }
  out->row_ptr[0] = 0;
  for (uint32_t v = 0; v < n; v++) out->row_ptr[v + 1] = out->row_ptr[v] + deg[v];
  parallel_for(n, 4096, [&](uint64_t b, uint64_t e) {
    for (uint64_t v = b; v < e; v++) {
      const uint32_t* r = tmp.data() + off[v];
      uint32_t o = out->row_ptr[v];
      for (uint32_t j = 0; j < deg[v]; j++) {
        out->col[o + j] = r[j];
        out->w[o + j] = wf((uint32_t)v, r[j]);
      }
    }
  });
  return ADDS_OK;
}

}  // namespace adds

using namespace adds;

extern "C" {

void adds_host_csr_free(adds_host_csr* g) {
  if (!g) return;
  free(g->row_ptr);
  free(g->col);
  free(g->w);
  g->row_ptr = g->col = g->w = nullptr;
  g->n = 0;
  g->m = 0;
}

adds_status adds_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                          adds_wkind wk, double lo, double hi, adds_host_csr* out) {
  if (!out || scale == 0 || scale > 31 || edge_factor == 0) return ADDS_EINVAL;
  memset(out, 0, sizeof(*out));
  const uint32_t n = 1u << scale;
  const uint64_t P = (uint64_t)n * edge_factor;
  // Quadrant thresholds as 16-bit integers: u < ta -> (0,0), < tb -> (0,1), < tc -> (1,0), else (1,1).
  const uint32_t ta = (uint32_t)std::lround(a * 65536.0);
  const uint32_t tb = (uint32_t)std::lround((a + b) * 65536.0);
  const uint32_t tc = (uint32_t)std::lround((a + b + c) * 65536.0);
  std::vector<uint32_t> src(P), dst(P);
  const uint64_t key = stream_key(seed, STREAM_EDGES);
  parallel_for(P, 1 << 15, [&](uint64_t b0, uint64_t e0) {
    for (uint64_t i = b0; i < e0; i++) {
      uint32_t u = 0, v = 0;
      uint64_t h = 0;
      for (uint32_t l = 0; l < scale; l++) {
        if ((l & 3) == 0) h = rnd(key, i * 8 + (l >> 2));
        uint32_t r = (uint32_t)(h >> (16 * (l & 3))) & 0xFFFF;
        uint32_t bu = (r >= tb) ? 1u : 0u;
        uint32_t bv = (r >= ta && r < tb) || (r >= tc) ? 1u : 0u;
        u = (u << 1) | bu;
        v = (v << 1) | bv;
      }
      src[i] = u;
      dst[i] = v;
    }
  });
  WeightFn wf{wk, lo, hi, stream_key(seed, STREAM_WEIGHTS)};
  return build_symmetric_csr(n, src, dst, wf, out);
}

adds_status adds_gen_grid(uint32_t rows, uint32_t cols, uint64_t shortcuts, uint64_t seed, adds_wkind wk,
                          double lo, double hi, adds_host_csr* out) {
  if (!out || rows == 0 || cols == 0) return ADDS_EINVAL;
  memset(out, 0, sizeof(*out));
  uint64_t n64 = (uint64_t)rows * cols;
  if (n64 >= 0xFFFFFFFFull) return ADDS_EINVAL;
  const uint32_t n = (uint32_t)n64;
  const uint64_t grid_pairs = (uint64_t)rows * (cols - 1) + (uint64_t)(rows - 1) * cols;
  std::vector<uint32_t> src(grid_pairs + shortcuts), dst(grid_pairs + shortcuts);
  uint64_t k = 0;
  for (uint32_t r = 0; r < rows; r++)
    for (uint32_t c = 0; c + 1 < cols; c++) {
      src[k] = r * cols + c;
      dst[k++] = r * cols + c + 1;
    }
  for (uint32_t r = 0; r + 1 < rows; r++)
    for (uint32_t c = 0; c < cols; c++) {
      src[k] = r * cols + c;
      dst[k++] = (r + 1) * cols + c;
    }
  const uint64_t skey = stream_key(seed, STREAM_SHORTCUT);
  for (uint64_t i = 0; i < shortcuts; i++) {
    uint64_t h = rnd(skey, i);
    src[k] = (uint32_t)(((h & 0xFFFFFFFFull) * n) >> 32);
    dst[k++] = (uint32_t)(((h >> 32) * n) >> 32);
  }
  WeightFn wf{wk, lo, hi, stream_key(seed, STREAM_WEIGHTS)};
  return build_symmetric_csr(n, src, dst, wf, out);
}

adds_status adds_gen_er(uint32_t n, double avg_degree, uint64_t seed, adds_wkind wk, double lo, double hi,
                        adds_host_csr* out) {
  if (!out || n == 0) return ADDS_EINVAL;
  memset(out, 0, sizeof(*out));
  const uint64_t P = (uint64_t)std::llround((double)n * avg_degree / 2.0);
  std::vector<uint32_t> src(P), dst(P);
  const uint64_t key = stream_key(seed, STREAM_EDGES);
  parallel_for(P, 1 << 16, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; i++) {
      uint64_t h = rnd(key, i);
      src[i] = (uint32_t)(((h & 0xFFFFFFFFull) * n) >> 32);
      dst[i] = (uint32_t)(((h >> 32) * n) >> 32);
    }
  });
  WeightFn wf{wk, lo, hi, stream_key(seed, STREAM_WEIGHTS)};
  return build_symmetric_csr(n, src, dst, wf, out);
}

adds_status adds_gen_weights(adds_host_csr* g, uint64_t seed, adds_wkind wk, double lo, double hi) {
  if (!g || !g->row_ptr) return ADDS_EINVAL;
  WeightFn wf{wk, lo, hi, stream_key(seed, STREAM_WEIGHTS)};
  parallel_for(g->n, 4096, [&](uint64_t b, uint64_t e) {
    for (uint64_t v = b; v < e; v++)
      for (uint32_t j = g->row_ptr[v]; j < g->row_ptr[v + 1]; j++) g->w[j] = wf((uint32_t)v, g->col[j]);
  });
  g->wtype = (wk == ADDS_WK_U32_UNIFORM) ? ADDS_W_U32 : ADDS_W_F32;
  return ADDS_OK;
}

uint32_t adds_max_degree_vertex(const adds_host_csr* g) {
  uint32_t best = 0, bd = 0;
  for (uint32_t v = 0; v < g->n; v++) {
    uint32_t d = g->row_ptr[v + 1] - g->row_ptr[v];
    if (d > bd) {
      bd = d;
      best = v;
    }
  }
  return best;
}

uint32_t adds_sample_component(const adds_host_csr* g, uint32_t root, uint32_t k, uint64_t seed, uint32_t* out) {
  if (root >= g->n) return 0;
  std::vector<uint8_t> seen(g->n, 0);
  std::vector<uint32_t> comp;
  comp.push_back(root);
  seen[root] = 1;
  for (size_t i = 0; i < comp.size(); i++) {
    uint32_t u = comp[i];
    for (uint32_t e = g->row_ptr[u]; e < g->row_ptr[u + 1]; e++) {
      uint32_t v = g->col[e];
      if (!seen[v]) {
        seen[v] = 1;
        comp.push_back(v);
      }
    }
  }
  std::sort(comp.begin(), comp.end());
  // partial Fisher-Yates over the sorted component with the counter PRNG
  const uint64_t key = stream_key(seed, STREAM_SAMPLE);
  uint32_t take = (uint32_t)std::min<uint64_t>(k, comp.size());
  for (uint32_t i = 0; i < take; i++) {
    uint64_t left = comp.size() - i;
    uint64_t j = i + ((rnd(key, i) >> 32) * left >> 32);
    std::swap(comp[i], comp[j]);
    out[i] = comp[i];
  }
  return take;
}

void adds_reached_counts(const adds_host_csr* g, const uint32_t* dist, uint64_t* n_r, uint64_t* m_r) {
  const uint32_t inf = (g->wtype == ADDS_W_F32) ? ADDS_INF_F32_BITS : ADDS_INF_U32;
  std::atomic<uint64_t> an{0}, am{0};
  parallel_for(g->n, 1 << 16, [&](uint64_t b, uint64_t e) {
    uint64_t ln = 0, lm = 0;
    for (uint64_t v = b; v < e; v++)
      if (dist[v] != inf) {
        ln++;
        lm += g->row_ptr[v + 1] - g->row_ptr[v];
      }
    an += ln;
    am += lm;
  });
  *n_r = an;
  *m_r = am;
}

adds_status adds_read_gr(const char* path, adds_wtype wt, adds_host_csr* out) {
  if (!path || !out) return ADDS_EINVAL;
  memset(out, 0, sizeof(*out));
  int fd = open(path, O_RDONLY);
  if (fd < 0) return ADDS_EIO;
  struct stat st;
  if (fstat(fd, &st) != 0) {
    close(fd);
    return ADDS_EIO;
  }
  const size_t len = (size_t)st.st_size;
  if (len < 32) {
    close(fd);
    return ADDS_EINVAL;
  }
  void* mp = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (mp == MAP_FAILED) return ADDS_EIO;
  const uint64_t* h = (const uint64_t*)mp;
  adds_status rc = ADDS_OK;
  const uint64_t version = h[0], esz = h[1], nn = h[2], ne = h[3];
  const uint64_t need = 32 + nn * 8 + ne * 4 + ((ne & 1) ? 4 : 0) + (esz ? ne * 4 : 0);
  if (version != 1 || (esz != 0 && esz != 4) || nn >= 0xFFFFFFFFull || ne > 0xFFFFFFFFull || len < need) {
    munmap(mp, len);
    return ADDS_EINVAL;
  }
  const uint64_t* outIdx = h + 4;
  const uint32_t* outs = (const uint32_t*)(outIdx + nn);
  const uint32_t* edata = outs + ne + ((ne & 1) ? 1 : 0);
  out->n = (uint32_t)nn;
  out->m = ne;
  out->wtype = wt;
  out->row_ptr = (uint32_t*)malloc((nn + 1) * 4);
  out->col = (uint32_t*)malloc(std::max<uint64_t>(ne, 1) * 4);
  out->w = (uint32_t*)malloc(std::max<uint64_t>(ne, 1) * 4);
  if (!out->row_ptr || !out->col || !out->w) {
    munmap(mp, len);
    adds_host_csr_free(out);
    return ADDS_ENOMEM;
  }
  out->row_ptr[0] = 0;
  uint64_t prev = 0;
  for (uint64_t i = 0; i < nn; i++) {
    uint64_t o = outIdx[i];
    if (o < prev || o > ne) rc = ADDS_EINVAL;
    out->row_ptr[i + 1] = (uint32_t)o;
    prev = o;
  }
  if (rc == ADDS_OK && prev != ne) rc = ADDS_EINVAL;
  if (rc == ADDS_OK) {
    std::atomic<int> bad{0};
    parallel_for(ne, 1 << 16, [&](uint64_t b, uint64_t e) {
      for (uint64_t i = b; i < e; i++) {
        uint32_t d = outs[i];
        // The reference prints a warning for dst >= n and keeps the edge (csr_graph.cu:269-271); a device
        // kernel would then index out of bounds, so such a file is rejected here.
        if (d >= nn) bad = 1;
        out->col[i] = d;
        out->w[i] = esz ? edata[i] : 0u;
      }
    });
    if (bad) rc = ADDS_EINVAL;
  }
  munmap(mp, len);
  if (rc != ADDS_OK) adds_host_csr_free(out);
  return rc;
}

adds_status adds_write_gr(const char* path, const adds_host_csr* g) {
  if (!path || !g) return ADDS_EINVAL;
  FILE* f = fopen(path, "wb");
  if (!f) return ADDS_EIO;
  uint64_t hdr[4] = {1, 4, g->n, g->m};
  bool ok = fwrite(hdr, 8, 4, f) == 4;
  std::vector<uint64_t> idx(g->n);
  for (uint32_t i = 0; i < g->n; i++) idx[i] = g->row_ptr[i + 1];
  ok = ok && fwrite(idx.data(), 8, g->n, f) == g->n;
  ok = ok && fwrite(g->col, 4, g->m, f) == g->m;
  if (g->m & 1) {
    uint32_t pad = 0;
    ok = ok && fwrite(&pad, 4, 1, f) == 1;
  }
  ok = ok && fwrite(g->w, 4, g->m, f) == g->m;
  ok = (fclose(f) == 0) && ok;
  return ok ? ADDS_OK : ADDS_EIO;
}

adds_status adds_write_dist_text(const char* path, const uint32_t* dist, uint32_t n, adds_wtype wt) {
  if (!path || (!dist && n)) return ADDS_EINVAL;
  FILE* f = strcmp(path, "-") == 0 ? stdout : fopen(path, "w");
  if (!f) return ADDS_EIO;
  std::vector<char> buf(1 << 20);
  size_t used = 0;
  bool ok = true;
  for (uint32_t i = 0; i < n; i++) {
    if (used + 96 > buf.size()) {
      ok = ok && fwrite(buf.data(), 1, used, f) == used;
      used = 0;
    }
    int len;
    if (wt == ADDS_W_F32) {
      if (dist[i] == ADDS_INF_F32_BITS)
        len = snprintf(buf.data() + used, 96, "%d INF\n", (int)i);
      else {
        float x;
        memcpy(&x, &dist[i], 4);
        len = snprintf(buf.data() + used, 96, "%d %f\n", (int)i, (double)x);
      }
    } else {
      if (dist[i] == ADDS_INF_U32)
        len = snprintf(buf.data() + used, 96, "%d INF\n", (int)i);
      else
        len = snprintf(buf.data() + used, 96, "%d %u\n", (int)i, dist[i]);
    }
    used += (size_t)len;
  }
  ok = ok && fwrite(buf.data(), 1, used, f) == used;
  if (f != stdout) ok = (fclose(f) == 0) && ok;
  else fflush(f);
  return ok ? ADDS_OK : ADDS_EIO;
}

}  // extern "C"
