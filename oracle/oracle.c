/* oracle.c -- CPU restatement of the ADDS path, used ONLY as a test/benchmark checker.
 *
 * This file is test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and never as the thing measured or shipped.  The product (adds_amd/) never links it.
 *
 * What it restates (reference file:line):
 *  - relaxation rule d[v] = min(d[v], d[u] + w) with a strict '>' test:
 *      ads_int/kernel.cu:112-121 (int32 add), ads_float/kernel.cu:120-130 (float add, min by dist,
 *      ads_float/csr_graph.h:23-49); unreached = INF (ads_int/csr_graph.h:23, INT32_MAX; float FLT_MAX).
 *  - the reference's correctness oracle is Galois Dijkstra run on one thread (run_all.sh:93-107,
 *      cpu_sssp_int.patch:133-159, cpu_sssp_float.patch:227-253): oracle_dijkstra_* is a textbook
 *      binary-heap Dijkstra with the same arithmetic.  Galois itself is an un-vendored dependency
 *      (IntelligentSoftwareSystems/Galois @ 38cd91cfb59a30cf0b4f7cf7d19f29d4d7188548, cloned by
 *      build_all.sh:45-65) and is not available here.
 *  - the tree gradient of the differentiable backward (no reference counterpart: ads_float/kernel.cu:82
 *      stores no predecessor), defined in oracle_tree_grad below.
 * The CPU comparison legs (Galois Dijkstra -t 1 and delta-step -t NUM_THREADS) live in baselines/, not
 * here: this file is the checker only.
 *
 * Distances are 32-bit words: uint32 (INF 0xFFFFFFFF) or IEEE binary32 bits (INF = +inf, 0x7F800000).
 * Non-negative floats order like their bit patterns, so one heap serves both.
 * Compile with -ffp-contract=off (see Makefile): every float add is one IEEE binary32 add.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define INF_U32 0xFFFFFFFFu
#define INF_F32 0x7F800000u

static float bits_f(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static uint32_t f_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

/* d + w in the variant's arithmetic.  Int: 64-bit sum, *ovf set when it reaches 2^32-1 (the reference's
 * int32 would already be undefined past 2^31-1). */
static uint32_t relax_add(int is_float, uint32_t d, uint32_t w, int* ovf) {
  if (is_float) return f_bits(bits_f(d) + bits_f(w));
  uint64_t s = (uint64_t)d + (uint64_t)w;
  if (s >= INF_U32) { *ovf = 1; return INF_U32 - 1; }
  return (uint32_t)s;
}

/* ---- binary min-heap of (key, vertex) with lazy deletion ---- */
typedef struct { uint32_t key, v; } hent;
typedef struct { hent* a; uint64_t n, cap; } heap;

static void hpush(heap* h, uint32_t key, uint32_t v) {
  if (h->n == h->cap) {
    h->cap = h->cap ? 2 * h->cap : 1024;
    h->a = (hent*)realloc(h->a, h->cap * sizeof(hent));
  }
  uint64_t i = h->n++;
  while (i > 0) {
    uint64_t p = (i - 1) / 2;
    if (h->a[p].key <= key) break;
    h->a[i] = h->a[p];
    i = p;
  }
  h->a[i].key = key;
  h->a[i].v = v;
}
static hent hpop(heap* h) {
  hent top = h->a[0], last = h->a[--h->n];
  uint64_t i = 0;
  for (;;) {
    uint64_t c = 2 * i + 1;
    if (c >= h->n) break;
    if (c + 1 < h->n && h->a[c + 1].key < h->a[c].key) c++;
    if (h->a[c].key >= last.key) break;
    h->a[i] = h->a[c];
    i = c;
  }
  if (h->n) h->a[i] = last;
  return top;
}

/* Dijkstra from k sources (all at distance 0) over one CSR.  Returns 1 if an int sum overflowed. */
int oracle_dijkstra(int is_float, uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                    const uint32_t* srcs, uint32_t k, uint32_t* dist) {
  const uint32_t inf = is_float ? INF_F32 : INF_U32;
  int ovf = 0;
  for (uint32_t v = 0; v < n; v++) dist[v] = inf;
  heap h = {0, 0, 0};
  for (uint32_t i = 0; i < k; i++) {
    dist[srcs[i]] = 0;
    hpush(&h, 0, srcs[i]);
  }
  while (h.n) {
    hent t = hpop(&h);
    if (t.key != dist[t.v]) continue; /* stale */
    for (uint32_t e = row[t.v]; e < row[t.v + 1]; e++) {
      uint32_t nd = relax_add(is_float, t.key, w[e], &ovf);
      if (nd < dist[col[e]]) {
        dist[col[e]] = nd;
        hpush(&h, nd, col[e]);
      }
    }
  }
  free(h.a);
  return ovf;
}

/* Chaotic Bellman-Ford: sweep all edges until nothing changes (brute force for tiny graphs). */
int oracle_bellman_ford(int is_float, uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                        const uint32_t* srcs, uint32_t k, uint32_t* dist) {
  const uint32_t inf = is_float ? INF_F32 : INF_U32;
  int ovf = 0, changed = 1;
  for (uint32_t v = 0; v < n; v++) dist[v] = inf;
  for (uint32_t i = 0; i < k; i++) dist[srcs[i]] = 0;
  while (changed) {
    changed = 0;
    for (uint32_t u = 0; u < n; u++) {
      if (dist[u] == inf) continue;
      for (uint32_t e = row[u]; e < row[u + 1]; e++) {
        uint32_t nd = relax_add(is_float, dist[u], w[e], &ovf);
        if (nd < dist[col[e]]) {
          dist[col[e]] = nd;
          changed = 1;
        }
      }
    }
  }
  return ovf;
}

/* Certificate that `dist` is the shortest-path solution from the k sources:
 *   0  ok
 *   1  a source does not have distance 0
 *   2  an edge violates dist[v] <= fl(dist[u] + w) (this includes a reached u with an INF neighbour)
 *   3  a vertex with finite distance is not reached from a source through tight edges
 *      (dist[v] == fl(dist[u] + w)); this also rejects finite distances on unreachable vertices.
 * With non-negative weights and a monotone add, (2) gives dist <= the true solution and (3) gives
 * dist >= it, so together they pin it exactly. */
int oracle_certify(int is_float, uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                   const uint32_t* srcs, uint32_t k, const uint32_t* dist) {
  const uint32_t inf = is_float ? INF_F32 : INF_U32;
  int ovf = 0;
  for (uint32_t i = 0; i < k; i++)
    if (dist[srcs[i]] != 0) return 1;
  for (uint32_t u = 0; u < n; u++) {
    if (dist[u] == inf) continue;
    for (uint32_t e = row[u]; e < row[u + 1]; e++)
      if (dist[col[e]] > relax_add(is_float, dist[u], w[e], &ovf)) return 2;
  }
  uint8_t* seen = (uint8_t*)calloc(n, 1);
  uint32_t* q = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint64_t qh = 0, qt = 0;
  for (uint32_t i = 0; i < k; i++)
    if (!seen[srcs[i]]) { seen[srcs[i]] = 1; q[qt++] = srcs[i]; }
  while (qh < qt) {
    uint32_t u = q[qh++];
    for (uint32_t e = row[u]; e < row[u + 1]; e++) {
      uint32_t v = col[e];
      if (!seen[v] && dist[v] == relax_add(is_float, dist[u], w[e], &ovf)) { seen[v] = 1; q[qt++] = v; }
    }
  }
  int rc = 0;
  for (uint32_t v = 0; v < n && !rc; v++)
    if (dist[v] != inf && !seen[v]) rc = 3;
  free(seen);
  free(q);
  return rc;
}

/* ---- backward reference ----
 * The reference has no backward pass (ads_float/kernel.cu:82 stores the vertex's own id as "parent"), so
 * this defines the one the device must reproduce bit for bit.  For each source s (dist_s = dist + s*n,
 * g_s = gdist + s*n), over the reached vertices (finite distance):
 *   tree: the source is the root.  Round 0: pe(v) = the smallest edge id e = (u -> v) with d[u] < d[v] and
 *     fl(d[u] + w_e) == d[v].  Round r >= 1: a vertex still without a parent takes the smallest e = (u -> v)
 *     with d[u] == d[v] and fl(d[u] + w_e) == d[v] whose tail was resolved (root, or parent chosen) in an
 *     earlier round -- zero-weight edges and adds that round away are tree edges too, and the rounds keep
 *     the parent pointers acyclic.  Rounds end when one resolves nothing.
 *   fixed point: M = max |g| and N = number of nonzero g over the reached vertex-instances of the batch;
 *     with M = m 2^eM (0.5 <= m < 1, frexpf) and 2^eN the least power of two >= N, the scale is
 *     S = 62 - eM - eN, so that N * M * 2^S < 2^62.  G = rint(g * 2^S) as a 64-bit integer (exact for g a
 *     multiple of 2^-S).
 *   gradient: GW[e] = sum over sources s and reached vertices t of G_s[t] * [e is on t's tree path to
 *     the root], an exact 64-bit integer sum (no order dependence); gw[e] = ldexpf((float)GW[e], -S):
 *     the exact sum of the scaled inputs, rounded once to float32 (round to nearest even).
 * Here the sums are accumulated children before parents (descending distance, later rounds first within
 * a distance), an order the device does not use. */
typedef struct { float d; uint32_t round, v; } ordent;
static int cmp_ord(const void* a, const void* b) {
  const ordent* x = (const ordent*)a;
  const ordent* y = (const ordent*)b;
  if (x->d != y->d) return x->d > y->d ? -1 : 1;
  if (x->round != y->round) return x->round > y->round ? -1 : 1;
  return (x->v > y->v) - (x->v < y->v);
}

int oracle_grad_scale(uint32_t n, uint32_t k, const float* dist, const float* gdist, uint64_t* nnz_out) {
  float mx = 0.0f;
  uint64_t nnz = 0;
  for (uint64_t i = 0; i < (uint64_t)k * n; i++) {
    if (f_bits(dist[i]) == INF_F32) continue;
    float a = gdist[i] < 0 ? -gdist[i] : gdist[i];
    if (a != 0.0f) nnz++;
    if (a > mx) mx = a;
  }
  if (nnz_out) *nnz_out = nnz;
  if (nnz == 0) return 0;
  int em = 0, en = 0;
  (void)frexpf(mx, &em);
  while (((uint64_t)1 << en) < nnz) en++;
  return 62 - em - en;
}

/* Returns 0, or -1 when a reached vertex of some source is not resolved (dist is not a shortest-path
 * solution from srcs) or a reached g is not finite. */
int oracle_tree_grad(uint32_t n, const uint32_t* row, const uint32_t* col, const float* w, uint32_t k,
                     const uint32_t* srcs, const float* dist, const float* gdist, float* gw, uint32_t* pe_out,
                     int* scale_out) {
  const uint64_t m = row[n];
  int rc = 0;
  uint32_t* rrow = (uint32_t*)calloc((size_t)n + 1, 4);
  uint32_t* reid = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
  uint32_t* rsrc = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
  for (uint64_t e = 0; e < m; e++) rrow[col[e] + 1]++;
  for (uint32_t v = 0; v < n; v++) rrow[v + 1] += rrow[v];
  uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
  memcpy(cur, rrow, sizeof(uint32_t) * n);
  for (uint32_t u = 0; u < n; u++)
    for (uint32_t e = row[u]; e < row[u + 1]; e++) {
      uint32_t p = cur[col[e]]++;
      reid[p] = e;
      rsrc[p] = u;
    }
  for (uint64_t i = 0; i < (uint64_t)k * n; i++)
    if (f_bits(dist[i]) != INF_F32 && !(gdist[i] - gdist[i] == 0.0f)) rc = -1; /* inf or nan */
  const int S = oracle_grad_scale(n, k, dist, gdist, NULL);
  if (scale_out) *scale_out = S;
  int64_t* GW = (int64_t*)calloc(m ? m : 1, sizeof(int64_t));
  uint32_t* pe = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n ? n : 1));
  uint32_t* rnd = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n ? n : 1));
  int64_t* A = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n ? n : 1));
  ordent* order = (ordent*)malloc(sizeof(ordent) * ((size_t)n ? n : 1));
  const uint32_t UNRES = 0xFFFFFFFFu, ROOT = 0xFFFFFFFEu;
  for (uint32_t s = 0; s < k && rc == 0; s++) {
    const float* d = dist + (uint64_t)s * n;
    const float* g = gdist + (uint64_t)s * n;
    for (uint32_t v = 0; v < n; v++) {
      pe[v] = UNRES;
      rnd[v] = UNRES;
      if (f_bits(d[v]) == INF_F32) continue;
      if (v == srcs[s]) { pe[v] = ROOT; rnd[v] = 0; continue; }
      for (uint32_t p = rrow[v]; p < rrow[v + 1]; p++) {
        float du = d[rsrc[p]];
        if (du < d[v] && du + w[reid[p]] == d[v]) { pe[v] = reid[p]; rnd[v] = 0; break; }
      }
    }
    for (uint32_t r = 1;; r++) {
      uint32_t changed = 0;
      for (uint32_t v = 0; v < n; v++) {
        if (f_bits(d[v]) == INF_F32 || pe[v] != UNRES) continue;
        for (uint32_t p = rrow[v]; p < rrow[v + 1]; p++) {
          uint32_t u = rsrc[p];
          if (d[u] == d[v] && d[u] + w[reid[p]] == d[v] && rnd[u] < r) { pe[v] = reid[p]; rnd[v] = r; changed++; break; }
        }
      }
      if (!changed) break;
    }
    if (pe_out)
      for (uint32_t v = 0; v < n; v++) pe_out[(uint64_t)s * n + v] = pe[v] >= ROOT ? 0xFFFFFFFFu : pe[v];
    uint32_t cnt = 0;
    for (uint32_t v = 0; v < n; v++) {
      if (f_bits(d[v]) == INF_F32) continue;
      if (pe[v] == UNRES) { rc = -1; break; }
      A[v] = (int64_t)rint(ldexp((double)g[v], S));
      order[cnt].d = d[v];
      order[cnt].round = rnd[v];
      order[cnt].v = v;
      cnt++;
    }
    if (rc) break;
    qsort(order, cnt, sizeof(ordent), cmp_ord);
    /* children come before parents in `order`: when v is reached, A[v] holds its whole subtree */
    for (uint32_t i = 0; i < cnt; i++) {
      uint32_t v = order[i].v;
      if (pe[v] == ROOT) continue;
      GW[pe[v]] += A[v];
      /* tail of the parent edge: the vertex whose row holds pe[v] (the last u with row[u] <= pe[v]
         whose row is not empty) */
      uint32_t lo = 0, hi = n;
      while (hi - lo > 1) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (row[mid] <= pe[v]) lo = mid; else hi = mid;
      }
      A[lo] += A[v];
    }
  }
  for (uint64_t e = 0; e < m; e++) gw[e] = 0.0f;
  if (rc == 0)
    for (uint64_t e = 0; e < m; e++) gw[e] = ldexpf((float)GW[e], -S);
  free(rrow); free(reid); free(rsrc); free(cur); free(pe); free(rnd); free(A); free(order); free(GW);
  return rc;
}
