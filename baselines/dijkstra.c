/* dijkstra.c -- single-thread Dijkstra: the reference's first CPU comparison, Galois
 * "-algo=dijkstra -t 1" (run_all.sh:93-107), restated as a CPU baseline (run_all.py's dj leg, the
 * benchmark's single-thread CPU leg).  The GPU library never links it.
 *
 * A 4-ary min-heap of (distance, vertex) with lazy deletion: an improved vertex is pushed again and the
 * stale entry is skipped when popped.  Int sums saturate at 2^32-2 (one below the unreached marker
 * 0xFFFFFFFF), float sums are IEEE single precision; the unreached marker is 0x7F800000 (+inf) for float.
 * Float distances are non-negative, so their bit patterns order like the values and the heap compares
 * keys as u32 in both variants.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint32_t key, v; } item;
typedef struct { item* a; size_t n, cap; } qheap;

static void up(qheap* h, item x) {
  if (h->n == h->cap) {
    h->cap = h->cap ? h->cap * 2 : 4096;
    h->a = (item*)realloc(h->a, h->cap * sizeof(item));
  }
  size_t i = h->n++;
  while (i) {
    size_t p = (i - 1) >> 2;
    if (h->a[p].key <= x.key) break;
    h->a[i] = h->a[p];
    i = p;
  }
  h->a[i] = x;
}

static item down(qheap* h) {
  item top = h->a[0];
  item x = h->a[--h->n];
  size_t i = 0;
  for (;;) {
    size_t c = 4 * i + 1;
    if (c >= h->n) break;
    size_t best = c;
    size_t end = c + 4 < h->n ? c + 4 : h->n;
    for (size_t j = c + 1; j < end; j++)
      if (h->a[j].key < h->a[best].key) best = j;
    if (h->a[best].key >= x.key) break;
    h->a[i] = h->a[best];
    i = best;
  }
  if (h->n) h->a[i] = x;
  return top;
}

static inline uint32_t relax(int is_float, uint32_t d, uint32_t w) {
  if (is_float) {
    float a, b;
    memcpy(&a, &d, 4);
    memcpy(&b, &w, 4);
    a += b;
    uint32_t r;
    memcpy(&r, &a, 4);
    return r;
  }
  uint64_t s = (uint64_t)d + w;
  return s >= 0xFFFFFFFFu ? 0xFFFFFFFEu : (uint32_t)s;
}

/* Distances from `src` into dist (n words).  *work (may be null) receives the number of vertices
 * settled, Galois' total_work for Dijkstra.  Returns 0, or 1 when src is out of range. */
int cpu_dijkstra(int is_float, uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                 uint32_t src, uint32_t* dist, uint64_t* work) {
  if (src >= n) return 1;
  const uint32_t inf = is_float ? 0x7F800000u : 0xFFFFFFFFu;
  for (uint32_t v = 0; v < n; v++) dist[v] = inf;
  qheap h = {0, 0, 0};
  uint64_t settled = 0;
  dist[src] = 0;
  up(&h, (item){0, src});
  while (h.n) {
    item t = down(&h);
    if (t.key != dist[t.v]) continue;
    settled++;
    for (uint32_t e = row[t.v]; e < row[t.v + 1]; e++) {
      uint32_t nd = relax(is_float, t.key, w[e]);
      uint32_t v = col[e];
      if (nd < dist[v]) {
        dist[v] = nd;
        up(&h, (item){nd, v});
      }
    }
  }
  free(h.a);
  if (work) *work = settled;
  return 0;
}
