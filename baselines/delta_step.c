/* delta_step.c -- parallel CPU delta-stepping SSSP: the reference's second CPU comparison,
 * Galois "-algo=deltaStep -t <threads>" (run_all.sh:77-91, NUM_THREADS = hardware threads), restated.
 * A CPU baseline: run_all.py's ds leg and the benchmark's parallel CPU leg time it; the GPU library
 * never links it.
 *
 * Galois' delta semantics (cpu_graph_list_* column 3): int bucket = dist >> shift
 * (BFS_SSSP.h UpdateRequestIndexer), float bucket = (unsigned)(dist / div) (cpu_sssp_float.patch:265-279).
 * Here both are `bucket = (unsigned)(dist / delta)` with delta a power of two for int.
 *
 * Bucket-synchronous: all threads drain bucket b in rounds (the round's frontier is split evenly,
 * improvements go to thread-local bucket lists); b advances to the smallest non-empty bucket.  Entries
 * carry the distance they were pushed with and are dropped when the vertex already holds a smaller one.
 * Distances use atomic min (compare-and-swap on the 32-bit word), so the result is the unique shortest
 * path solution whatever the schedule.
 */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define INF_U32 0xFFFFFFFFu
#define INF_F32 0x7F800000u
#define NSLOT 4096u /* bucket ring per thread; bucket b lives in slot b % NSLOT */

typedef struct { uint32_t v, d; } ent;
typedef struct { ent* a; uint64_t n, cap; } vec;

static void vpush(vec* x, uint32_t v, uint32_t d) {
  if (x->n == x->cap) {
    x->cap = x->cap ? 2 * x->cap : 256;
    x->a = (ent*)realloc(x->a, x->cap * sizeof(ent));
  }
  x->a[x->n].v = v;
  x->a[x->n].d = d;
  x->n++;
}

static float bf(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
/* int sums saturate at 2^32-2, one below the unreached marker, as the GPU solver's do */
static uint32_t sat_add(uint32_t d, uint32_t w) {
  uint64_t s = (uint64_t)d + w;
  return s >= INF_U32 ? INF_U32 - 1 : (uint32_t)s;
}
static uint32_t fb(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

typedef struct shared_s {
  int is_float;
  uint32_t n;
  const uint32_t *row, *col, *w;
  uint32_t* dist;
  double inv_delta;
  unsigned nthreads;
  vec* slots;           /* [nthreads][NSLOT] */
  ent* frontier;        /* merged frontier of the current round */
  uint64_t fsize, fcap;
  uint64_t cur;         /* current bucket */
  int done;
  pthread_barrier_t bar;
} shared_t;

static uint64_t bkey(const shared_t* S, uint32_t d) {
  double x = S->is_float ? (double)bf(d) : (double)d;
  return (uint64_t)(x * S->inv_delta);
}

static void relax_min(shared_t* S, vec* mine, uint32_t v, uint32_t nd) {
  uint32_t old = __atomic_load_n(&S->dist[v], __ATOMIC_RELAXED);
  while (nd < old) {
    if (__atomic_compare_exchange_n(&S->dist[v], &old, nd, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
      vpush(&mine[bkey(S, nd) % NSLOT], v, nd);
      return;
    }
  }
}

typedef struct { shared_t* S; unsigned id; uint64_t work; } targ;

/* Thread 0 merges the slot lists of bucket `cur` into the frontier (entries of other buckets that share
 * the slot stay), picks the next bucket when empty, and every thread relaxes its share. */
static void* worker(void* p) {
  targ* t = (targ*)p;
  shared_t* S = t->S;
  vec* mine = S->slots + (uint64_t)t->id * NSLOT;
  for (;;) {
    pthread_barrier_wait(&S->bar);
    if (t->id == 0) {
      S->fsize = 0;
      for (int pass = 0; pass < 2 && S->fsize == 0; pass++) {
        if (pass == 1) { /* bucket drained: move to the smallest non-empty bucket */
          uint64_t best = UINT64_MAX;
          for (unsigned th = 0; th < S->nthreads; th++)
            for (unsigned s = 0; s < NSLOT; s++) {
              vec* x = &S->slots[(uint64_t)th * NSLOT + s];
              for (uint64_t i = 0; i < x->n; i++) {
                uint64_t k = bkey(S, x->a[i].d);
                if (k < best) best = k;
              }
            }
          if (best == UINT64_MAX) { S->done = 1; break; }
          S->cur = best;
        }
        unsigned slot = (unsigned)(S->cur % NSLOT);
        for (unsigned th = 0; th < S->nthreads; th++) {
          vec* x = &S->slots[(uint64_t)th * NSLOT + slot];
          uint64_t keep = 0;
          for (uint64_t i = 0; i < x->n; i++) {
            ent e = x->a[i];
            if (bkey(S, e.d) == S->cur) {
              if (S->fsize == S->fcap) {
                S->fcap = S->fcap ? 2 * S->fcap : 4096;
                S->frontier = (ent*)realloc(S->frontier, S->fcap * sizeof(ent));
              }
              S->frontier[S->fsize++] = e;
            } else {
              x->a[keep++] = e;
            }
          }
          x->n = keep;
        }
      }
    }
    pthread_barrier_wait(&S->bar);
    if (S->done) break;
    uint64_t lo = S->fsize * t->id / S->nthreads, hi = S->fsize * (t->id + 1) / S->nthreads;
    for (uint64_t i = lo; i < hi; i++) {
      uint32_t u = S->frontier[i].v, du = S->frontier[i].d;
      if (__atomic_load_n(&S->dist[u], __ATOMIC_RELAXED) < du) continue; /* stale */
      t->work++; /* Galois WORK_COUNT (cpu_sssp_int.patch:144): requests processed, stale ones excluded */
      for (uint32_t e = S->row[u]; e < S->row[u + 1]; e++) {
        uint32_t nd = S->is_float ? fb(bf(du) + bf(S->w[e])) : sat_add(du, S->w[e]);
        relax_min(S, mine, S->col[e], nd);
      }
    }
  }
  return NULL;
}

/* Returns 0 on success.  delta > 0 (distance units).  *work (may be null) receives the number of
 * non-stale requests processed, the count Galois reports as total_work. */
int cpu_delta_step(int is_float, uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                      uint32_t src, double delta, unsigned nthreads, uint32_t* dist, uint64_t* work) {
  if (nthreads < 1) nthreads = 1;
  shared_t S;
  memset(&S, 0, sizeof(S));
  S.is_float = is_float;
  S.n = n;
  S.row = row;
  S.col = col;
  S.w = w;
  S.dist = dist;
  S.inv_delta = 1.0 / delta;
  S.nthreads = nthreads;
  S.slots = (vec*)calloc((uint64_t)nthreads * NSLOT, sizeof(vec));
  for (uint32_t v = 0; v < n; v++) dist[v] = is_float ? INF_F32 : INF_U32;
  dist[src] = 0;
  vpush(&S.slots[0], src, 0);
  S.cur = 0;
  pthread_barrier_init(&S.bar, NULL, nthreads);
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
  targ* ta = (targ*)malloc(sizeof(targ) * nthreads);
  for (unsigned i = 0; i < nthreads; i++) {
    ta[i].S = &S;
    ta[i].id = i;
    ta[i].work = 0;
    if (i) pthread_create(&th[i], NULL, worker, &ta[i]);
  }
  worker(&ta[0]);
  for (unsigned i = 1; i < nthreads; i++) pthread_join(th[i], NULL);
  if (work) {
    *work = 0;
    for (unsigned i = 0; i < nthreads; i++) *work += ta[i].work;
  }
  pthread_barrier_destroy(&S.bar);
  for (uint64_t i = 0; i < (uint64_t)nthreads * NSLOT; i++) free(S.slots[i].a);
  free(S.slots);
  free(S.frontier);
  free(th);
  free(ta);
  return 0;
}
