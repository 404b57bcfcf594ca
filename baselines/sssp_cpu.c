/* sssp_cpu: the command line that run_all.sh gives the reference's Galois sssp-cpu (run_dj / run_ds,
 * run_all.sh:80-110), over this directory's CPU baselines.
 *
 *   sssp_cpu (int|float) <graph.gr> [-noverify] -algo=(dijkstra|deltaStep) -startNode=<s>
 *            -delta=<x> -t <threads>
 *
 * Like the patched Galois binary it writes {dj|ds}_{int|float}_final_dist/<graph basename> in the
 * working directory ("%d %u" / "%d %f" / "%d INF" lines, cpu_sssp_int.patch:129-153,
 * cpu_sssp_float.patch:225-248) and appends "<basename> <ms> <work>" to fd 3 (cpu_sssp_int.patch:155-158).
 * ms is the algorithm alone (steady clock, like Galois execTime); work is Galois' total_work: requests
 * processed, stale ones excluded (the WORK_COUNT / total_work counters the patches add), which for
 * Dijkstra is one per reached vertex.
 *
 * -delta follows Galois: for int graphs -delta=k is the shift of the bucket index dist >> k (bucket width
 * 2^k, BFS_SSSP.h UpdateRequestIndexer); for float graphs -delta=x is the divisor of (unsigned)(dist / x)
 * (cpu_sssp_float.patch:13, 265-279).  The graph lists' third column holds these values
 * (cpu_graph_list_int, cpu_graph_list_float).  Dijkstra runs on one thread (run_all.sh passes -t 1). */
#define _POSIX_C_SOURCE 200809L
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

int cpu_read_gr(const char*, uint32_t*, uint64_t*, uint32_t**, uint32_t**, uint32_t**);
int cpu_dijkstra(int, uint32_t, const uint32_t*, const uint32_t*, const uint32_t*, uint32_t, uint32_t*, uint64_t*);
int cpu_delta_step(int, uint32_t, const uint32_t*, const uint32_t*, const uint32_t*, uint32_t, double, unsigned,
                   uint32_t*, uint64_t*);
int cpu_write_dist(const char*, int, uint32_t, const uint32_t*);

static int usage(const char* a0) {
  fprintf(stderr,
          "usage: %s (int|float) <graph.gr> [-noverify] -algo=(dijkstra|deltaStep) -startNode=<s> -delta=<x> "
          "-t <threads>\n",
          a0);
  return 2;
}

/* value of "-name=v" or "-name v" at argv[*i] (advancing *i past a separate value); NULL if not -name */
static const char* opt(int argc, char** argv, int* i, const char* name) {
  const char* a = argv[*i];
  if (a[0] != '-') return NULL;
  a += (a[1] == '-') ? 2 : 1;
  size_t l = strlen(name);
  if (strncmp(a, name, l)) return NULL;
  if (a[l] == '=') return a + l + 1;
  if (a[l] == 0 && *i + 1 < argc) return argv[++*i];
  return NULL;
}

int main(int argc, char** argv) {
  if (argc < 3 || (strcmp(argv[1], "int") && strcmp(argv[1], "float"))) return usage(argv[0]);
  const int is_float = !strcmp(argv[1], "float");
  const char* path = NULL;
  const char* algo = "dijkstra";
  uint32_t src = 0;
  double delta_arg = is_float ? 1.0 : 13.0; /* Galois defaults: divisor 1 / shift 13 */
  unsigned threads = 1;
  for (int i = 2; i < argc; i++) {
    const char* v;
    if (!strcmp(argv[i], "-noverify")) continue;
    if ((v = opt(argc, argv, &i, "algo"))) algo = v;
    else if ((v = opt(argc, argv, &i, "startNode"))) src = (uint32_t)strtoul(v, 0, 10);
    else if ((v = opt(argc, argv, &i, "delta"))) delta_arg = strtod(v, 0);
    else if ((v = opt(argc, argv, &i, "t"))) threads = (unsigned)strtoul(v, 0, 10);
    else if (argv[i][0] != '-' && !path) path = argv[i];
    else return usage(argv[0]);
  }
  const int ds = !strcmp(algo, "deltaStep");
  if (!path || (!ds && strcmp(algo, "dijkstra"))) return usage(argv[0]);
  const double delta = is_float ? delta_arg : (double)(1ull << (unsigned)delta_arg);
  if (!(delta > 0) || (!is_float && (delta_arg < 0 || delta_arg > 31))) {
    fprintf(stderr, "bad -delta %g\n", delta_arg);
    return 1;
  }
  uint32_t n, *row, *col, *w;
  uint64_t m;
  if (cpu_read_gr(path, &n, &m, &row, &col, &w)) {
    fprintf(stderr, "cannot read %s\n", path);
    return 1;
  }
  if (src >= n) {
    fprintf(stderr, "startNode %u out of range (n = %u)\n", src, n);
    return 1;
  }
  uint32_t* dist = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint64_t work = 0;
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  int rc = ds ? cpu_delta_step(is_float, n, row, col, w, src, delta, threads ? threads : 1, dist, &work)
              : cpu_dijkstra(is_float, n, row, col, w, src, dist, &work);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (rc) {
    fprintf(stderr, "%s failed\n", ds ? "deltaStep" : "dijkstra");
    return 1;
  }
  const double ms = (t1.tv_sec - t0.tv_sec) * 1e3 + (t1.tv_nsec - t0.tv_nsec) * 1e-6;
  printf("total_work %llu\ntime %.3f\n", (unsigned long long)work, ms);

  const char* base = strrchr(path, '/');
  base = base ? base + 1 : path;
  char out[4096];
  snprintf(out, sizeof(out), "%s_%s_final_dist/%s", ds ? "ds" : "dj", is_float ? "float" : "int", base);
  if (cpu_write_dist(out, is_float, n, dist)) {
    fprintf(stderr, "cannot write %s (does the directory exist?)\n", out);
    return 1;
  }
  FILE* d3 = fopen("/dev/fd/3", "a");
  if (d3) {
    fprintf(d3, "%s %.3f %llu\n", base, ms, (unsigned long long)work);
    fclose(d3);
  } else {
    fprintf(stderr, "fd 3 is not open; result line not written\n");
  }
  free(dist);
  free(row);
  free(col);
  free(w);
  return 0;
}
