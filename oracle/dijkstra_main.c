/* oracle_dijkstra: the CPU oracle as a command, standing where the reference ran Galois Dijkstra
 * (run_all.sh:93-107, "-algo=dijkstra -t 1") to produce dj_{int,float}_final_dist/<graph>.
 * Test infrastructure only.
 *
 *   oracle_dijkstra (int|float) <src> <graph.gr> <out-file>
 *
 * Prints the Dijkstra time in ms (steady clock around the algorithm only, like Galois execTime,
 * cpu_sssp_int.patch:115) on stdout. */
#define _POSIX_C_SOURCE 200809L
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

int oracle_read_gr(const char*, uint32_t*, uint64_t*, uint32_t**, uint32_t**, uint32_t**);
int oracle_dijkstra(int, uint32_t, const uint32_t*, const uint32_t*, const uint32_t*, const uint32_t*, uint32_t,
                    uint32_t*);
int oracle_write_dist(const char*, int, uint32_t, const uint32_t*);

int main(int argc, char** argv) {
  if (argc != 5 || (strcmp(argv[1], "int") && strcmp(argv[1], "float"))) {
    fprintf(stderr, "usage: %s (int|float) <src> <graph.gr> <out-file>\n", argv[0]);
    return 2;
  }
  int is_float = !strcmp(argv[1], "float");
  uint32_t src = (uint32_t)strtoul(argv[2], 0, 10), n;
  uint64_t m;
  uint32_t *row, *col, *w;
  if (oracle_read_gr(argv[3], &n, &m, &row, &col, &w)) {
    fprintf(stderr, "cannot read %s\n", argv[3]);
    return 1;
  }
  if (src >= n) {
    fprintf(stderr, "source out of range\n");
    return 1;
  }
  uint32_t* dist = (uint32_t*)malloc(sizeof(uint32_t) * n);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  int ovf = oracle_dijkstra(is_float, n, row, col, w, &src, 1, dist);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  printf("%.3f\n", (t1.tv_sec - t0.tv_sec) * 1e3 + (t1.tv_nsec - t0.tv_nsec) / 1e6);
  if (ovf) fprintf(stderr, "warning: integer distance overflow\n");
  if (oracle_write_dist(argv[4], is_float, n, dist)) return 1;
  free(dist); free(row); free(col); free(w);
  return 0;
}
