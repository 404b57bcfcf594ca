// adds_convert: text graph formats -> Galois .gr v1, the input of the drop-in `sssp` binaries.
//
//   adds_convert (mtx|dimacs|gr) (int|float) <in> <out.gr>
//
// mtx: MatrixMarket coordinate (SuiteSparse); dimacs: DIMACS shortest-path ("p sp", "a u v w"); gr:
// re-writes a .gr file (for example to change how its weights are read).  Weights are |value|: the
// nearest float or integer (adds_read_mtx / adds_read_dimacs in include/adds.h).  Prints "n m" on stdout.
#include <cstdio>
#include <cstring>

#include "../../include/adds.h"

int main(int argc, char** argv) {
  if (argc != 5 || (strcmp(argv[2], "int") && strcmp(argv[2], "float"))) {
    fprintf(stderr, "usage: %s (mtx|dimacs|gr) (int|float) <in> <out.gr>\n", argv[0]);
    return 2;
  }
  const adds_wtype wt = strcmp(argv[2], "float") ? ADDS_W_U32 : ADDS_W_F32;
  adds_host_csr g;
  adds_status s;
  if (!strcmp(argv[1], "mtx")) s = adds_read_mtx(argv[3], wt, &g);
  else if (!strcmp(argv[1], "dimacs")) s = adds_read_dimacs(argv[3], wt, &g);
  else if (!strcmp(argv[1], "gr")) s = adds_read_gr(argv[3], wt, &g);
  else {
    fprintf(stderr, "unknown input format %s\n", argv[1]);
    return 2;
  }
  if (s == ADDS_OK) s = adds_write_gr(argv[4], &g);
  if (s != ADDS_OK) {
    fprintf(stderr, "adds_convert: %s (%s)\n", adds_strerror(s), adds_last_error_detail());
    return 1;
  }
  printf("%u %llu\n", g.n, (unsigned long long)g.m);
  adds_host_csr_free(&g);
  return 0;
}
