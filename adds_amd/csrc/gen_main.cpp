// adds_gen: writes the seeded synthetic graphs as Galois .gr files for the drop-in `sssp` binaries.
//
//   adds_gen rmat  <scale> <edge_factor> <seed> <wkind> <lo> <hi> out.gr
//   adds_gen grid  <rows> <cols> <shortcuts> <seed> <wkind> <lo> <hi> out.gr
//   adds_gen er    <n> <avg_degree> <seed> <wkind> <lo> <hi> out.gr
//
// wkind: u32 (uniform integer in [lo,hi]), f32 (uniform float in [lo,hi)), lognormal (exp(N(0,lo^2))).
// RMAT uses a/b/c = 0.57/0.19/0.19.  Prints "n m max_degree_vertex" on stdout.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/adds.h"

static bool wkind(const char* s, adds_wkind* k) {
  if (!strcmp(s, "u32")) *k = ADDS_WK_U32_UNIFORM;
  else if (!strcmp(s, "f32")) *k = ADDS_WK_F32_UNIFORM;
  else if (!strcmp(s, "lognormal")) *k = ADDS_WK_F32_LOGNORMAL;
  else return false;
  return true;
}

int main(int argc, char** argv) {
  adds_host_csr g;
  adds_status s = ADDS_EINVAL;
  adds_wkind wk;
  const char* out = nullptr;
  if (argc == 9 && !strcmp(argv[1], "rmat") && wkind(argv[5], &wk)) {
    s = adds_gen_rmat((uint32_t)atoi(argv[2]), (uint32_t)atoi(argv[3]), 0.57, 0.19, 0.19, strtoull(argv[4], 0, 10),
                      wk, atof(argv[6]), atof(argv[7]), &g);
    out = argv[8];
  } else if (argc == 10 && !strcmp(argv[1], "grid") && wkind(argv[6], &wk)) {
    s = adds_gen_grid((uint32_t)atoi(argv[2]), (uint32_t)atoi(argv[3]), strtoull(argv[4], 0, 10),
                      strtoull(argv[5], 0, 10), wk, atof(argv[7]), atof(argv[8]), &g);
    out = argv[9];
  } else if (argc == 9 && !strcmp(argv[1], "er") && wkind(argv[5], &wk)) {
    s = adds_gen_er((uint32_t)atoi(argv[2]), atof(argv[3]), strtoull(argv[4], 0, 10), wk, atof(argv[6]),
                    atof(argv[7]), &g);
    out = argv[8];
  } else {
    fprintf(stderr,
            "usage: %s rmat <scale> <ef> <seed> <wkind> <lo> <hi> out.gr\n"
            "       %s grid <rows> <cols> <shortcuts> <seed> <wkind> <lo> <hi> out.gr\n"
            "       %s er <n> <avg_degree> <seed> <wkind> <lo> <hi> out.gr\n",
            argv[0], argv[0], argv[0]);
    return 2;
  }
  if (s != ADDS_OK) {
    fprintf(stderr, "adds_gen: %s\n", adds_strerror(s));
    return 1;
  }
  s = adds_write_gr(out, &g);
  if (s != ADDS_OK) {
    fprintf(stderr, "adds_gen: %s: %s\n", out, adds_strerror(s));
    return 1;
  }
  printf("%u %llu %u\n", g.n, (unsigned long long)g.m, adds_max_degree_vertex(&g));
  adds_host_csr_free(&g);
  return 0;
}
