/* gr_io.c -- the .gr v1 reader and the distance-text writer of the CPU baselines.
 *
 * .gr v1 (csr_graph.cu:236-281): four u64 words (version 1, edge-data size, nodes, edges), nodes u64
 * row ends, edges u32 destinations, one u32 of padding when edges is odd, then edges u32 weights when
 * the edge-data size is non-zero.  Distance text (cpu_sssp_int.patch:129-153, cpu_sssp_float.patch:
 * 225-248): one "<v> <d>" line per vertex, "%u" for int, "%f" for float, "INF" when unreached. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Returns 0 and malloc'd arrays (row n+1, col m, w m; free with free()), or non-zero on a missing file
 * (1), a bad header (2), a short file (3) or row ends that are not non-decreasing up to m (4). */
int cpu_read_gr(const char* path, uint32_t* n_out, uint64_t* m_out, uint32_t** row_out, uint32_t** col_out,
                uint32_t** w_out) {
  FILE* f = fopen(path, "rb");
  if (!f) return 1;
  uint64_t h[4];
  if (fread(h, sizeof(uint64_t), 4, f) != 4 || h[0] != 1 || h[2] >= 0xFFFFFFFFull || h[3] > 0xFFFFFFFFull) {
    fclose(f);
    return 2;
  }
  const uint64_t n = h[2], m = h[3];
  uint64_t* ends = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
  uint32_t* row = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1));
  uint32_t* col = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
  uint32_t* w = (uint32_t*)calloc(m ? m : 1, sizeof(uint32_t));
  int rc = fread(ends, sizeof(uint64_t), n, f) == n && fread(col, sizeof(uint32_t), m, f) == m ? 0 : 3;
  if (!rc && (m & 1)) {
    uint32_t pad;
    if (fread(&pad, sizeof(pad), 1, f) != 1) rc = 3;
  }
  if (!rc && h[1] && fread(w, sizeof(uint32_t), m, f) != m) rc = 3;
  fclose(f);
  row[0] = 0;
  for (uint64_t i = 0; i < n && !rc; i++) {
    if (ends[i] < row[i] || ends[i] > m) rc = 4;
    row[i + 1] = (uint32_t)ends[i];
  }
  if (!rc && n && row[n] != m) rc = 4;
  free(ends);
  if (rc) {
    free(row);
    free(col);
    free(w);
    return rc;
  }
  *n_out = (uint32_t)n;
  *m_out = m;
  *row_out = row;
  *col_out = col;
  *w_out = w;
  return 0;
}

int cpu_write_dist(const char* path, int is_float, uint32_t n, const uint32_t* dist) {
  FILE* f = fopen(path, "w");
  if (!f) return 1;
  const uint32_t inf = is_float ? 0x7F800000u : 0xFFFFFFFFu;
  for (uint32_t v = 0; v < n; v++) {
    if (dist[v] == inf) {
      fprintf(f, "%u INF\n", v);
    } else if (is_float) {
      float x;
      memcpy(&x, &dist[v], 4);
      fprintf(f, "%u %f\n", v, (double)x);
    } else {
      fprintf(f, "%u %u\n", v, dist[v]);
    }
  }
  return fclose(f) ? 1 : 0;
}
