// Drop-in `sssp` binary: same command line, input, output and fd-3 line as the reference's ads_int/sssp
// and ads_float/sssp (built from this file with ADDS_FLOAT=0/1).
//
//   sssp [-q] [-g gpunum] [-o output-file|-] [-s startNode] [-l] [-d x] graph-file
//
// Options follow parse_args (ads_int/support.cu:30-76, getopt "g:qo:" + "ls:d:"): -s sets the source
// (:18-23), -g the device (:48-56), -o the output file or stdout (:84-87); -q, -l and -d are accepted and
// ignored as in the reference.  With no arguments it prints usage and exits 1 (kernel.cu:291-294).
// The solve is repeated RUN_LOOP times (8 int, 1 float: ads_int/kernel.cu:213, ads_float/kernel.cu:225)
// and fd 3 receives "%s %.3f %llu\n" = graph basename, mean ms, mean work (kernel.cu:282), where work is
// vertices expanded (int, kernel.cu:54) or edges scanned (float, ads_float/kernel.cu:59-60).
// Deliberate differences: failures print a message and exit non-zero instead of asserting, and a closed
// fd 3 is reported on stderr instead of crashing.
// Built with ADDS_NF=1 (sssp_nf_int / sssp_nf_float) it stands in for the reference's near-far GPU
// comparison binaries instead (nf_int.patch / nf_float.patch: same command line, RUN_LOOP 8, fd-3 line
// with work = worklist items processed) and solves with adds_sssp_nearfar at its default delta.
#include <errno.h>
#include <fcntl.h>
#include <getopt.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/adds.h"

#ifndef ADDS_FLOAT
#define ADDS_FLOAT 0
#endif
#ifndef ADDS_NF
#define ADDS_NF 0
#endif

extern "C" const char* adds_last_error_detail(void);

static void usage(const char* argv0) {
  fprintf(stderr, "usage: %s [-q] [-g gpunum] [-o output-file] [-s startNode] graph-file \n", argv0);
}

static int fail(const char* what, adds_status s) {
  fprintf(stderr, "sssp: %s: %s (%s)\n", what, adds_strerror(s), adds_last_error_detail());
  return 1;
}

#if !ADDS_NF
// ---- tuned parameters ----
// The reference adapts delta inside the binary (wl.h:395-449, 968-1310); this solver is tuned offline
// (adds_autotune) and the binary looks the result up by workload key in tuned_params.json (committed
// entries) and then tune_cache.json (local tuning runs), both in the package directory above bin/, or
// in the file ADDS_TUNE_FILE names ("none": defaults).  The key is the one adds_amd/autotune.py
// workload_key() writes: n, m, weight type, the mean of every max(1, m / 65536)-th weight (%.4g) and the
// number of sources.
static std::string tuned_key(const adds_host_csr& h) {
  const uint64_t step = h.m / 65536 > 1 ? h.m / 65536 : 1;
  double s = 0;
  uint64_t cnt = 0;
  for (uint64_t i = 0; i < h.m; i += step, cnt++) {
    if (h.wtype == ADDS_W_F32) {
      float f;
      memcpy(&f, &h.w[i], 4);
      s += f;
    } else {
      s += h.w[i];
    }
  }
  char mean[64];
  snprintf(mean, sizeof(mean), "%.4g", cnt ? s / (double)cnt : 0.0);
  return "n=" + std::to_string(h.n) + ";m=" + std::to_string(h.m) + ";w=" +
         (h.wtype == ADDS_W_F32 ? "f32" : "u32") + ";meanw=" + mean + ";k=1";
}

static bool read_file(const std::string& path, std::string& out) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  char buf[1 << 16];
  size_t r;
  while ((r = fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, r);
  fclose(f);
  return true;
}

// Fills p from the "params" object of `key` in a JSON cache written by autotune.py (a flat object of
// numbers); returns false when the key is absent.
static bool params_from_json(const std::string& text, const std::string& key, adds_params* p) {
  const size_t at = text.find("\"" + key + "\"");
  if (at == std::string::npos) return false;
  const size_t po = text.find("\"params\"", at);
  const size_t open = po == std::string::npos ? po : text.find('{', po);
  const size_t close = open == std::string::npos ? open : text.find('}', open);
  if (close == std::string::npos) return false;
  const std::string obj = text.substr(open, close - open);
  auto num = [&](const char* name, double* v) {
    const size_t f = obj.find("\"" + std::string(name) + "\"");
    if (f == std::string::npos) return;
    const size_t colon = obj.find(':', f);
    if (colon != std::string::npos) *v = strtod(obj.c_str() + colon + 1, nullptr);
  };
  double delta = 0, light = 0, wg = 0, chunk = 0, nb = 0, parts = 0, conc = 0, te = 0, flags = 0, win = 0, tail = 0,
         hub = 0, pdelta = 0, order = 0, early = 0;
  num("delta", &delta);
  num("light_delta", &light);
  num("wg_per_cu", &wg);
  num("chunk", &chunk);
  num("nbuckets", &nb);
  num("parts", &parts);
  num("concurrency", &conc);
  num("task_edges", &te);
  num("flags", &flags);
  num("edge_windows", &win);
  num("tail_items", &tail);
  num("hub_degree", &hub);
  num("pull_delta", &pdelta);
  num("order_degree", &order);
  num("pull_early", &early);
  p->delta = (float)delta;
  p->light_delta = (float)light;
  p->wg_per_cu = (uint32_t)wg;
  p->chunk = (uint32_t)chunk;
  p->nbuckets = (uint32_t)nb;
  p->parts = (uint32_t)parts;
  p->concurrency = (uint32_t)conc;
  p->task_edges = (uint32_t)te;
  p->flags = (uint32_t)flags & (ADDS_F_NO_LOCAL | ADDS_F_NO_WAIT | ADDS_F_NO_PULL | ADDS_F_PULL_ALL_LIGHT |
                                ADDS_F_PULL_SPLIT | ADDS_F_LOCAL_SPARSE | ADDS_F_FULL_FIRST);  // tuning flags only
  p->edge_windows = (uint32_t)win;
  p->tail_items = (uint32_t)tail;
  p->hub_degree = (uint32_t)hub;
  p->pull_delta = (float)pdelta;
  p->order_degree = (uint32_t)order;
  p->pull_early = (uint32_t)early;
  return true;
}

static std::string lookup_tuned(const adds_host_csr& h, adds_params* p) {
  std::vector<std::string> files;
  const char* env = getenv("ADDS_TUNE_FILE");
  if (env) {
    if (strcmp(env, "none") == 0) return "";
    files.push_back(env);
  } else {
    char exe[4096];
    const ssize_t len = readlink("/proc/self/exe", exe, sizeof(exe) - 1);
    if (len > 0) {
      exe[len] = '\0';
      std::string dir(exe);
      dir = dir.substr(0, dir.find_last_of('/'));   // .../bin
      dir = dir.substr(0, dir.find_last_of('/'));   // the package directory
      files.push_back(dir + "/tuned_params.json");
      files.push_back(dir + "/tune_cache.json");
    }
  }
  const std::string key = tuned_key(h);
  for (const std::string& f : files) {
    std::string text;
    if (read_file(f, text) && params_from_json(text, key, p)) return f;
  }
  return "";
}
#endif

int main(int argc, char** argv) {
  if (argc == 1) {
    usage(argv[0]);
    return 1;
  }
  int device = 0;
  long start_node = 0;
  const char* output = nullptr;
  int c;
  while ((c = getopt(argc, argv, "g:qo:ls:d:")) != -1) {
    switch (c) {
      case 'q': break;
      case 'o': output = optarg; break;
      case 'g': {
        char* end;
        errno = 0;
        device = (int)strtol(optarg, &end, 10);
        if (errno != 0 || *end != '\0') {
          fprintf(stderr, "Invalid GPU device '%s'. An integer must be specified.\n", optarg);
          return EXIT_FAILURE;
        }
        break;
      }
      case 's':
        start_node = atol(optarg);
        if (start_node < 0) {
          fprintf(stderr, "sssp: start node must be >= 0\n");
          return 1;
        }
        break;
      case 'l':
      case 'd': break;
      default: usage(argv[0]); return 255;
    }
  }
  if (optind >= argc) {
    usage(argv[0]);
    return 255;
  }
  const char* input = argv[optind];
  const adds_wtype wt = ADDS_FLOAT ? ADDS_W_F32 : ADDS_W_U32;

  adds_status s = adds_set_device(device);
  if (s != ADDS_OK) return fail("device", s);
  adds_host_csr h;
  s = adds_read_gr(input, wt, &h);
  if (s != ADDS_OK) return fail(input, s);
  if ((uint64_t)start_node >= h.n) {
    fprintf(stderr, "sssp: start node %ld out of range (n=%u)\n", start_node, h.n);
    return 1;
  }
  printf("nnodes=%u, nedges=%llu.\n", h.n, (unsigned long long)h.m);
  adds_graph* g = nullptr;
  s = adds_graph_create(h.row_ptr, h.col, h.w, h.n, h.m, wt, &g);
  if (s != ADDS_OK) return fail("graph", s);

  const int run_loop = (ADDS_FLOAT && !ADDS_NF) ? 1 : 8;
  std::vector<uint32_t> dist(h.n);
  adds_params p;
  memset(&p, 0, sizeof(p));
#if !ADDS_NF
  const std::string tuned_from = lookup_tuned(h, &p);
  if (!tuned_from.empty()) printf("tuned parameters from %s\n", tuned_from.c_str());
  // a symmetric graph (every edge with its reverse) lets settled buckets' heavy edges be pulled
  int symmetric = 0;
  s = adds_graph_check_symmetric(g, &symmetric);
  if (s != ADDS_OK) return fail("symmetry check", s);
  printf("symmetric: %s\n", symmetric ? "yes" : "no");
#endif
  p.flags |= ADDS_F_HOST_OUT;
  double agg_ms = 0;
  unsigned long long agg_work = 0;
  for (int loop = 0; loop < run_loop; loop++) {
    adds_stats st;
#if ADDS_NF
    s = adds_sssp_nearfar(g, (uint32_t)start_node, dist.data(), 0.f, ADDS_F_HOST_OUT, &st, nullptr);
#else
    s = adds_sssp(g, (uint32_t)start_node, dist.data(), &p, &st, nullptr);
#endif
    if (s != ADDS_OK) return fail("solve", s);
    unsigned long long work = (ADDS_FLOAT && !ADDS_NF) ? st.edges_scanned : st.vertices_expanded;
    printf("%s Measured time for sample = %.3fs\n", input, st.ms / 1000.0);
    printf("total work is %llu\n", work);
    agg_ms += st.ms;
    agg_work += work;
  }
  // basename of the input path, as readFromGR's strtok loop keeps the last '/'-separated token
  std::string base(input);
  size_t slash = base.find_last_of('/');
  if (slash != std::string::npos) base = base.substr(slash + 1);
  if (fcntl(3, F_GETFD) != -1) {
    FILE* d3 = fdopen(dup(3), "a");
    if (d3) {
      fprintf(d3, "%s %.3f %llu\n", base.c_str(), agg_ms / run_loop, agg_work / (unsigned long long)run_loop);
      fclose(d3);
    }
  } else {
    fprintf(stderr, "sssp: fd 3 is not open; result line not written\n");
  }
  if (output) {
    s = adds_write_dist_text(output, dist.data(), h.n, wt);
    if (s != ADDS_OK) return fail(output, s);
  }
  adds_graph_destroy(g);
  adds_host_csr_free(&h);
  return 0;
}
