// C ABI of the ADDS solver (include/adds.h): device-resident graphs, solver workspace, solve
// orchestration (dist fill + persistent kernel + cleanup on one stream), status and watchdog,
// batched/union solves and HIP-graph capture.
//
// Reference seams replaced here: CSRGraph::copy_to_gpu / allocOnDevice (csr_graph.cu:86-105, 309-328),
// gg_main's grid sizing, worklist sizing and timed solve (ads_int/kernel.cu:185-288), worklist::alloc /
// reinit / set_param (wl.h:203-449) and the weight profiler (kernel.cu:151-183, 238-257).
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include <algorithm>

#include "../../include/adds.h"
#include "device.h"

using namespace adds;

static thread_local std::string g_detail;

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) {                                                                 \
      g_detail = std::string(#expr) + ": " + hipGetErrorString(_e);                         \
      return (_e == hipErrorOutOfMemory || _e == hipErrorMemoryAllocation) ? ADDS_ENOMEM    \
                                                                           : ADDS_EHIP;     \
    }                                                                                       \
  } while (0)

constexpr uint32_t TRACE_CAP = 1u << 16;

struct Workspace {
  uint32_t* st = nullptr;
  uint64_t* pool = nullptr;
  uint32_t* tab = nullptr;
  uint32_t* hrec = nullptr;  // edge-range records
  uint32_t hrec_cap = 0;
  uint32_t* srcs = nullptr;
  uint32_t nq = 0, nb = 0, maxv = 0, pool_chunks = 0, srcs_cap = 0;
  uint32_t* host_state = nullptr;       // pinned copy of the state words before the queues
  uint32_t* host_abort = nullptr;       // pinned, device-visible abort word
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;  // solve start, solve end, state copied
  hipEvent_t evk0 = nullptr, evk1 = nullptr;                // around the persistent kernel
  hipStream_t last_stream = nullptr;
  bool pending = false;                 // a NO_SYNC solve has not been collected
  bool kernel_events = false;           // evk0/evk1 bracket the last solve's persistent kernel
  bool pool_fixed = false;              // the last solve's pool size was set by the caller
  uint32_t pool_shift = 0;              // default pool sizes are doubled this often (grown after overflows)
  uint64_t* trace = nullptr;            // bucket timeline (ADDS_TRACE=<file>), TRACE_CAP records
  uint64_t trace_t0 = 0;
  double timeout_ms = 0;
  uint4* ptask = nullptr;               // pull phases: long-row tasks
  uint32_t ptask_cap = 0;
  uint64_t* pstage = nullptr;           // re-bucketing pulls: staged queue entries, one per vertex-instance
  uint32_t* pcnt = nullptr;             // ... and their count per 256-instance chunk
  uint64_t pstage_cap = 0;
};

// Every graph handle gets a number no other handle of this process has had, renewed whenever its
// weights change: a cached union of graphs is keyed by (handle, number), so neither a weight update nor
// a new handle allocated at a freed handle's address can match a stale union.
static std::atomic<uint64_t> g_next_gen{1};

struct adds_graph {
  int device = 0;
  uint64_t gen = 0;
  uint32_t n = 0;
  uint64_t m = 0;
  adds_wtype wt = ADDS_W_U32;
  // Declared symmetric (every edge (u, v, w) has its reverse (v, u, w)): pull phases may then read a
  // vertex's in-edges from its own row.  sym_dev mirrors it on the device, where captured solves read it.
  bool symmetric = false;
  uint32_t* sym_dev = nullptr;
  uint32_t* row = nullptr;
  uint32_t* col = nullptr;
  uint32_t* w = nullptr;
  float ave_wt = 1.f;  // mean sampled weight (the reference's profiler statistic)
  // largest weight (bits; float weights are non-negative, so bits order them), when known: with a light
  // threshold at or above it no edge is heavy and no pull phase can be requested
  bool max_w_known = false;
  uint32_t max_w = 0;
  // light-first copy of the rows for the solver (col/w stay in input order for the backward pass and
  // the device-array accessor); valid for split_delta
  uint2* scw = nullptr;
  uint32_t* lcnt = nullptr;
  uint4* rinfo = nullptr;  // per-vertex row record (row start, row end, light edges)
  uint4* phead = nullptr;  // per-vertex first two heavy edges of the split rows (pull phases)
  uint32_t* split_aux = nullptr;
  float split_delta = 0.f;
  Workspace ws;
  uint32_t* outbuf = nullptr;  // device staging for host outputs
  uint64_t outbuf_cap = 0;
  // transpose (backward)
  uint32_t* rrow = nullptr;
  uint32_t* reid = nullptr;
  uint32_t* rsrc = nullptr;
  float* rw = nullptr;          // weights in transposed order, valid for weight generation rw_gen
  uint64_t rw_gen = 0;
  uint8_t* bwd = nullptr;       // backward workspace (per vertex-instance arrays, gradient sums, counters)
  uint64_t bwd_bytes = 0;
  unsigned long long* bwd_host = nullptr;  // pinned copy of the counters
  // batched union cache: the members' (handle, generation) pairs it was built from
  std::vector<std::pair<const adds_graph*, uint64_t>> union_key;
  adds_graph* union_g = nullptr;
  std::vector<uint32_t> union_off;
};

// A solve captured in a HIP graph.  It owns its workspace (queues, pool, chunk maps, range records,
// status words), so nothing an eager solve on the same graph reallocates can move under it; a batched
// capture also owns its block-diagonal union of the member graphs (a snapshot taken at capture).
struct adds_solve_graph {
  adds_graph* g = nullptr;       // the graph solved: the caller's, or the owned union
  adds_graph* owned = nullptr;   // batched capture: the union (destroyed with the solve graph)
  Workspace ws;
  float split_light = 0.f;       // light threshold of the light-first rows the capture was made for
  uint32_t* srcs = nullptr;      // batched capture: sources mapped into the union (device)
  uint64_t* scatter = nullptr;   // batched capture: per instance (destination pointer, union offset | n)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" const char* adds_strerror(adds_status s) {
  switch (s) {
    case ADDS_OK: return "ok";
    case ADDS_EINVAL: return "invalid argument";
    case ADDS_ENOMEM: return "out of memory";
    case ADDS_EOVERFLOW: return "worklist capacity exceeded";
    case ADDS_ETIMEOUT: return "solve timed out or was aborted";
    case ADDS_EHIP: return "HIP runtime error";
    case ADDS_EIO: return "I/O error";
    case ADDS_ENODEV: return "no GPU device";
  }
  return "unknown status";
}

extern "C" const char* adds_last_error_detail(void) { return g_detail.c_str(); }
namespace adds {
void set_error_detail(const std::string& s) { g_detail = s; }
float graph_mean_weight(const adds_graph* g) { return g->ave_wt; }
}  // namespace adds

extern "C" adds_status adds_set_device(int device) {
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) return ADDS_ENODEV;
  if (device < 0 || device >= cnt) return ADDS_EINVAL;
  HIPCHK(hipSetDevice(device));
  return ADDS_OK;
}

extern "C" uint32_t adds_device_cu_count(void) {
  int dev = 0, cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return (uint32_t)cu;
}

static adds_status free_workspace(Workspace& ws) {
  if (ws.st) (void)hipFree(ws.st);
  if (ws.pool) (void)hipFree(ws.pool);
  if (ws.tab) (void)hipFree(ws.tab);
  if (ws.hrec) (void)hipFree(ws.hrec);
  if (ws.srcs) (void)hipFree(ws.srcs);
  if (ws.trace) (void)hipFree(ws.trace);
  if (ws.ptask) (void)hipFree(ws.ptask);
  if (ws.pstage) (void)hipFree(ws.pstage);
  if (ws.pcnt) (void)hipFree(ws.pcnt);
  if (ws.host_state) (void)hipHostFree(ws.host_state);
  if (ws.host_abort) (void)hipHostFree(ws.host_abort);
  if (ws.ev0) (void)hipEventDestroy(ws.ev0);
  if (ws.ev1) (void)hipEventDestroy(ws.ev1);
  if (ws.ev2) (void)hipEventDestroy(ws.ev2);
  if (ws.evk0) (void)hipEventDestroy(ws.evk0);
  if (ws.evk1) (void)hipEventDestroy(ws.evk1);
  ws = Workspace();
  return ADDS_OK;
}

static void release_graph_memory(adds_graph* g) {
  free_workspace(g->ws);
  if (g->row) (void)hipFree(g->row);
  if (g->col) (void)hipFree(g->col);
  if (g->w) (void)hipFree(g->w);
  if (g->outbuf) (void)hipFree(g->outbuf);
  if (g->sym_dev) (void)hipFree(g->sym_dev);
  g->sym_dev = nullptr;
  if (g->rrow) (void)hipFree(g->rrow);
  if (g->reid) (void)hipFree(g->reid);
  if (g->rsrc) (void)hipFree(g->rsrc);
  if (g->rw) (void)hipFree(g->rw);
  if (g->bwd) (void)hipFree(g->bwd);
  if (g->bwd_host) (void)hipHostFree(g->bwd_host);
  g->rw = nullptr;
  g->bwd = nullptr;
  g->bwd_bytes = 0;
  g->bwd_host = nullptr;
  if (g->scw) (void)hipFree(g->scw);
  if (g->lcnt) (void)hipFree(g->lcnt);
  if (g->rinfo) (void)hipFree(g->rinfo);
  if (g->phead) (void)hipFree(g->phead);
  g->rinfo = nullptr;
  g->phead = nullptr;
  if (g->split_aux) (void)hipFree(g->split_aux);
  g->scw = nullptr;
  g->lcnt = g->split_aux = nullptr;
  g->split_delta = 0.f;
  if (g->union_g) {
    release_graph_memory(g->union_g);
    delete g->union_g;
  }
  g->row = g->col = g->w = g->outbuf = g->rrow = g->reid = g->rsrc = nullptr;
  g->union_g = nullptr;
}

// Mean of up to 2^20 evenly spaced weights: the statistic the reference's profiler_kernel samples
// (ads_int/kernel.cu:151-183), computed once per graph on the host copy.
static float sample_mean_weight(const uint32_t* w, uint64_t m, adds_wtype wt) {
  if (m == 0) return 1.f;
  const uint64_t samples = std::min<uint64_t>(m, 1u << 20);
  double s = 0;
  for (uint64_t i = 0; i < samples; i++) {
    uint32_t x = w[(i * m) / samples];
    if (wt == ADDS_W_F32) {
      float f;
      memcpy(&f, &x, 4);
      s += f;
    } else {
      s += x;
    }
  }
  double mean = s / (double)samples;
  return mean > 0 ? (float)mean : 1.f;
}

static adds_status alloc_graph(uint32_t n, uint64_t m, adds_wtype wt, adds_graph** out) {
  if (!out || n == 0 || n >= 0xFFFFFFFFu || m >= 0xFFFFFFFFull) return ADDS_EINVAL;
  int dev = 0;
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) return ADDS_ENODEV;
  HIPCHK(hipGetDevice(&dev));
  adds_graph* g = new adds_graph();
  g->device = dev;
  g->gen = g_next_gen.fetch_add(1);
  g->n = n;
  g->m = m;
  g->wt = wt;
  hipError_t e = hipMalloc(&g->row, ((size_t)n + 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&g->col, std::max<uint64_t>(m, 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&g->w, std::max<uint64_t>(m, 1) * 4);
  // the solver's light-first copy of the rows: allocated with the graph (filled by the first solve for
  // its light threshold), so that no solve needs to allocate it
  if (e == hipSuccess) e = hipMalloc(&g->scw, (m + 1) * 8);  // one pad entry: the last 16-byte pair is in bounds
  if (e == hipSuccess) e = hipMalloc(&g->lcnt, std::max<uint64_t>(n, 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&g->rinfo, std::max<uint64_t>(n, 1) * 16);
  if (e == hipSuccess) e = hipMalloc(&g->phead, std::max<uint64_t>(n, 1) * 16);
  if (e == hipSuccess) e = hipMalloc(&g->split_aux, split_aux_words(m) * 4);
  if (e == hipSuccess) e = hipMalloc(&g->sym_dev, 256);
  if (e == hipSuccess) e = launch_fill_u32(g->sym_dev, 1, 0u, nullptr);
  if (e != hipSuccess) {
    g_detail = std::string("graph allocation: ") + hipGetErrorString(e);
    release_graph_memory(g);
    delete g;
    return ADDS_ENOMEM;
  }
  *out = g;
  return ADDS_OK;
}

extern "C" adds_status adds_graph_create(const uint32_t* row_ptr, const uint32_t* col, const void* w, uint32_t n,
                                         uint64_t m, adds_wtype wt, adds_graph** out) {
  if (!row_ptr || (m && (!col || !w)) || !out) return ADDS_EINVAL;
  if (row_ptr[0] != 0 || row_ptr[n] != m) {
    g_detail = "row_ptr[0] must be 0 and row_ptr[n] must equal m";
    return ADDS_EINVAL;
  }
  for (uint32_t v = 0; v < n; v++)
    if (row_ptr[v + 1] < row_ptr[v]) {
      g_detail = "row_ptr is not monotone";
      return ADDS_EINVAL;
    }
  const uint32_t* wb = (const uint32_t*)w;
  uint32_t maxw = 0;
  for (uint64_t i = 0; i < m; i++) {
    if (col[i] >= n) {
      g_detail = "edge destination out of range";
      return ADDS_EINVAL;
    }
    maxw = std::max(maxw, wt == ADDS_W_F32 ? (wb[i] & 0x7FFFFFFFu) : wb[i]);
    if (wt == ADDS_W_F32) {
      uint32_t x = wb[i] & 0x7FFFFFFFu;  // -0.0 is accepted (it adds like +0.0)
      if (x >= 0x7F800000u || (wb[i] >> 31 && x != 0)) {
        g_detail = "float weights must be finite and not negative";
        return ADDS_EINVAL;
      }
    }
  }
  adds_status rc = alloc_graph(n, m, wt, out);
  if (rc != ADDS_OK) return rc;
  adds_graph* g = *out;
  HIPCHK(hipMemcpy(g->row, row_ptr, ((size_t)n + 1) * 4, hipMemcpyHostToDevice));
  if (m) {
    HIPCHK(hipMemcpy(g->col, col, m * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g->w, w, m * 4, hipMemcpyHostToDevice));
  }
  g->ave_wt = sample_mean_weight(wb, m, wt);
  g->max_w = maxw;
  g->max_w_known = true;
  return ADDS_OK;
}

extern "C" adds_status adds_graph_create_device(const uint32_t* row_ptr_dev, const uint32_t* col_dev,
                                                const void* w_dev, uint32_t n, uint64_t m, adds_wtype wt,
                                                adds_graph** out) {
  if (!row_ptr_dev || (m && (!col_dev || !w_dev)) || !out) return ADDS_EINVAL;
  adds_status rc = alloc_graph(n, m, wt, out);
  if (rc != ADDS_OK) return rc;
  adds_graph* g = *out;
  HIPCHK(hipMemcpy(g->row, row_ptr_dev, ((size_t)n + 1) * 4, hipMemcpyDeviceToDevice));
  if (m) {
    HIPCHK(hipMemcpy(g->col, col_dev, m * 4, hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(g->w, w_dev, m * 4, hipMemcpyDeviceToDevice));
    const uint64_t samples = std::min<uint64_t>(m, 1u << 16);
    std::vector<uint32_t> hw(samples);
    for (uint64_t i = 0; i < samples; i++)
      HIPCHK(hipMemcpy(&hw[i], (const uint32_t*)w_dev + (i * m) / samples, 4, hipMemcpyDeviceToHost));
    g->ave_wt = sample_mean_weight(hw.data(), samples, wt);
  }
  return ADDS_OK;
}

extern "C" adds_status adds_graph_set_weights_device(adds_graph* g, const void* w_dev, void* stream) {
  if (!g || (!w_dev && g->m)) return ADDS_EINVAL;
  HIPCHK(hipSetDevice(g->device));
  if (g->m) HIPCHK(hipMemcpyAsync(g->w, w_dev, g->m * 4, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  if (g->split_delta != 0.f && g->n)  // re-split in place (same buffers, so captured solves see the new weights)
    HIPCHK(launch_split(g->row, g->col, g->w, g->n, g->split_delta, g->split_delta, 1.0f / g->split_delta,
                        g->wt == ADDS_W_F32, g->scw, g->lcnt, g->split_aux, g->rinfo, g->phead,
                        (hipStream_t)stream));
  g->max_w_known = false;
  // new weights need not be symmetric: pull phases stop (also in solves captured earlier) until the
  // caller declares the graph symmetric again
  if (g->symmetric) {
    HIPCHK(launch_fill_u32(g->sym_dev, 1, 0u, (hipStream_t)stream));
    g->symmetric = false;
  }
  // a new generation: every cached union that holds this graph (on any handle) no longer matches
  g->gen = g_next_gen.fetch_add(1);
  if (g->union_g) {  // weights changed: the union copy is stale
    release_graph_memory(g->union_g);
    delete g->union_g;
    g->union_g = nullptr;
    g->union_key.clear();
  }
  return ADDS_OK;
}

extern "C" adds_status adds_graph_set_symmetric(adds_graph* g, int symmetric) {
  if (!g) return ADDS_EINVAL;
  HIPCHK(hipSetDevice(g->device));
  g->symmetric = symmetric != 0;
  HIPCHK(launch_fill_u32(g->sym_dev, 1, g->symmetric ? 1u : 0u, nullptr));
  HIPCHK(hipDeviceSynchronize());
  return ADDS_OK;
}

// Counts the edges that break "symmetric with strictly ascending rows": an edge (u, v, w) whose reverse
// (v, u, w) is missing (binary search for u in v's row; equal weight bits), or whose destination is not
// above its predecessor's in the row.  One edge per thread; its row is found by binary search over the
// row offsets (the lanes of a wavefront mostly share one row, so those probes are broadcasts).
__global__ static void k_check_sym(const uint32_t* row, const uint32_t* col, const uint32_t* w, uint32_t n,
                                   uint64_t m, unsigned long long* bad) {
  __shared__ unsigned long long s_bad[256 / 64];
  unsigned long long nb = 0;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = n;  // row[lo] <= e < row[hi]
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (row[mid] <= e) lo = mid;
      else hi = mid;
    }
    const uint32_t u = lo, v = col[e], wt = w[e];
    if (e > row[u] && col[e - 1] >= v) nb++;
    uint32_t a = row[v], b = row[v + 1];
    const uint32_t end = b;
    while (a < b) {
      const uint32_t mid = a + (b - a) / 2;
      if (col[mid] < u) a = mid + 1;
      else b = mid;
    }
    if (!(a < end && col[a] == u && w[a] == wt)) nb++;
  }
  for (int d = 32; d >= 1; d >>= 1) nb += __shfl_xor(nb, d, 64);
  if ((threadIdx.x & 63) == 0) s_bad[threadIdx.x >> 6] = nb;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (uint32_t i = 0; i < blockDim.x / 64; i++) t += s_bad[i];
    if (t) atomicAdd(bad, t);
  }
}

extern "C" adds_status adds_graph_check_symmetric(adds_graph* g, int* symmetric) {
  if (!g || !symmetric) return ADDS_EINVAL;
  HIPCHK(hipSetDevice(g->device));
  unsigned long long* bad = nullptr;
  HIPCHK(hipMalloc(&bad, 8));
  unsigned long long hb = 0;
  hipError_t e = hipMemcpy(bad, &hb, 8, hipMemcpyHostToDevice);
  if (e == hipSuccess && g->m) {
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((g->m + 255) / 256, 8192);
    hipLaunchKernelGGL(k_check_sym, dim3(blocks), dim3(256), 0, nullptr, g->row, g->col, g->w, g->n, g->m, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&hb, bad, 8, hipMemcpyDeviceToHost);
  (void)hipFree(bad);
  if (e != hipSuccess) {
    g_detail = std::string("symmetry check: ") + hipGetErrorString(e);
    return ADDS_EHIP;
  }
  *symmetric = hb == 0 ? 1 : 0;
  return adds_graph_set_symmetric(g, *symmetric);
}

extern "C" void adds_graph_destroy(adds_graph* g) {
  if (!g) return;
  (void)hipSetDevice(g->device);
  (void)hipDeviceSynchronize();
  release_graph_memory(g);
  delete g;
}

extern "C" adds_status adds_graph_info(const adds_graph* g, uint32_t* n, uint64_t* m, adds_wtype* wt) {
  if (!g) return ADDS_EINVAL;
  if (n) *n = g->n;
  if (m) *m = g->m;
  if (wt) *wt = g->wt;
  return ADDS_OK;
}

extern "C" adds_status adds_graph_device_arrays(const adds_graph* g, const uint32_t** row_ptr,
                                                const uint32_t** col, const void** w) {
  if (!g) return ADDS_EINVAL;
  if (row_ptr) *row_ptr = g->row;
  if (col) *col = g->col;
  if (w) *w = g->w;
  return ADDS_OK;
}

// ------------------------------------------------------------------------------------------------
// parameters and workspace
// ------------------------------------------------------------------------------------------------
struct Resolved {
  float delta, light_delta;
  uint32_t wg_per_cu, chunk, nb, parts, hub_degree, task_edges, concur, flags, windows, tail_items;
  uint64_t pool_entries;
  bool pool_fixed;
  double timeout_ms;
  uint32_t pull_slots, ptask_cap;  // pull phases per solve (0: none) and the long-row task list size
  float pull_delta;                // bucket width after the first pull (0: delta)
  uint64_t pull_min;               // heavy edges a bucket needs before a pull phase is considered
  uint32_t pull_at;                // ADDS_F_PULL_SPLIT: the bucket of the forced first pull (PULL_AT_OFF: off)
  uint32_t order_degree;           // ... and the light degree from which vertices wait for their bucket there
  uint32_t pull_early;             // ... and the pending-item count at which the pull may come early
  uint64_t total_n;                // vertex-instances of the solve
};

// Default bucket width.  The reference derives it from the mean weight and degree
// (wl.h:427-449: nearest power of two to ave_wt/ave_degree/32, at least 4, then adapted online); this
// solver has no online adaptation.  Measured optima on this solver (RMAT scales 16-24, 2-D grids,
// Erdos-Renyi) sit near ave_wt * 64 / ave_degree up to about 4M vertices and shrink as n^-1.5 beyond
// (a larger graph fills the GPU with a narrower bucket): ave_wt * 64 / ave_degree * min(1, (2^22/n)^1.5).
// For int graphs it is rounded to a power of two like the reference's shift.
static float default_delta(const adds_graph* g) {
  double ave_deg = g->n ? (double)g->m / (double)g->n : 1.0;
  if (ave_deg < 1) ave_deg = 1;
  const double size = g->n ? std::min(1.0, std::pow(4194304.0 / (double)g->n, 1.5)) : 1.0;
  double d = (double)g->ave_wt * 64.0 / ave_deg * size;
  if (g->wt == ADDS_W_U32) {
    double p = std::exp2(std::round(std::log2(std::max(d, 1.0))));
    return (float)p;
  }
  return (float)std::max(d, 1e-30);
}

// Default claim size: small graphs need more, smaller claims to spread their few items over the
// wavefronts; large ones fewer, larger claims (less contention on the queue counters).
static uint32_t default_chunk(uint64_t total_n) {
  return total_n <= (256u << 10) ? 64u : total_n <= (8u << 20) ? 128u : 256u;
}

static adds_status resolve(const adds_graph* g, const adds_params* p, uint64_t total_n, uint32_t pool_shift,
                           Resolved& r) {
  adds_params z;
  memset(&z, 0, sizeof(z));
  if (!p) p = &z;
  r.delta = p->delta > 0 ? p->delta : default_delta(g);
  r.light_delta = p->light_delta > 0 ? p->light_delta : r.delta;
  r.wg_per_cu = p->wg_per_cu ? p->wg_per_cu : 1;
  r.chunk = p->chunk ? p->chunk : default_chunk(total_n);
  r.nb = p->nbuckets ? p->nbuckets : 64;
  r.parts = p->parts ? p->parts : 8;
  r.hub_degree = p->hub_degree ? p->hub_degree : 1024;
  r.task_edges = p->task_edges ? p->task_edges : 64;
  r.concur = p->concurrency ? p->concurrency : 2;
  r.windows = p->edge_windows ? p->edge_windows : 4;
  r.tail_items = p->tail_items;
  r.flags = p->flags;
  r.timeout_ms = p->timeout_ms ? (double)p->timeout_ms : 120000.0;
  r.pool_fixed = p->pool_entries != 0;
  // default: 4 entries per vertex-instance + 1 per 16 edge-instances + 8M, doubled after each overflow
  // on this graph (remembered)
  const uint64_t total_m = g->n ? (total_n / g->n) * g->m : 0;
  r.pool_entries = p->pool_entries ? p->pool_entries
                                   : (4 * total_n + total_m / 16 + (8u << 20)) << std::min(pool_shift, 6u);
  if (!(r.delta > 0) || !std::isfinite(r.delta) || !(r.light_delta >= r.delta) || r.wg_per_cu > 64 || r.chunk > 256 ||
      (r.nb & (r.nb - 1)) || r.nb < 4 || r.nb > 256 || (r.parts != 1 && r.parts != 2 && r.parts != 4 &&
      r.parts != 8) || (r.concur + 3) * r.parts > MAX_PROBE_QUEUES || r.concur < 1 || r.concur > 16 ||
      r.task_edges == 0 || r.task_edges > 65536 || r.hub_degree < 64 || (r.windows != 4 && r.windows != 8)) {
    g_detail = "parameter out of range";
    return ADDS_EINVAL;
  }
  if (r.pool_entries < CH) r.pool_entries = CH;
  // Pull phases (symmetric graphs only): one per solve, for the first bucket whose heavy queues hold
  // at least max(total_n / 2, 2^20) heavy edges -- in practice bucket 0 of a power-law graph with a
  // high-degree source.  Every long row of every instance may become tasks of at least PULL_LONG edges.
  r.pull_slots = (g->symmetric && !(r.flags & ADDS_F_NO_PULL)) ? 1u : 0u;
  // no heavy edge (every weight at or below the light threshold): no pull can be asked for, so its
  // launches are left out
  if (g->max_w_known) {
    const bool all_light = g->wt == ADDS_W_F32 ? __builtin_bit_cast(float, g->max_w) <= r.light_delta
                                               : (double)g->max_w <= std::floor((double)r.light_delta);
    if (all_light) r.pull_slots = 0;
  }
  r.pull_delta = p->pull_delta;
  if (!(r.pull_delta >= 0.f) || !std::isfinite(r.pull_delta)) {
    g_detail = "parameter out of range";
    return ADDS_EINVAL;
  }
  r.pull_min = std::max<uint64_t>(total_n / 2, 1u << 20);
  r.total_n = total_n;
  // Split first bucket: the last bucket of width delta that no heavy edge can reach, i.e. one below the
  // bucket of the smallest heavy weight (the kernel's bucket key is monotone, and a heavy edge's target
  // distance is at least its weight).  Heavy: w > floor(light_delta) (int), w > light_delta (float).
  r.pull_at = PULL_AT_OFF;
  if ((r.flags & ADDS_F_PULL_SPLIT) && r.pull_slots) {
    float hmin;
    if (g->wt == ADDS_W_F32) {
      hmin = std::nextafter(r.light_delta, INFINITY);
    } else {
      const double lh = std::floor((double)r.light_delta) + 1.0;
      hmin = lh >= 4294967295.0 ? INFINITY : (float)(uint32_t)lh;
    }
    const float inv = 1.0f / r.delta;  // the kernel's inv_delta
    const float x = std::fmin(hmin * inv, 4.0e9f);
    const uint32_t key = std::isfinite(hmin) ? (uint32_t)x : 0u;
    if (key >= 1) r.pull_at = key - 1;
  }
  r.order_degree = r.pull_at != PULL_AT_OFF ? p->order_degree : 0u;
  // an early pull leaves items queued below the split point; the launches after it must relax every edge
  // of whatever they expand (all light, after a re-bucketing pull), and the queues' contents are carried
  r.pull_early = (r.pull_at != PULL_AT_OFF && r.pull_delta > 0.f && (r.flags & ADDS_F_PULL_ALL_LIGHT))
                     ? p->pull_early : 0u;
  r.ptask_cap = (uint32_t)std::min<uint64_t>(total_m / PULL_LONG + 4096, 1u << 24);
  return ADDS_OK;
}

static adds_status ensure_workspace(Workspace& ws, const Resolved& r, uint32_t nsrc) {
  const uint32_t nq = r.nb * r.parts;
  const uint64_t chunks64 = (r.pool_entries + CH - 1) / CH;
  if (chunks64 >= (1u << 24)) return ADDS_EINVAL;  // chunk ids are 24-bit in the kernel's chunk-map cache
  const uint32_t chunks = (uint32_t)chunks64;
  if (!ws.host_state) {
    HIPCHK(hipHostMalloc((void**)&ws.host_state, HOST_STATE_WORDS * 4, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&ws.host_abort, 256, hipHostMallocCoherent | hipHostMallocMapped));
    *ws.host_abort = 0;
    HIPCHK(hipEventCreate(&ws.ev0));
    HIPCHK(hipEventCreate(&ws.ev1));
    HIPCHK(hipEventCreate(&ws.ev2));
    HIPCHK(hipEventCreate(&ws.evk0));
    HIPCHK(hipEventCreate(&ws.evk1));
  }
  if (ws.nq != nq || ws.nb != r.nb || ws.pool_chunks < chunks || !ws.st) {
    // (re)allocate state, pool and chunk maps, then put them in their empty state
    HIPCHK(hipDeviceSynchronize());
    if (ws.st) (void)hipFree(ws.st);
    if (ws.pool) (void)hipFree(ws.pool);
    if (ws.tab) (void)hipFree(ws.tab);
    if (ws.hrec) (void)hipFree(ws.hrec);
    ws.st = nullptr;
    ws.pool = nullptr;
    ws.tab = nullptr;
    ws.hrec = nullptr;
    const uint32_t pc = std::max(chunks, ws.pool_chunks);
    HIPCHK(hipMalloc(&ws.st, (size_t)state_words(nq, r.nb) * 4));
    HIPCHK(hipMalloc(&ws.pool, (size_t)pc * CH * 8));
    HIPCHK(hipMalloc(&ws.tab, (size_t)2 * nq * pc * 4));
    // one range record per 16 pool entries (a record is only written for a range of > hub_degree edges)
    ws.hrec_cap = (uint32_t)std::min<uint64_t>((uint64_t)pc * CH / 16, REF_MAX_RECORDS);
    HIPCHK(hipMalloc(&ws.hrec, (size_t)ws.hrec_cap * HREC_WORDS * 4));
    HIPCHK(launch_fill_u32(ws.st, state_words(nq, r.nb), 0, nullptr));
    HIPCHK(launch_fill_u64(ws.pool, (uint64_t)pc * CH, EMPTY64, nullptr));
    HIPCHK(launch_fill_u32(ws.tab, (uint64_t)2 * nq * pc, TAB_UNSET, nullptr));
    HIPCHK(hipDeviceSynchronize());
    ws.nq = nq;
    ws.nb = r.nb;
    ws.pool_chunks = pc;
    ws.maxv = pc;
  }
  if (r.pull_slots && ws.ptask_cap < r.ptask_cap) {
    HIPCHK(hipDeviceSynchronize());
    if (ws.ptask) (void)hipFree(ws.ptask);
    ws.ptask = nullptr;
    ws.ptask_cap = 0;
    HIPCHK(hipMalloc(&ws.ptask, ((size_t)r.ptask_cap + CARRY_UINT4) * 16));  // then the carry ranges
    ws.ptask_cap = r.ptask_cap;
  }
  if (r.pull_slots && r.pull_delta > 0.f && ws.pstage_cap < r.total_n) {
    HIPCHK(hipDeviceSynchronize());
    if (ws.pstage) (void)hipFree(ws.pstage);
    if (ws.pcnt) (void)hipFree(ws.pcnt);
    ws.pstage = nullptr;
    ws.pcnt = nullptr;
    ws.pstage_cap = 0;
    HIPCHK(hipMalloc(&ws.pstage, (size_t)r.total_n * 8));
    HIPCHK(hipMalloc(&ws.pcnt, ((size_t)r.total_n / 256 + 1) * 4));  // the pull's chunks of 256 vertex-instances
    ws.pstage_cap = r.total_n;
  }
  if (ws.srcs_cap < nsrc) {
    if (ws.srcs) (void)hipFree(ws.srcs);
    ws.srcs = nullptr;
    HIPCHK(hipMalloc(&ws.srcs, (size_t)nsrc * 4));
    ws.srcs_cap = nsrc;
  }
  return ADDS_OK;
}

// Light-first rows for light threshold `delta` (computed once per graph and threshold; solves with
// another threshold re-split).
// The buffers are allocated with the graph; this launches kernels only.
static adds_status ensure_split(adds_graph* g, float delta, hipStream_t s) {
  if (g->split_delta != delta) {
    // heavy weight bands of width delta above delta (the pull phases read the same two values)
    HIPCHK(launch_split(g->row, g->col, g->w, g->n, delta, delta, 1.0f / delta, g->wt == ADDS_W_F32, g->scw,
                        g->lcnt, g->split_aux, g->rinfo, g->phead, s));
    g->split_delta = delta;
  }
  return ADDS_OK;
}

static uint32_t inf_bits(adds_wtype wt) { return wt == ADDS_W_F32 ? ADDS_INF_F32_BITS : ADDS_INF_U32; }

extern "C" adds_status adds_graph_split_rows(adds_graph* g, float light_delta, const void** cw,
                                             const uint32_t** light_counts) {
  if (!g || !(light_delta > 0.f)) return ADDS_EINVAL;
  HIPCHK(hipSetDevice(g->device));
  if (g->n) {
    const adds_status rc = ensure_split(g, light_delta, nullptr);
    if (rc != ADDS_OK) return rc;
    HIPCHK(hipStreamSynchronize(nullptr));
  }
  if (cw) *cw = g->scw;
  if (light_counts) *light_counts = g->lcnt;
  return ADDS_OK;
}

static KArgs make_args(const adds_graph* g, const Workspace& ws, const Resolved& r, uint32_t* dist, uint64_t total_n,
                       uint32_t grid) {
  KArgs a;
  memset(&a, 0, sizeof(a));
  a.cw = g->scw;
  a.rinfo = g->rinfo;
  a.phead = g->phead;
  a.dist = dist;
  a.st = ws.st;
  a.pool = ws.pool;
  a.tab = ws.tab;
  a.hrec = ws.hrec;
  a.hrec_cap = ws.hrec_cap;
  a.host_abort = ws.host_abort;
  a.n = g->n;
  a.total_n = (uint32_t)total_n;
  a.pool_chunks = (uint32_t)std::min<uint64_t>(ws.pool_chunks, (r.pool_entries + CH - 1) / CH);
  a.maxv = ws.maxv;
  a.nb = r.nb;
  a.parts = r.parts;
  a.inv_delta = 1.0f / r.delta;
  // replicated solves may place each replica's items in one part (ADDS_F_REPLICA_PARTS)
  a.inst_part = (total_n > g->n && r.parts > 1 && (r.flags & ADDS_F_REPLICA_PARTS)) ? 1u : 0u;
  a.n_magic = g->n > 1 ? (uint32_t)((((uint64_t)1 << 32) + g->n - 1) / g->n) : 0u;
  a.chunk = r.chunk;
  // guided claims: a wavefront takes its share of what is published among the wavefronts of one
  // part, at least claim_min items.  Measured on RMAT graphs, full claims (claim_min = chunk) are
  // fastest: fewer claims, less contention.
  a.claim_min = r.chunk;
  a.chunk_heavy = r.chunk;
  a.claim_div = std::max<uint32_t>(1, grid * WAVES_PER_WG / r.parts);
  a.hub_degree = r.hub_degree;
  a.task_edges = r.task_edges;
  a.windows = r.windows;
  a.tail_items = r.tail_items;
  a.concur = std::min(r.concur, r.nb - 2);
  a.total_waves = grid * WAVES_PER_WG;
  int rate_khz = 100000;
  (void)hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, g->device);
  if (rate_khz <= 0) rate_khz = 100000;
  a.timeout_ticks = (uint64_t)(r.timeout_ms * (double)rate_khz);
  a.cached_probe = (r.flags & ADDS_F_ATOMIC_PROBE) ? 0u : 1u;
  a.sticky = 1;
  a.local_cont = (r.flags & ADDS_F_NO_LOCAL) ? 0u : (r.flags & ADDS_F_LOCAL_SPARSE) ? 2u : 1u;
  a.full_first = (r.flags & ADDS_F_FULL_FIRST) ? 1u : 0u;
  a.nowait = (r.flags & ADDS_F_NO_WAIT) ? 1u : 0u;
  a.trace = ws.trace;
  a.trace_cap = TRACE_CAP;
  a.trace_slice = std::max(1, rate_khz / 20);  // 50 us bins
  // pull phases need the task list (the allocation-free solve has none) and the graph's symmetric word
  a.pull_slots = (ws.ptask && g->sym_dev) ? r.pull_slots : 0u;
  a.pull_min = r.pull_min;
  a.pull_eager = (r.flags & ADDS_F_PULL_EAGER) ? 1u : 0u;
  a.pull_ok = g->sym_dev;
  a.inv_delta_resume = r.pull_delta > 0.f ? 1.0f / r.pull_delta : 0.f;
  a.resume_all_light = (r.flags & ADDS_F_PULL_ALL_LIGHT) ? 1u : 0u;
  a.pull_at = r.pull_at;
  a.order_degree = r.order_degree;
  a.band_lo = g->split_delta;  // the split's bands (ensure_split has run for r.light_delta)
  a.band_inv = 1.0f / g->split_delta;
  a.ptask = ws.ptask;
  a.ptask_cap = ws.ptask_cap;
  a.carry = ws.ptask ? (uint2*)(ws.ptask + ws.ptask_cap) : nullptr;
  a.pstage = (ws.pstage && ws.pstage_cap >= total_n) ? ws.pstage : nullptr;
  a.pcnt = a.pstage ? ws.pcnt : nullptr;
  a.pull_early = r.pull_early;
  a.launch_idx = 0;
  return a;
}

static uint32_t grid_for(const Resolved& r) {
  uint32_t cu = adds_device_cu_count();
  if (cu == 0) cu = 1;
  return cu * r.wg_per_cu;
}

// The bucket timeline buffer when ADDS_TRACE names a file (eager solves only).
static adds_status ensure_trace(Workspace& ws) {
  if (!ws.trace && getenv("ADDS_TRACE"))
    HIPCHK(hipMalloc(&ws.trace, (4 + 4 * (size_t)TRACE_CAP + 8 * (size_t)TRACE_HBINS * TRACE_HCOPIES) * 8));
  return ADDS_OK;
}

// Enqueues dist fill + sources, the persistent solve and the cleanup on `s` (kernel launches only).
static adds_status enqueue_solve(const adds_graph* g, const Workspace& ws, const Resolved& r, uint32_t* dist,
                                 uint64_t total_n, const uint32_t* srcs_dev, uint32_t k, bool replicated,
                                 hipStream_t s, bool kernel_events) {
  const uint32_t grid = grid_for(r);
  KArgs a = make_args(g, ws, r, dist, total_n, grid);
  if (a.trace) {
    HIPCHK(launch_fill_u64(a.trace, 1, 0, s));
    HIPCHK(launch_fill_u64(a.trace + 4 + 4 * (size_t)TRACE_CAP, 8 * (size_t)TRACE_HBINS * TRACE_HCOPIES, 0, s));
  }
  HIPCHK(launch_init(a, srcs_dev, k, replicated ? 1 : 0, inf_bits(g->wt), s));
  if (kernel_events) HIPCHK(hipEventRecord(ws.evk0, s));
  HIPCHK(launch_solve(a, g->wt == ADDS_W_F32, replicated, grid, s));
  // Pull phase l and the launch that continues after it; both return at once unless the previous
  // launch asked for the pull (the sequence is fixed, so it is captured in HIP graphs as it is).
  for (uint32_t l = 1; l <= a.pull_slots; l++) {
    HIPCHK(launch_pull(a, g->wt == ADDS_W_F32, replicated, l, grid_for(r) / r.wg_per_cu * 4, s));
    KArgs b = a;
    b.launch_idx = l;
    HIPCHK(launch_solve(b, g->wt == ADDS_W_F32, replicated, grid, s));
  }
  if (kernel_events) HIPCHK(hipEventRecord(ws.evk1, s));
  HIPCHK(launch_cleanup(a, s));
  return ADDS_OK;
}

// Waits for ev1 with a host-side watchdog: past the device timeout (+5 s) the abort word is raised,
// which the persistent kernel polls.
static adds_status wait_solve(Workspace& ws) {
  auto t0 = std::chrono::steady_clock::now();
  bool aborted = false;
  for (;;) {
    hipError_t e = hipEventQuery(ws.ev2);
    if (e == hipSuccess) break;
    if (e != hipErrorNotReady) {
      g_detail = std::string("solve: ") + hipGetErrorString(e);
      return ADDS_EHIP;
    }
    double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (!aborted && el > ws.timeout_ms + 5000.0) {
      __atomic_store_n(ws.host_abort, 1u, __ATOMIC_SEQ_CST);
      aborted = true;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(el < 10 ? 5 : 200));
  }
  if (aborted) __atomic_store_n(ws.host_abort, 0u, __ATOMIC_SEQ_CST);
  return ADDS_OK;
}

// Appends the last solve's bucket timeline to the file named by ADDS_TRACE: one line per record,
// "t_us kind bucket prev_head pool_chunks" (kind 0 = start of a launch, 1 = heavy phase opened, 2 = head
// moved, 3 = done, 5 = pull requested (prev_head: heavy edges queued), 6 = pull phase decided (prev_head:
// 1 pulled, 2 pushed, 3 pulled and re-bucketed)), after a
// header line per solve.  A diagnostic aid; solves are unaffected apart from the records' writes.
static void dump_trace(const adds_graph* g, const Workspace& ws) {
  const char* path = getenv("ADDS_TRACE");
  if (!path) return;
  uint64_t cnt = 0;
  if (hipMemcpy(&cnt, ws.trace, 8, hipMemcpyDeviceToHost) != hipSuccess) return;
  cnt = std::min<uint64_t>(cnt, TRACE_CAP);
  std::vector<uint64_t> rec(4 * cnt);
  if (cnt && hipMemcpy(rec.data(), ws.trace + 4, cnt * 32, hipMemcpyDeviceToHost) != hipSuccess) return;
  FILE* f = fopen(path, "a");
  if (!f) return;
  int rate_khz = 100000;
  (void)hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, g->device);
  const double us_per_tick = 1000.0 / (rate_khz > 0 ? rate_khz : 100000);
  fprintf(f, "# solve n=%u records=%llu\n", g->n, (unsigned long long)cnt);
  const uint64_t t0 = cnt ? rec[0] : 0;
  for (uint64_t i = 0; i < cnt; i++)
    fprintf(f, "%.2f %llu %llu %llu %llu\n", (double)(rec[4 * i] - t0) * us_per_tick,
            (unsigned long long)(rec[4 * i + 1] >> 56), (unsigned long long)(rec[4 * i + 1] & 0xFFFFFFFFull),
            (unsigned long long)rec[4 * i + 2], (unsigned long long)rec[4 * i + 3]);
  // activity histogram: per 50 us bin of each wavefront's own clock, wavefront-us probing, on items,
  // on edges and idle, then edges relaxed and items expanded
  std::vector<uint64_t> hist(8 * (size_t)TRACE_HBINS * TRACE_HCOPIES);
  if (hipMemcpy(hist.data(), ws.trace + 4 + 4 * (size_t)TRACE_CAP, hist.size() * 8, hipMemcpyDeviceToHost) ==
      hipSuccess) {
    fprintf(f, "# hist bin_us=50: bin probe_us items_us edges_us idle_us edges items flush_us\n");
    for (uint32_t b = 0; b < TRACE_HBINS; b++) {
      uint64_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (uint32_t c = 0; c < TRACE_HCOPIES; c++)
        for (int i = 0; i < 8; i++) h[i] += hist[8 * ((size_t)c * TRACE_HBINS + b) + i];
      if (!(h[0] | h[1] | h[2] | h[3])) continue;
      fprintf(f, "h %u %.0f %.0f %.0f %.0f %llu %llu %.0f\n", b, h[0] * us_per_tick, h[1] * us_per_tick,
              h[2] * us_per_tick, h[3] * us_per_tick, (unsigned long long)h[4], (unsigned long long)h[5],
              h[6] * us_per_tick);
    }
  }
  fclose(f);
}

// The solve's counters and outcome from the host copy of the state words before the queues.  ms / kernel_ms are
// left to the caller.
static void read_stats(const uint32_t* hs, adds_stats* stats) {
  uint64_t sv[NSTATS] = {};
  for (uint32_t c = 0; c < STAT_COPIES; c++)
    for (uint32_t i = 0; i < NSTATS; i++) sv[i] += ((const uint64_t*)(hs + stat_copy(c)))[i];
  stats->vertices_expanded = sv[ST_EXPANDED];
  stats->edges_scanned = sv[ST_EDGES];
  stats->relaxations = sv[ST_RELAX];
  stats->pushes = sv[ST_PUSH];
  stats->stale_skips = sv[ST_STALE];
  stats->hub_tasks = sv[ST_TASKS];
  stats->rebuckets = sv[ST_REBUCKET];
  stats->pool_chunks_used = hs[S_POOL];
  stats->probes = sv[ST_PROBE];
  stats->cas_failures = sv[ST_CASFAIL];
  stats->idle_iterations = sv[ST_IDLE];
  stats->late_items = sv[ST_LATE];
  stats->ticks_probe = sv[ST_T_PROBE];
  stats->ticks_items = sv[ST_T_ITEMS];
  stats->ticks_edges = sv[ST_T_EDGES];
  stats->ticks_idle = sv[ST_T_IDLE];
  stats->local_items = sv[ST_LOCAL];
  stats->heavy_edges = sv[ST_HEDGES];
  stats->ticks_flush = sv[ST_T_FLUSH];
  stats->pulls = hs[PULL_DONE];
  stats->pull_edges = sv[ST_PULL_EDGES];
  stats->pull_improved = sv[ST_PULL_IMPROVED];
}

static adds_status solve_outcome(const uint32_t* hs, double timeout_ms) {
  const uint32_t err = hs[S_DONE + 1];
  if (err == DERR_OVERFLOW) {
    g_detail = "worklist pool exhausted (" + std::to_string(hs[S_POOL]) + " chunks requested)";
    return ADDS_EOVERFLOW;
  }
  if (err == DERR_TIMEOUT || err == DERR_ABORT) {
    g_detail = err == DERR_ABORT ? "solve aborted by the host watchdog"
                                 : "device watchdog ended the solve after " + std::to_string((long)timeout_ms) +
                                       " ms (or a bounded wait expired)";
    return ADDS_ETIMEOUT;
  }
  if (err != 0) {
    g_detail = "solver error code " + std::to_string(err);
    return ADDS_EHIP;
  }
  if (hs[S_DONE] != 1) {
    g_detail = "solver exited without completing";
    return ADDS_EHIP;
  }
  return ADDS_OK;
}

static adds_status collect(const adds_graph* g, Workspace& ws, adds_stats* stats) {
  adds_status rc = wait_solve(ws);
  ws.pending = false;
  if (rc != ADDS_OK) return rc;
  const uint32_t* hs = ws.host_state;
  if (stats) {
    read_stats(hs, stats);
    float ms = 0, kms = 0;
    (void)hipEventElapsedTime(&ms, ws.ev0, ws.ev1);
    if (ws.kernel_events) (void)hipEventElapsedTime(&kms, ws.evk0, ws.evk1);
    stats->ms = ms;
    stats->kernel_ms = kms;
  }
  if (ws.trace) dump_trace(g, ws);
  rc = solve_outcome(hs, ws.timeout_ms);
  if (rc == ADDS_EOVERFLOW && !ws.pool_fixed && ws.pool_shift < 6) ws.pool_shift++;
  return rc;
}

// One synchronous-or-not solve of k sources into dist (device, [k][n] when replicated).
static adds_status solve_core(adds_graph* g, const uint32_t* srcs_host, uint32_t k, uint32_t* dist_dev,
                              bool replicated, const adds_params* p, adds_stats* stats, hipStream_t s,
                              bool sync) {
  const uint64_t total_n = replicated ? (uint64_t)k * g->n : g->n;
  if (total_n >= 0xFFFFFFFFull) return ADDS_EINVAL;
  Resolved r;
  adds_status rc = resolve(g, p, total_n, g->ws.pool_shift, r);
  if (rc != ADDS_OK) return rc;
  for (uint32_t i = 0; i < k; i++)
    if (srcs_host[i] >= g->n) {
      g_detail = "source out of range";
      return ADDS_EINVAL;
    }
  if ((uint64_t)k > r.pool_entries) {
    g_detail = "more sources than worklist pool entries";
    return ADDS_EINVAL;
  }
  for (int attempt = 0; attempt < 4; attempt++) {
    Workspace& ws = g->ws;
    rc = ensure_workspace(ws, r, k);
    if (rc == ADDS_OK) rc = ensure_trace(ws);
    if (rc == ADDS_OK) rc = ensure_split(g, r.light_delta, s);
    if (rc != ADDS_OK) return rc;
    ws.timeout_ms = r.timeout_ms;
    ws.pool_fixed = r.pool_fixed;
    HIPCHK(hipMemcpyAsync(ws.srcs, srcs_host, (size_t)k * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(ws.ev0, s));
    rc = enqueue_solve(g, ws, r, dist_dev, total_n, ws.srcs, k, replicated, s, true);
    ws.kernel_events = true;
    if (rc != ADDS_OK) return rc;
    HIPCHK(hipEventRecord(ws.ev1, s));
    HIPCHK(hipMemcpyAsync(ws.host_state, ws.st, HOST_STATE_WORDS * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(ws.ev2, s));
    ws.last_stream = s;
    ws.pending = true;
    if (!sync) return ADDS_OK;
    rc = collect(g, ws, stats);
    if (rc != ADDS_EOVERFLOW || r.pool_fixed) return rc;
    r.pool_entries *= 2;  // grow the pool and solve again
  }
  return rc;
}

static adds_status ensure_outbuf(adds_graph* g, uint64_t words) {
  if (g->outbuf_cap >= words) return ADDS_OK;
  if (g->outbuf) (void)hipFree(g->outbuf);
  g->outbuf = nullptr;
  g->outbuf_cap = 0;
  HIPCHK(hipMalloc(&g->outbuf, words * 4));
  g->outbuf_cap = words;
  return ADDS_OK;
}

extern "C" adds_status adds_sssp_multi(adds_graph* g, const uint32_t* srcs, uint32_t k, void* dist,
                                       const adds_params* p, adds_stats* stats, void* stream) {
  if (!g || !srcs || k == 0 || !dist) return ADDS_EINVAL;
  HIPCHK(hipSetDevice(g->device));
  hipStream_t s = (hipStream_t)stream;
  const uint32_t flags = p ? p->flags : 0;
  const bool host_out = flags & ADDS_F_HOST_OUT;
  const bool sync = host_out || !(flags & ADDS_F_NO_SYNC);
  uint32_t* d = (uint32_t*)dist;
  if (host_out) {
    adds_status rc = ensure_outbuf(g, (uint64_t)k * g->n);
    if (rc != ADDS_OK) return rc;
    d = g->outbuf;
  }
  adds_status rc = solve_core(g, srcs, k, d, k > 1, p, stats, s, sync);
  if (rc != ADDS_OK) return rc;
  if (host_out) {
    HIPCHK(hipMemcpyAsync(dist, d, (size_t)k * g->n * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return ADDS_OK;
}

extern "C" adds_status adds_sssp(adds_graph* g, uint32_t src, void* dist, const adds_params* p, adds_stats* stats,
                                 void* stream) {
  return adds_sssp_multi(g, &src, 1, dist, p, stats, stream);
}

extern "C" adds_status adds_last_status(adds_graph* g, adds_stats* stats) {
  if (!g) return ADDS_EINVAL;
  if (!g->ws.pending) return ADDS_OK;
  HIPCHK(hipSetDevice(g->device));
  return collect(g, g->ws, stats);
}

// ---- batched over different graphs: block-diagonal union CSR ----
__global__ static void k_offset_copy(uint32_t* dst, const uint32_t* src, uint64_t count, uint32_t add) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i] + add;
}

static adds_status build_union(adds_graph* const* graphs, uint32_t k, adds_graph** out, std::vector<uint32_t>& off) {
  uint64_t N = 0, M = 0;
  off.assign(k + 1, 0);
  std::vector<uint64_t> eoff(k + 1, 0);
  for (uint32_t i = 0; i < k; i++) {
    if (graphs[i]->wt != graphs[0]->wt) return ADDS_EINVAL;
    // a graph repeated in the list gets its own copy so that every instance has its own vertices
    N += graphs[i]->n;
    M += graphs[i]->m;
    if (N >= 0xFFFFFFFFull || M >= 0xFFFFFFFFull) return ADDS_EINVAL;
    off[i + 1] = (uint32_t)N;
    eoff[i + 1] = M;
  }
  adds_graph* u = nullptr;
  adds_status rc = alloc_graph((uint32_t)N, M, graphs[0]->wt, &u);
  if (rc != ADDS_OK) return rc;
  double wsum = 0;
  for (uint32_t i = 0; i < k; i++) {
    const adds_graph* g = graphs[i];
    hipLaunchKernelGGL(k_offset_copy, dim3(256), dim3(256), 0, nullptr, u->row + off[i], g->row, (uint64_t)g->n,
                       (uint32_t)eoff[i]);
    if (g->m) {
      hipLaunchKernelGGL(k_offset_copy, dim3(1024), dim3(256), 0, nullptr, u->col + eoff[i], g->col, g->m, off[i]);
      HIPCHK(hipMemcpy(u->w + eoff[i], g->w, g->m * 4, hipMemcpyDeviceToDevice));
    }
    wsum += (double)g->ave_wt * (double)g->m;
  }
  const uint32_t last = (uint32_t)M;
  HIPCHK(hipMemcpy(u->row + N, &last, 4, hipMemcpyHostToDevice));
  HIPCHK(hipDeviceSynchronize());
  u->ave_wt = M ? (float)(wsum / (double)M) : 1.f;
  // block-diagonal: symmetric when every member is; its largest weight is known when every member's is
  bool sym = true, mk = true;
  for (uint32_t i = 0; i < k; i++) {
    sym &= graphs[i]->symmetric;
    mk &= graphs[i]->max_w_known;
    u->max_w = std::max(u->max_w, graphs[i]->max_w);
  }
  u->max_w_known = mk;
  if (sym) {
    u->symmetric = true;
    HIPCHK(launch_fill_u32(u->sym_dev, 1, 1u, nullptr));
    HIPCHK(hipDeviceSynchronize());
  }
  *out = u;
  return ADDS_OK;
}

// The block-diagonal union of `graphs` (cached on the first graph's handle) and the sources mapped into it.
static adds_status get_union(adds_graph* const* graphs, const uint32_t* srcs, uint32_t k, adds_graph** out,
                             std::vector<uint32_t>& usrc) {
  adds_graph* g0 = graphs[0];
  // keyed by (handle, generation): a member whose weights changed, or a new handle at the address of a
  // destroyed one, does not match
  std::vector<std::pair<const adds_graph*, uint64_t>> key(k);
  for (uint32_t i = 0; i < k; i++) key[i] = {graphs[i], graphs[i]->gen};
  bool hit = g0->union_g && g0->union_key == key && g0->union_off.size() == (size_t)k + 1;
  for (uint32_t i = 0; hit && i < k; i++) hit = g0->union_off[i + 1] - g0->union_off[i] == graphs[i]->n;
  if (!hit) {
    if (g0->union_g) {
      release_graph_memory(g0->union_g);
      delete g0->union_g;
      g0->union_g = nullptr;
    }
    adds_status rc = build_union(graphs, k, &g0->union_g, g0->union_off);
    if (rc != ADDS_OK) return rc;
    g0->union_key = key;
  }
  usrc.assign(k, 0);
  for (uint32_t i = 0; i < k; i++) {
    if (srcs[i] >= graphs[i]->n) return ADDS_EINVAL;
    usrc[i] = g0->union_off[i] + srcs[i];
  }
  *out = g0->union_g;
  return ADDS_OK;
}

extern "C" adds_status adds_sssp_batched(adds_graph* const* graphs, const uint32_t* srcs, uint32_t k,
                                         void* const* dists, const adds_params* p, adds_stats* stats, void* stream) {
  if (!graphs || !srcs || !dists || k == 0) return ADDS_EINVAL;
  bool same = true;
  for (uint32_t i = 0; i < k; i++) {
    if (!graphs[i] || !dists[i]) return ADDS_EINVAL;
    same &= graphs[i] == graphs[0];
  }
  adds_graph* g0 = graphs[0];
  HIPCHK(hipSetDevice(g0->device));
  hipStream_t s = (hipStream_t)stream;
  const uint32_t flags = p ? p->flags : 0;
  const bool host_out = flags & ADDS_F_HOST_OUT;
  const hipMemcpyKind kind = host_out ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (same) {
    // contiguous [k][n] solve, then one copy per instance
    adds_status rc = ensure_outbuf(g0, (uint64_t)k * g0->n);
    if (rc != ADDS_OK) return rc;
    adds_params q = p ? *p : adds_params{};
    q.flags &= ~(ADDS_F_HOST_OUT | ADDS_F_NO_SYNC);
    rc = solve_core(g0, srcs, k, g0->outbuf, k > 1, &q, stats, s, true);
    if (rc != ADDS_OK) return rc;
    for (uint32_t i = 0; i < k; i++)
      HIPCHK(hipMemcpyAsync(dists[i], g0->outbuf + (size_t)i * g0->n, (size_t)g0->n * 4, kind, s));
    HIPCHK(hipStreamSynchronize(s));
    return ADDS_OK;
  }
  std::vector<uint32_t> usrc;
  adds_graph* u = nullptr;
  adds_status rc = get_union(graphs, srcs, k, &u, usrc);
  if (rc != ADDS_OK) return rc;
  rc = ensure_outbuf(u, u->n);
  if (rc != ADDS_OK) return rc;
  adds_params q = p ? *p : adds_params{};
  q.flags &= ~(ADDS_F_HOST_OUT | ADDS_F_NO_SYNC);
  rc = solve_core(u, usrc.data(), k, u->outbuf, false, &q, stats, s, true);
  if (rc != ADDS_OK) return rc;
  for (uint32_t i = 0; i < k; i++)
    HIPCHK(hipMemcpyAsync(dists[i], u->outbuf + g0->union_off[i], (size_t)graphs[i]->n * 4, kind, s));
  HIPCHK(hipStreamSynchronize(s));
  return ADDS_OK;
}

// ---- HIP graph capture ----
// Copies each instance's rows of a union solve's output into the caller's buffers: tab[2i] is the
// destination, tab[2i+1] = (offset in the union << 32) | vertices.
__global__ static void k_scatter_union(const uint32_t* src, const uint64_t* tab, uint32_t k) {
  for (uint32_t i = blockIdx.y; i < k; i += gridDim.y) {
    uint32_t* dst = (uint32_t*)tab[2 * i];
    const uint32_t off = (uint32_t)(tab[2 * i + 1] >> 32), n = (uint32_t)tab[2 * i + 1];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) dst[j] = src[off + j];
  }
}

extern "C" void adds_solve_graph_destroy(adds_solve_graph* sg) {
  if (!sg) return;
  if (sg->g) (void)hipSetDevice(sg->g->device);
  (void)hipDeviceSynchronize();
  if (sg->exec) (void)hipGraphExecDestroy(sg->exec);
  if (sg->graph) (void)hipGraphDestroy(sg->graph);
  free_workspace(sg->ws);
  if (sg->srcs) (void)hipFree(sg->srcs);
  if (sg->scatter) (void)hipFree(sg->scatter);
  if (sg->owned) {
    release_graph_memory(sg->owned);
    delete sg->owned;
  }
  delete sg;
}

// Captures the solve of k sources (device array) on sg->g into sg's HIP graph, with sg's own workspace:
// dist fill + sources, the persistent kernel, the worklist cleanup, and, for a batched capture, the copy
// of every instance's distances out of the union's output.
static adds_status capture_solve(adds_solve_graph* sg, const adds_params* p, const uint32_t* srcs_dev, uint32_t k,
                                 bool repl, uint32_t* dist_dev) {
  adds_graph* g = sg->g;
  const uint64_t total_n = repl ? (uint64_t)k * g->n : g->n;
  if (total_n >= 0xFFFFFFFFull) return ADDS_EINVAL;
  Resolved r;
  adds_status rc = resolve(g, p, total_n, 0, r);
  if (rc != ADDS_OK) return rc;
  if ((uint64_t)k > r.pool_entries) {
    g_detail = "more sources than worklist pool entries";
    return ADDS_EINVAL;
  }
  rc = ensure_workspace(sg->ws, r, 0);
  if (rc == ADDS_OK) rc = ensure_split(g, r.light_delta, nullptr);
  if (rc != ADDS_OK) return rc;
  HIPCHK(hipDeviceSynchronize());
  sg->ws.timeout_ms = r.timeout_ms;
  sg->ws.pool_fixed = true;  // a replay cannot grow its pool: an overflow is reported by the status call
  sg->split_light = r.light_delta;
  hipStream_t s;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = enqueue_solve(g, sg->ws, r, dist_dev, total_n, srcs_dev, k, repl, s, false);
    if (rc == ADDS_OK && sg->scatter) {
      hipLaunchKernelGGL(k_scatter_union, dim3(64, std::min<uint32_t>(k, 1024)), dim3(256), 0, s, dist_dev,
                         (const uint64_t*)sg->scatter, k);
      if (hipGetLastError() != hipSuccess) rc = ADDS_EHIP;
    }
    hipError_t e2 = hipStreamEndCapture(s, &sg->graph);
    if (rc == ADDS_OK && e2 != hipSuccess) e = e2;
  }
  if (e == hipSuccess && rc == ADDS_OK) e = hipGraphInstantiate(&sg->exec, sg->graph, nullptr, nullptr, 0);
  (void)hipStreamDestroy(s);
  if (e != hipSuccess) {
    g_detail = std::string("graph capture: ") + hipGetErrorString(e);
    return ADDS_EHIP;
  }
  return rc;
}

extern "C" adds_status adds_solve_graph_create(adds_graph* g, const uint32_t* srcs_dev, uint32_t k, void* dist_dev,
                                               const adds_params* p, adds_solve_graph** out) {
  if (!g || !srcs_dev || !k || !dist_dev || !out) return ADDS_EINVAL;
  HIPCHK(hipSetDevice(g->device));
  adds_solve_graph* sg = new adds_solve_graph();
  sg->g = g;
  const adds_status rc = capture_solve(sg, p, srcs_dev, k, k > 1, (uint32_t*)dist_dev);
  if (rc != ADDS_OK) {
    adds_solve_graph_destroy(sg);
    return rc;
  }
  *out = sg;
  return ADDS_OK;
}

extern "C" adds_status adds_solve_graph_create_batched(adds_graph* const* graphs, const uint32_t* srcs, uint32_t k,
                                                       void* const* dists, const adds_params* p,
                                                       adds_solve_graph** out) {
  if (!graphs || !srcs || !dists || !k || !out) return ADDS_EINVAL;
  for (uint32_t i = 0; i < k; i++) {
    if (!graphs[i] || !dists[i]) return ADDS_EINVAL;
    if (srcs[i] >= graphs[i]->n) {
      g_detail = "source out of range";
      return ADDS_EINVAL;
    }
  }
  HIPCHK(hipSetDevice(graphs[0]->device));
  adds_solve_graph* sg = new adds_solve_graph();
  std::vector<uint32_t> off;
  adds_status rc = build_union(graphs, k, &sg->owned, off);
  if (rc == ADDS_OK) {
    sg->g = sg->owned;
    rc = ensure_outbuf(sg->owned, sg->owned->n);
  }
  if (rc == ADDS_OK) {
    std::vector<uint32_t> usrc(k);
    std::vector<uint64_t> tab(2 * (size_t)k);
    for (uint32_t i = 0; i < k; i++) {
      usrc[i] = off[i] + srcs[i];
      tab[2 * i] = (uint64_t)(uintptr_t)dists[i];
      tab[2 * i + 1] = ((uint64_t)off[i] << 32) | graphs[i]->n;
    }
    hipError_t e = hipMalloc(&sg->srcs, (size_t)k * 4);
    if (e == hipSuccess) e = hipMalloc(&sg->scatter, (size_t)k * 16);
    if (e == hipSuccess) e = hipMemcpy(sg->srcs, usrc.data(), (size_t)k * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(sg->scatter, tab.data(), (size_t)k * 16, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      g_detail = std::string("batched solve graph: ") + hipGetErrorString(e);
      rc = e == hipErrorOutOfMemory ? ADDS_ENOMEM : ADDS_EHIP;
    }
  }
  if (rc == ADDS_OK) rc = capture_solve(sg, p, sg->srcs, k, false, sg->owned->outbuf);
  if (rc != ADDS_OK) {
    adds_solve_graph_destroy(sg);
    return rc;
  }
  *out = sg;
  return ADDS_OK;
}

extern "C" adds_status adds_solve_graph_launch(adds_solve_graph* sg, void* stream) {
  if (!sg || !sg->exec) return ADDS_EINVAL;
  adds_graph* g = sg->g;
  HIPCHK(hipSetDevice(g->device));
  hipStream_t s = (hipStream_t)stream;
  Workspace& ws = sg->ws;
  // an eager solve with another light threshold may have re-split the rows since the capture
  const adds_status rc = ensure_split(g, sg->split_light, s);
  if (rc != ADDS_OK) return rc;
  HIPCHK(hipEventRecord(ws.ev0, s));
  ws.kernel_events = false;
  HIPCHK(hipGraphLaunch(sg->exec, s));
  HIPCHK(hipEventRecord(ws.ev1, s));
  HIPCHK(hipMemcpyAsync(ws.host_state, ws.st, HOST_STATE_WORDS * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(ws.ev2, s));
  ws.pending = true;
  return ADDS_OK;
}

extern "C" adds_status adds_solve_graph_status(adds_solve_graph* sg, adds_stats* stats) {
  if (!sg) return ADDS_EINVAL;
  if (!sg->ws.pending) return ADDS_OK;
  HIPCHK(hipSetDevice(sg->g->device));
  return collect(sg->g, sg->ws, stats);
}

// ---- backward ----
static adds_status ensure_transpose(adds_graph* g) {
  if (g->rrow) return ADDS_OK;
  const uint32_t n = g->n;
  const uint64_t m = g->m;
  std::vector<uint32_t> row(n + 1), col(m);
  HIPCHK(hipMemcpy(row.data(), g->row, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost));
  if (m) HIPCHK(hipMemcpy(col.data(), g->col, m * 4, hipMemcpyDeviceToHost));
  std::vector<uint32_t> rrow(n + 1, 0), reid(m), rsrc(m);
  for (uint64_t e = 0; e < m; e++) rrow[col[e] + 1]++;
  for (uint32_t v = 0; v < n; v++) rrow[v + 1] += rrow[v];
  std::vector<uint32_t> cur(rrow.begin(), rrow.end() - 1);
  for (uint32_t u = 0; u < n; u++)
    for (uint32_t e = row[u]; e < row[u + 1]; e++) {  // ascending e per destination row
      uint32_t p = cur[col[e]]++;
      reid[p] = e;
      rsrc[p] = u;
    }
  HIPCHK(hipMalloc(&g->rrow, ((size_t)n + 1) * 4));
  HIPCHK(hipMalloc(&g->reid, std::max<uint64_t>(m, 1) * 4));
  HIPCHK(hipMalloc(&g->rsrc, std::max<uint64_t>(m, 1) * 4));
  HIPCHK(hipMemcpy(g->rrow, rrow.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice));
  if (m) {
    HIPCHK(hipMemcpy(g->reid, reid.data(), m * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g->rsrc, rsrc.data(), m * 4, hipMemcpyHostToDevice));
  }
  return ADDS_OK;
}

// Backward workspace for k sources: P and C/Acc (8 bytes per vertex-instance each), cnt (4), the list
// of nonzero g entries (4 per 32 vertex-instances), GW (8 per edge), counters and sources.
struct BwdLayout {
  uint64_t p, c, cnt, list, segup, attach, part, gw, ctr, srcs, bytes;
  uint32_t list_cap;
};
static BwdLayout bwd_layout(uint32_t n, uint64_t m, uint32_t k) {
  auto up = [](uint64_t x) { return (x + 255) & ~255ull; };
  const uint64_t kn = (uint64_t)k * n;
  BwdLayout L;
  L.list_cap = (uint32_t)(kn / 32 + 1);
  L.p = 0;
  L.c = up(L.p + 8 * kn);
  L.cnt = up(L.c + 8 * kn);
  L.list = up(L.cnt + 4 * kn);
  L.segup = up(L.list + 4 * (uint64_t)L.list_cap);
  L.attach = up(L.segup + 4 * (uint64_t)L.list_cap);
  L.part = up(L.attach + 4 * (uint64_t)L.list_cap);
  L.gw = up(L.part + 32 * (uint64_t)BWD_MAX_PARENT_BLOCKS);
  L.ctr = up(L.gw + 8 * std::max<uint64_t>(m, 1));
  L.srcs = up(L.ctr + 64);
  L.bytes = up(L.srcs + 4 * (uint64_t)k);
  return L;
}

// Fixed-point scale of a batch: M = largest |g| = f 2^eM (0.5 <= f < 1), 2^eN >= nonzero count:
// S = 62 - eM - eN keeps every sum below 2^62.
static int grad_scale(uint32_t maxbits, uint64_t nnz) {
  float mx;
  memcpy(&mx, &maxbits, 4);
  int em = 0, en = 0;
  (void)frexpf(mx, &em);
  while ((1ull << en) < nnz) en++;
  return 62 - em - en;
}

extern "C" adds_status adds_sssp_backward_f32(adds_graph* g, const uint32_t* srcs, uint32_t k, const float* dist,
                                              const float* grad_dist, float* grad_w, double* ms, void* stream) {
  if (!g || !srcs || !k || !dist || !grad_dist || !grad_w) return ADDS_EINVAL;
  if (g->wt != ADDS_W_F32) {
    g_detail = "backward needs a float graph";
    return ADDS_EINVAL;
  }
  if ((uint64_t)k * g->n >= 0xFFFFFFFFull) return ADDS_EINVAL;
  for (uint32_t i = 0; i < k; i++)
    if (srcs[i] >= g->n) {
      g_detail = "source out of range";
      return ADDS_EINVAL;
    }
  HIPCHK(hipSetDevice(g->device));
  hipStream_t s = (hipStream_t)stream;
  adds_status rc = ensure_transpose(g);
  if (rc != ADDS_OK) return rc;
  if (!g->rw) HIPCHK(hipMalloc(&g->rw, std::max<uint64_t>(g->m, 1) * 4));
  if (g->rw_gen != g->gen) {
    HIPCHK(launch_bwd_tweights(g->reid, (const float*)g->w, g->rw, g->m, s));
    g->rw_gen = g->gen;
  }
  const BwdLayout L = bwd_layout(g->n, g->m, k);
  if (g->bwd_bytes < L.bytes) {
    HIPCHK(hipStreamSynchronize(s));
    if (g->bwd) (void)hipFree(g->bwd);
    g->bwd = nullptr;
    g->bwd_bytes = 0;
    HIPCHK(hipMalloc(&g->bwd, L.bytes));
    g->bwd_bytes = L.bytes;
  }
  if (!g->bwd_host) HIPCHK(hipHostMalloc((void**)&g->bwd_host, 64, hipHostMallocDefault));
  const uint64_t kn = (uint64_t)k * g->n;
  BwdArgs b;
  memset(&b, 0, sizeof(b));
  b.rrow = g->rrow;
  b.reid = g->reid;
  b.rsrc = g->rsrc;
  b.rw = g->rw;
  b.dist = dist;
  b.gdist = grad_dist;
  b.srcs = (const uint32_t*)(g->bwd + L.srcs);
  b.P = (uint2*)(g->bwd + L.p);
  b.C = (uint2*)(g->bwd + L.c);
  b.Acc = (long long*)(g->bwd + L.c);  // the dense form runs after the rounds: C is free by then
  b.cnt = (uint32_t*)(g->bwd + L.cnt);
  b.list = (uint32_t*)(g->bwd + L.list);
  b.list_cap = L.list_cap;
  b.owner = b.cnt;  // the sparse form's owner words use the dense form's counters
  b.seg_up = (uint32_t*)(g->bwd + L.segup);
  b.attach = (uint32_t*)(g->bwd + L.attach);
  b.part = (unsigned long long*)(g->bwd + L.part);
  b.GW = (unsigned long long*)(g->bwd + L.gw);
  b.gw = grad_w;
  b.ctr = (unsigned long long*)(g->bwd + L.ctr);
  b.n = g->n;
  b.k = k;
  b.m = g->m;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto counters = [&]() -> adds_status {
    HIPCHK(hipMemcpyAsync(g->bwd_host, b.ctr, 48, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return ADDS_OK;
  };
  auto finish = [&](adds_status r) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return r;
  };
  HIPCHK(hipMemcpyAsync((void*)b.srcs, srcs, (size_t)k * 4, hipMemcpyHostToDevice, s));
  HIPCHK(hipEventRecord(e0, s));
  HIPCHK(launch_fill_u64((uint64_t*)b.ctr, 8, 0, s));
  HIPCHK(launch_fill_u64((uint64_t*)b.GW, g->m, 0, s));
  HIPCHK(launch_bwd_parent(b, s));
  rc = counters();
  if (rc != ADDS_OK) return finish(rc);
  if (g->bwd_host[3]) {
    g_detail = "grad_dist holds a non-finite value at a reached vertex";
    return finish(ADDS_EINVAL);
  }
  // equal-distance rounds, while vertices are left without a parent
  for (uint64_t unresolved = g->bwd_host[0]; unresolved;) {
    HIPCHK(launch_fill_u64((uint64_t*)b.ctr + 5, 1, 0, s));
    HIPCHK(launch_bwd_round(b, s));
    rc = counters();
    if (rc != ADDS_OK) return finish(rc);
    if (g->bwd_host[5] == 0 || g->bwd_host[5] > unresolved) {
      g_detail = "dist is not a shortest-path solution from srcs (a reached vertex has no tight path)";
      return finish(ADDS_EINVAL);
    }
    unresolved -= g->bwd_host[5];
  }
  const uint64_t nnz = g->bwd_host[1], listed = g->bwd_host[4];
  b.S = nnz ? grad_scale((uint32_t)g->bwd_host[2], nnz) : 0;
  if (nnz) {
    if (listed <= b.list_cap) {
      HIPCHK(launch_fill_u32(b.owner, kn, 0xFFFFFFFFu, s));
      HIPCHK(launch_fill_u64((uint64_t*)b.Acc, kn, 0, s));
      HIPCHK(launch_bwd_sparse(b, (uint32_t)listed, s));
    } else {
      HIPCHK(launch_fill_u64((uint64_t*)b.Acc, kn, 0, s));
      HIPCHK(launch_fill_u32(b.cnt, kn, 0, s));
      HIPCHK(launch_bwd_climb(b, s));
    }
  }
  HIPCHK(launch_bwd_convert(b, s));
  HIPCHK(hipEventRecord(e1, s));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0;
  (void)hipEventElapsedTime(&t, e0, e1);
  if (ms) *ms = t;
  return finish(ADDS_OK);
}

// ---- autotuner ----
// Coordinate descent over the solver's tuning space for one workload (graph + sources), timed with the
// persistent-kernel events.  Every candidate's distances are compared word for word with the first
// candidate's on the device: a tuning point never changes results, and a mismatch is an error.
__global__ static void k_count_diff(const uint32_t* a, const uint32_t* b, uint64_t n, unsigned long long* out) {
  unsigned long long c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    c += a[i] != b[i];
  if (c) atomicAdd(out, c);
}

static adds_status autotune_core(adds_graph* g, const uint32_t* srcs, uint32_t k, bool repl, uint32_t trials,
                                 adds_params* best, double* best_ms) {
  HIPCHK(hipSetDevice(g->device));
  if (trials == 0) trials = 3;
  const uint64_t words = (repl ? (uint64_t)k : 1) * g->n;
  uint32_t *ref = nullptr, *cur = nullptr;
  unsigned long long* diff = nullptr;
  HIPCHK(hipMalloc(&ref, words * 4));
  HIPCHK(hipMalloc(&cur, words * 4));
  HIPCHK(hipMalloc(&diff, 8));
  adds_status rc = ADDS_OK;
  bool have_ref = false;
  auto measure = [&](const adds_params& p, double& ms) -> adds_status {
    std::vector<double> t;
    for (uint32_t i = 0; i < trials + 1; i++) {  // first run warms the workspace for this shape
      adds_stats st;
      adds_status r = solve_core(g, srcs, k, have_ref ? cur : ref, repl, &p, &st, nullptr, true);
      if (r != ADDS_OK) return r;
      if (i) t.push_back(st.kernel_ms);
      if (!have_ref) {
        have_ref = true;
        continue;
      }
      HIPCHK(hipMemset(diff, 0, 8));
      hipLaunchKernelGGL(k_count_diff, dim3(1024), dim3(256), 0, nullptr, ref, cur, words, diff);
      unsigned long long nd = 0;
      HIPCHK(hipMemcpy(&nd, diff, 8, hipMemcpyDeviceToHost));
      if (nd) {
        // name the point and the first differing words (index: first point's word / this point's)
        std::vector<uint32_t> hr(words), hc(words);
        std::string where;
        if (hipMemcpy(hr.data(), ref, words * 4, hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(hc.data(), cur, words * 4, hipMemcpyDeviceToHost) == hipSuccess)
          for (uint64_t j = 0, shown = 0; j < words && shown < 4; j++)
            if (hr[j] != hc[j]) {
              where += " " + std::to_string(j) + ":" + std::to_string(hr[j]) + "/" + std::to_string(hc[j]);
              shown++;
            }
        char pt[256];
        snprintf(pt, sizeof(pt), "delta=%g light=%g pull_delta=%g order=%u nb=%u parts=%u concur=%u chunk=%u tail=%u flags=%u",
                 p.delta, p.light_delta, p.pull_delta, p.order_degree, p.nbuckets, p.parts, p.concurrency, p.chunk,
                 p.tail_items, p.flags);
        g_detail = "autotuner: distances differ between tuning points (" + std::to_string(nd) + " words, run " +
                   std::to_string(i) + " of " + pt + ";" + where + ")";
        return ADDS_EHIP;
      }
    }
    std::sort(t.begin(), t.end());
    ms = t[t.size() / 2];
    return ADDS_OK;
  };
  adds_params p;
  memset(&p, 0, sizeof(p));
  Resolved r0;
  rc = resolve(g, &p, words, g->ws.pool_shift, r0);  // words = vertex-instances of the solve
  if (rc == ADDS_OK) {
    p.delta = r0.delta;
    p.wg_per_cu = r0.wg_per_cu;
    p.chunk = r0.chunk;
    p.nbuckets = r0.nb;
    p.parts = r0.parts;
    p.concurrency = r0.concur;
    double bms = 0;
    rc = measure(p, bms);
    if (getenv("ADDS_TUNE_VERBOSE"))
      fprintf(stderr, "[tune] default delta=%g -> %.3f ms (%s)\n", p.delta, bms, adds_strerror(rc));
    const bool is_int = g->wt == ADDS_W_U32;
    std::vector<std::pair<double, adds_params>> tried;  // every point that ran, for the final playoff
    tried.push_back({bms, p});
    auto try_point = [&](adds_params q) {
      if (rc != ADDS_OK) return;
      double ms = 0;
      // a point far slower than the best so far is cut off by the device watchdog and skipped
      q.timeout_ms = (uint32_t)std::min(120000.0, std::max(2000.0, 20.0 * bms));
      adds_status r = measure(q, ms);
      q.timeout_ms = 0;
      // points that need more memory than planned, or that are cut off, are skipped
      if (r == ADDS_EOVERFLOW || r == ADDS_ETIMEOUT) {
        if (getenv("ADDS_TUNE_VERBOSE"))
          fprintf(stderr, "[tune] delta=%g nb=%u concur=%u wg/cu=%u chunk=%u -> skipped (%s)\n", q.delta,
                  q.nbuckets, q.concurrency, q.wg_per_cu, q.chunk, adds_strerror(r));
        return;
      }
      if (r != ADDS_OK) {
        rc = r;
        return;
      }
      if (getenv("ADDS_TUNE_VERBOSE"))
        fprintf(stderr, "[tune] delta=%g light=%g pull_delta=%g order=%u nb=%u parts=%u concur=%u wg/cu=%u chunk=%u task_edges=%u hub=%u windows=%u tail=%u flags=%u -> %.3f ms\n",
                q.delta, q.light_delta, q.pull_delta, q.order_degree, q.nbuckets, q.parts, q.concurrency, q.wg_per_cu,
                q.chunk, q.task_edges, q.hub_degree, q.edge_windows, q.tail_items, q.flags, ms);
      tried.push_back({ms, q});
      if (ms < bms) {
        bms = ms;
        p = q;
      }
    };
    // 1) bucket width: powers of two around the default (int keeps power-of-two widths), each also with
    //    every edge light: many sources on one graph do best with narrow buckets and no heavy phases, a
    //    corner that one change at a time from the default does not reliably reach
    const float d0 = p.delta;
    for (int j = -3; j <= 3 && rc == ADDS_OK; j++) {
      adds_params q = p;
      q.delta = d0 * std::ldexp(1.0f, j);
      if (is_int && q.delta < 1.0f) continue;
      if (j != 0) try_point(q);
      q.light_delta = 1e30f;
      try_point(q);
    }
    // pull phases (symmetric graphs): a wider bucket for the sparse rest of the solve after the pull
    if (g->symmetric)
      for (float f : {16.0f, 64.0f}) {
        adds_params q = p;
        q.pull_delta = p.delta * f;
        try_point(q);
      }
    // light threshold above the bucket width: all edges light (no heavy phases) pays where vertices are
    // rarely expanded twice (many sources on one graph)
    for (float ld : {4.0f * p.delta, 1e30f}) {
      adds_params q = p;
      q.light_delta = ld;
      try_point(q);
    }
    // 2) bucket ring length and open buckets
    for (uint32_t nb : {32u, 128u, 256u}) {
      adds_params q = p;
      q.nbuckets = nb;
      try_point(q);
    }
    // (sub-queues per bucket, open buckets): fewer parts leave probe lanes for more open buckets
    const uint32_t pc_pairs[][2] = {{8, 1}, {8, 2}, {8, 3}, {4, 2}, {4, 3}, {4, 4}, {2, 4}};
    const adds_params p2 = p;
    for (const auto& pc : pc_pairs) {
      adds_params q = p2;
      q.parts = pc[0];
      q.concurrency = pc[1];
      if (q.parts == p2.parts && q.concurrency == p2.concurrency) continue;
      try_point(q);
    }
    // 3) persistent workgroups per CU (more than one has never won on gfx950), then the claim size
    for (uint32_t w : {2u}) {
      adds_params q = p;
      q.wg_per_cu = w;
      try_point(q);
    }
    for (uint32_t c : {64u, 128u}) {
      adds_params q = p;
      q.chunk = c;
      try_point(q);
    }
    // 4) edges per chunk of a split range
    for (uint32_t te : {128u, 256u}) {
      adds_params q = p;
      q.task_edges = te;
      try_point(q);
    }
    // longer ranges relaxed by the wavefront that meets them (fewer references)
    for (uint32_t hd : {4096u, 16384u}) {
      adds_params q = p;
      q.hub_degree = hd;
      try_point(q);
    }
    // 5) more edge windows in flight
    {
      adds_params q = p;
      q.edge_windows = 8;
      try_point(q);
    }
    // 6) the sparse tail as one bucket
    for (uint32_t t : {4096u, 65536u}) {
      adds_params q = p;
      q.tail_items = t;
      try_point(q);
    }
    // 7) local continuation off; flushes that do not wait for their atomicMins
    {
      adds_params q = p;
      q.flags |= ADDS_F_NO_LOCAL;
      try_point(q);
    }
    {
      adds_params q = p;
      q.flags |= ADDS_F_NO_WAIT;
      try_point(q);
    }
    if (repl) {
      adds_params q = p;
      q.flags |= ADDS_F_REPLICA_PARTS;
      try_point(q);
    }
    // 8) a second, shorter pass over the choices that interact most (the best bucket width depends on
    //    how long ranges are split and on the tail, and so on), from the point reached so far
    for (float f : {0.5f, 2.0f}) {
      adds_params q = p;
      q.delta = p.delta * f;
      if (is_int && q.delta < 1.0f) continue;
      if (q.light_delta != 0.0f && q.light_delta < q.delta) q.light_delta = 0.0f;
      try_point(q);
    }
    // with a pull the best bucket width before it is narrower (its rest is handled at pull_delta), and the
    // rest may do best with every edge light
    if (g->symmetric) {
      for (float f : {0.0f, 8.0f, 32.0f, 64.0f}) {
        adds_params q = p;
        q.pull_delta = p.delta * f;
        if (q.pull_delta == p.pull_delta) continue;
        try_point(q);
      }
      if (p.pull_delta > 0.f) {
        adds_params q = p;
        q.flags ^= ADDS_F_PULL_ALL_LIGHT;
        try_point(q);
      }
    }
    // split first bucket (symmetric graphs): the light threshold kept, narrower buckets below it, all
    // their heavy edges pulled at once, optionally only the vertices of many light edges in bucket order;
    // then more open buckets and local continuation for the sparse stretches
    if (g->symmetric) {
      const adds_params p4 = p;
      const float lt = p4.light_delta > 0.f ? p4.light_delta : p4.delta;
      if (lt < 1e29f && !(p4.flags & ADDS_F_NO_PULL)) {
        // (the light threshold itself, and half of it: the pull then comes earlier, where the banded scan
        // makes it cheap)
        for (float lf : {1.0f, 0.5f})
          for (float div : {1.0f, 8.0f, 12.0f, 16.0f})
            for (uint32_t od : {0u, 12u, 16u, 32u}) {
              if (div == 1.0f && od) continue;
              adds_params q = p4;
              q.flags |= ADDS_F_PULL_SPLIT;
              q.light_delta = lt * lf;
              q.delta = q.light_delta / div;
              if (is_int && q.delta < 1.0f) continue;
              q.order_degree = od;
              try_point(q);
            }
        if (p.flags & ADDS_F_PULL_SPLIT) {
          // more open buckets: the narrow buckets' hubs are expanded ahead of the head, and the head's
          // walk through the narrow buckets blocks less (4 parts leave probe lanes for 8 buckets)
          const adds_params p5 = p;
          const uint32_t pc_open[][2] = {{8, 4}, {4, 4}, {4, 8}, {4, 12}};
          for (const auto& pc : pc_open) {
            adds_params q = p5;
            q.parts = pc[0];
            q.concurrency = pc[1];
            if (q.parts == p5.parts && q.concurrency == p5.concurrency) continue;
            if ((q.concurrency + 3) * q.parts > MAX_PROBE_QUEUES) continue;
            try_point(q);
          }
          {
            adds_params q = p;
            q.flags = (q.flags & ~ADDS_F_NO_LOCAL) | ADDS_F_LOCAL_SPARSE;
            try_point(q);
          }
          // an early pull (its queued items carried over), with every edge light after it
          if (p.pull_delta > 0.f && (p.flags & ADDS_F_PULL_ALL_LIGHT)) {
            const adds_params p6 = p;
            for (uint32_t e : {65536u, 262144u}) {
              adds_params q = p6;
              q.pull_early = e;
              try_point(q);
            }
          }
        }
      }
    }
    for (uint32_t hd : {0u, 4096u, 16384u}) {
      adds_params q = p;
      if (q.hub_degree == hd) continue;
      q.hub_degree = hd;
      try_point(q);
    }
    {
      adds_params q = p;
      q.edge_windows = q.edge_windows == 8 ? 4 : 8;
      try_point(q);
    }
    // the light threshold and the tail together: they interact (on a grid with many sources a
    // threshold of 4-8 bucket widths pays only with a long tail; measured 26 ms against 30 ms for
    // either change alone), and the first pass changes one at a time
    {
      const adds_params p3 = p;
      for (float ld : {0.0f, 4.0f * p3.delta, 8.0f * p3.delta, 1e30f})
        for (uint32_t t : {0u, 4096u, 65536u}) {
          adds_params q = p3;
          if (q.light_delta == ld && q.tail_items == t) continue;
          q.light_delta = ld;
          q.tail_items = t;
          try_point(q);
        }
    }
    // the two flags again: the first pass can settle on a point from which the no-wait flush or local
    // continuation looks worse than it is
    for (uint32_t f : {ADDS_F_NO_LOCAL, ADDS_F_NO_WAIT, ADDS_F_REPLICA_PARTS, ADDS_F_FULL_FIRST}) {
      if (f == ADDS_F_REPLICA_PARTS && !repl) continue;
      adds_params q = p;
      q.flags ^= f;
      try_point(q);
    }
    // 9) playoff: single timings differ by a few percent from run to run, about as much as the best
    //    points differ, so the three fastest points are timed again with twice the trials
    if (rc == ADDS_OK && tried.size() > 1) {
      std::stable_sort(tried.begin(), tried.end(),
                       [](const std::pair<double, adds_params>& x, const std::pair<double, adds_params>& y) {
                         return x.first < y.first;
                       });
      const uint32_t t0 = trials;
      trials = 2 * t0;
      double pbest = 0;
      for (size_t i = 0; i < std::min<size_t>(3, tried.size()) && rc == ADDS_OK; i++) {
        double ms = 0;
        adds_params q = tried[i].second;
        q.timeout_ms = 0;
        const adds_status r = measure(q, ms);
        if (r == ADDS_EOVERFLOW || r == ADDS_ETIMEOUT) continue;  // not picked
        if (r != ADDS_OK) {
          rc = r;
          break;
        }
        if (getenv("ADDS_TUNE_VERBOSE")) fprintf(stderr, "[tune] playoff %zu -> %.3f ms\n", i, ms);
        if (pbest == 0 || ms < pbest) {
          pbest = ms;
          p = q;
        }
      }
      trials = t0;
      if (pbest > 0) bms = pbest;
    }
    *best = p;
    if (best_ms) *best_ms = bms;
  }
  (void)hipFree(ref);
  (void)hipFree(cur);
  (void)hipFree(diff);
  return rc;
}

extern "C" adds_status adds_autotune(adds_graph* g, const uint32_t* srcs, uint32_t k, uint32_t trials,
                                     adds_params* best, double* best_ms) {
  if (!g || !srcs || !k || !best) return ADDS_EINVAL;
  return autotune_core(g, srcs, k, k > 1, trials, best, best_ms);
}

extern "C" adds_status adds_autotune_batched(adds_graph* const* graphs, const uint32_t* srcs, uint32_t k,
                                             uint32_t trials, adds_params* best, double* best_ms) {
  if (!graphs || !srcs || !k || !best) return ADDS_EINVAL;
  for (uint32_t i = 0; i < k; i++)
    if (!graphs[i]) return ADDS_EINVAL;
  HIPCHK(hipSetDevice(graphs[0]->device));
  std::vector<uint32_t> usrc;
  adds_graph* u = nullptr;
  adds_status rc = get_union(graphs, srcs, k, &u, usrc);
  if (rc != ADDS_OK) return rc;
  return autotune_core(u, usrc.data(), k, false, trials, best, best_ms);
}

// ---- allocation-free single-source solve ----
// adds_sssp_noalloc runs adds_sssp's kernels with their workspace -- state words, queue pool, chunk maps,
// range records, the source word -- in `scratch`, a file-scope device array of a fixed size that every
// GPU the library runs on holds until the process ends.  In host code `scratch` names a host-side stand-in,
// not device memory: its device address is taken once per device with hipGetSymbolAddress (with that
// device current) and only that address is used afterwards.  Every call on a device shares its array, so
// calls that overlap on one device race; each call leaves the pool and chunk maps empty again (the cleanup
// kernel) and clears the state words at its start (the init kernel).
constexpr uint64_t NA_POOL_ENTRIES = 8ull << 20;                 // 64 MiB of queue entries
constexpr uint32_t NA_CHUNKS = (uint32_t)(NA_POOL_ENTRIES / CH);  // chunk-map entries per queue
constexpr uint32_t NA_MAX_QUEUES = 512;                           // nbuckets * parts
constexpr uint64_t NA_STATE_WORDS = 1ull << 19;                   // 2 MiB
constexpr uint32_t NA_HREC = NA_CHUNKS * CH / 16;                 // range records (one per 16 entries)
constexpr uint64_t NA_OFF_POOL = 0;
constexpr uint64_t NA_OFF_TAB = NA_OFF_POOL + NA_POOL_ENTRIES * 8;
constexpr uint64_t NA_OFF_HREC = NA_OFF_TAB + (uint64_t)2 * NA_MAX_QUEUES * NA_CHUNKS * 4;
constexpr uint64_t NA_OFF_STATE = NA_OFF_HREC + (uint64_t)NA_HREC * HREC_WORDS * 4;
constexpr uint64_t NA_OFF_SRC = NA_OFF_STATE + NA_STATE_WORDS * 4;
constexpr uint64_t NA_BYTES = NA_OFF_SRC + 256;  // 98 MiB + 256 bytes
static_assert(NA_OFF_STATE % 256 == 0 && NA_OFF_SRC % 256 == 0, "scratch regions are 256-byte aligned");

static __device__ __align__(256) uint64_t scratch[NA_BYTES / 8];

constexpr int NA_MAX_DEVICES = 64;
static std::mutex g_na_mu;
static uint8_t* g_na_ptr[NA_MAX_DEVICES];  // device address of `scratch` per device, once set up
static bool g_na_clean[NA_MAX_DEVICES];    // the pool and chunk maps are in their empty state

// Called with `dev` current.  Takes the device address of `scratch` (first call on the device) and puts
// the pool and chunk maps in their empty state when a previous call may have left them otherwise (first
// call, or a call that failed before its cleanup kernel was enqueued).  Fill kernels, no memset nodes.
static adds_status noalloc_setup(int dev, hipStream_t s, uint8_t** base) {
  if (dev < 0 || dev >= NA_MAX_DEVICES) {
    g_detail = "allocation-free solve: device index out of range";
    return ADDS_EINVAL;
  }
  if (!g_na_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      g_detail = std::string("allocation-free solve: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return ADDS_EHIP;
    }
    g_na_ptr[dev] = (uint8_t*)ptr;
    g_na_clean[dev] = false;
  }
  uint8_t* b = g_na_ptr[dev];
  if (!g_na_clean[dev]) {
    HIPCHK(launch_fill_u64((uint64_t*)(b + NA_OFF_POOL), NA_POOL_ENTRIES, EMPTY64, s));
    HIPCHK(launch_fill_u32((uint32_t*)(b + NA_OFF_TAB), (uint64_t)2 * NA_MAX_QUEUES * NA_CHUNKS, TAB_UNSET, s));
    g_na_clean[dev] = true;
  }
  *base = b;
  return ADDS_OK;
}

extern "C" adds_status adds_sssp_noalloc(adds_graph* g, uint32_t src, void* dist, const adds_params* p,
                                         adds_stats* stats, void* stream) {
  if (!g || !dist) return ADDS_EINVAL;
  if (p && (p->flags & (ADDS_F_HOST_OUT | ADDS_F_NO_SYNC))) {
    g_detail = "allocation-free solve: dist must be device memory and the call is synchronous";
    return ADDS_EINVAL;
  }
  if (src >= g->n) {
    g_detail = "source out of range";
    return ADDS_EINVAL;
  }
  HIPCHK(hipSetDevice(g->device));
  Resolved r;
  adds_status rc = resolve(g, p, g->n, 0, r);
  if (rc != ADDS_OK) return rc;
  // the sizes the scratch holds: a worklist of 4 entries per vertex + 1 per 16 edges (the default rule
  // without its slack, or the caller's pool_entries) and at most NA_MAX_QUEUES queues per phase
  const uint64_t need = p && p->pool_entries ? p->pool_entries : 4ull * g->n + g->m / 16;
  const uint32_t nq = r.nb * r.parts;
  if (need > NA_POOL_ENTRIES || nq > NA_MAX_QUEUES || state_words(nq, r.nb) > NA_STATE_WORDS) {
    g_detail = "graph or parameters too large for the allocation-free solve: it needs " + std::to_string(need) +
               " worklist entries and " + std::to_string(nq) + " queues per phase; the scratch holds " +
               std::to_string(NA_POOL_ENTRIES) + " and " + std::to_string(NA_MAX_QUEUES) + " (use adds_sssp)";
    return ADDS_EINVAL;
  }
  r.pool_entries = NA_POOL_ENTRIES;
  hipStream_t s = (hipStream_t)stream;
  int dev = g->device;
  std::lock_guard<std::mutex> lk(g_na_mu);
  uint8_t* b = nullptr;
  rc = noalloc_setup(dev, s, &b);
  if (rc != ADDS_OK) return rc;
  Workspace ws;  // views into the scratch; nothing here is allocated
  ws.st = (uint32_t*)(b + NA_OFF_STATE);
  ws.pool = (uint64_t*)(b + NA_OFF_POOL);
  ws.tab = (uint32_t*)(b + NA_OFF_TAB);
  ws.hrec = (uint32_t*)(b + NA_OFF_HREC);
  ws.hrec_cap = NA_HREC;
  ws.nq = nq;
  ws.nb = r.nb;
  ws.pool_chunks = NA_CHUNKS;
  ws.maxv = NA_CHUNKS;
  ws.timeout_ms = r.timeout_ms;
  uint32_t* srcw = (uint32_t*)(b + NA_OFF_SRC);
  const auto t0 = std::chrono::steady_clock::now();
  g_na_clean[dev] = false;  // until the cleanup kernel is enqueued
  rc = ensure_split(g, r.light_delta, s);  // kernels only: the graph owns its row buffers
  if (rc == ADDS_OK) {
    const hipError_t e = launch_fill_u32(srcw, 1, src, s);
    if (e != hipSuccess) {
      g_detail = std::string("allocation-free solve: ") + hipGetErrorString(e);
      rc = ADDS_EHIP;
    }
  }
  if (rc == ADDS_OK) rc = enqueue_solve(g, ws, r, (uint32_t*)dist, g->n, srcw, 1, false, s, false);
  if (rc != ADDS_OK) return rc;
  g_na_clean[dev] = true;
  uint32_t hs[HOST_STATE_WORDS];
  HIPCHK(hipMemcpyAsync(hs, ws.st, sizeof(hs), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    read_stats(hs, stats);
    stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return solve_outcome(hs, r.timeout_ms);
}

// ---- cross-XCD ordering litmus (diagnostic) ----
extern "C" adds_status adds_litmus(uint32_t pairs, uint32_t rounds, uint32_t variant, uint64_t* forbidden,
                                   uint64_t* lost, uint64_t* cross_xcd_pairs) {
  if (!pairs || !rounds || variant > 3 || !forbidden || !lost) return ADDS_EINVAL;
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) return ADDS_ENODEV;
  uint32_t* mem = nullptr;
  uint32_t* xcc = nullptr;
  unsigned long long* out = nullptr;
  HIPCHK(hipMalloc(&mem, (size_t)pairs * 128 * 4));
  HIPCHK(hipMalloc(&xcc, (size_t)pairs * 2 * 4));
  HIPCHK(hipMalloc(&out, 16));
  HIPCHK(launch_fill_u32(mem, (uint64_t)pairs * 128, 0, nullptr));
  for (uint32_t p = 0; p < pairs; p++) HIPCHK(hipMemset(mem + (size_t)p * 128 + 96, 0xFF, 4));
  HIPCHK(hipMemset(out, 0, 16));
  HIPCHK(launch_litmus(mem, pairs, rounds, variant, out, xcc, nullptr));
  unsigned long long h[2] = {0, 0};
  std::vector<uint32_t> hx((size_t)pairs * 2);
  HIPCHK(hipMemcpy(h, out, 16, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(hx.data(), xcc, hx.size() * 4, hipMemcpyDeviceToHost));
  (void)hipFree(mem);
  (void)hipFree(xcc);
  (void)hipFree(out);
  *forbidden = h[0];
  *lost = h[1];
  if (cross_xcd_pairs) {
    uint64_t c = 0;
    for (uint32_t p = 0; p < pairs; p++) c += hx[2 * p] != hx[2 * p + 1];
    *cross_xcd_pairs = c;
  }
  return ADDS_OK;
}

extern "C" adds_status adds_xcd_map(uint32_t workgroups, uint32_t threads, uint32_t* xcc) {
  if (!workgroups || !threads || threads > 1024 || !xcc) return ADDS_EINVAL;
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) return ADDS_ENODEV;
  uint32_t* d = nullptr;
  HIPCHK(hipMalloc(&d, (size_t)workgroups * 4));
  HIPCHK(launch_xcd_map(d, workgroups, threads, nullptr));
  HIPCHK(hipMemcpy(xcc, d, (size_t)workgroups * 4, hipMemcpyDeviceToHost));
  (void)hipFree(d);
  return ADDS_OK;
}

extern "C" adds_status adds_gather_ceiling(uint32_t n, uint64_t m, double* gathers_per_s) {
  if (!n || !m || m > (1ull << 31) || !gathers_per_s) return ADDS_EINVAL;
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) return ADDS_ENODEV;
  const uint32_t cu = std::max(1u, adds_device_cu_count());
  uint2* cw = nullptr;
  uint32_t* d = nullptr;
  uint32_t* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  adds_status rc = ADDS_OK;
  double best = 0.0;
  do {
    if (hipMalloc(&cw, m * sizeof(uint2)) != hipSuccess || hipMalloc(&d, (size_t)n * 4) != hipSuccess ||
        hipMalloc(&out, (size_t)cu * 8 * 256 * 4) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || launch_gather_probe_init(cw, m, n, d, nullptr) != hipSuccess) {
      rc = ADDS_EHIP;
      break;
    }
    // one to eight workgroups of 256 threads per CU; the best rate over the shapes
    for (uint32_t wpc = 1; wpc <= 8 && rc == ADDS_OK; wpc *= 2) {
      const uint32_t grid = cu * wpc;
      const int R = 5;
      bool ok = true;
      for (int r = 0; r < 2; r++) ok = ok && launch_gather_probe(cw, d, m, out, grid, nullptr) == hipSuccess;
      (void)hipEventRecord(e0, nullptr);
      for (int r = 0; r < R; r++) ok = ok && launch_gather_probe(cw, d, m, out, grid, nullptr) == hipSuccess;
      (void)hipEventRecord(e1, nullptr);
      if (!ok) {
        rc = ADDS_EHIP;
        break;
      }
      float ms = 0.0f;
      if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.0f) {
        rc = ADDS_EHIP;
        break;
      }
      best = std::max(best, (double)m * R / (ms * 1e-3));
    }
  } while (false);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(cw);
  (void)hipFree(d);
  (void)hipFree(out);
  if (rc == ADDS_OK) *gathers_per_s = best;
  return rc;
}
