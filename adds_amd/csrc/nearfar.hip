// Bulk-synchronous near-far SSSP: the GPU comparison baseline the reference evaluates ADDS against (its
// "nf" runs, nf_int.patch / nf_float.patch over Galois' GPU sssp: a near worklist relaxed in rounds under a
// threshold that moves up by DELTA = 32 * mean weight / mean degree, nf_int.patch:334, while vertices
// improved past the threshold wait in a far pile).  Built here on the same graph handles as the ADDS
// solver so that the two run on the same hardware and inputs.
//
// One round: k_nf_items relaxes the near worklist -- each wavefront takes 64 items, flattens the edges
// of rows up to NF_BIG long into 64-edge windows (relaxed 16 or 4 at a time, their loads overlapped), and
// turns longer rows into NF_BIG-edge tasks that
// k_nf_tasks relaxes one per wavefront, so a hub does not hold one wavefront for its whole row.  An
// improved vertex joins the next near worklist (once per round, a stamp per vertex) when its distance
// is below the threshold, else the far pile (once per phase).  When a round leaves the near worklist
// empty, the threshold moves to the first multiple of delta above the far pile's smallest live
// distance (phases with nothing to do are skipped) and k_nf_split moves the far entries below it into
// the near worklist.  The host reads three counters after every round: the rounds are bulk-synchronous.
//
// Distances are 32-bit words, uint32 or the bits of non-negative floats (which order like the floats),
// lowered with atomicMin; relaxation is d[u] + w (float: one IEEE add, no contraction), strict <, so the
// result is the unique shortest-path solution whatever the schedule.  Integer sums saturate below INF.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "../../include/adds.h"
#include "device.h"
#include "host_util.h"

namespace adds {
namespace {

constexpr uint32_t NF_BIG = 1024;  // rows longer than this are relaxed as NF_BIG-edge tasks
constexpr uint32_t NF_WAVES = 4;   // wavefronts per workgroup
constexpr uint32_t INF32 = 0xFFFFFFFFu;

// counters (u32 words of NfArgs::cnt)
enum { C_NEAR = 0, C_FAR = 1, C_TASK = 2, C_FARNEXT = 3, C_MIN = 4, C_WORDS = 8 };
// u64 statistics (NfArgs::st)
enum { S_EDGES = 0, S_RELAX = 1, S_PUSH = 2, S_WORDS = 4 };

struct NfArgs {
  const uint32_t* row;
  const uint32_t* col;
  const uint32_t* w;
  uint32_t* dist;
  const uint32_t* in;  // near worklist of this round
  uint32_t n_in;
  uint32_t* out;       // next round's near worklist
  uint32_t* far;       // far pile of this phase
  uint2* tasks;        // (vertex, chunk) for rows longer than NF_BIG
  uint32_t* near_stamp;
  uint32_t* far_stamp;
  uint32_t* cnt;
  unsigned long long* st;
  uint32_t round, phase, thr;
  uint32_t is_float;
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t nf_add(uint32_t du, uint32_t w, uint32_t is_float) {
  if (is_float) return __float_as_uint(__uint_as_float(du) + __uint_as_float(w));
  const uint64_t s = (uint64_t)du + w;
  return s >= INF32 ? INF32 - 1 : (uint32_t)s;
}

// Appends val for every lane with f set: one atomic per wavefront.
__device__ __forceinline__ void wave_append(bool f, uint32_t val, uint32_t* list, uint32_t* counter) {
  const uint64_t m = __ballot(f);
  if (!m) return;
  const uint32_t lane = lane_id();
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, (int)leader, 64);
  if (f) list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1))] = val;
}

// Appends v[j] for every (lane, j) with f[j] set: one reservation per wavefront for all UN windows.
template <int UN>
__device__ __forceinline__ void wave_append_n(const bool (&f)[UN], const uint32_t (&v)[UN], uint32_t* list,
                                              uint32_t* counter) {
  uint64_t m[UN];
  uint32_t tot = 0;
#pragma unroll
  for (int j = 0; j < UN; j++) {
    m[j] = __ballot(f[j]);
    tot += (uint32_t)__popcll(m[j]);
  }
  if (!tot) return;
  const uint32_t lane = lane_id();
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(counter, tot);
  base = __shfl(base, 0, 64);
#pragma unroll
  for (int j = 0; j < UN; j++) {
    if (f[j]) list[base + (uint32_t)__popcll(m[j] & ((1ull << lane) - 1))] = v[j];
    base += (uint32_t)__popcll(m[j]);
  }
}

// Relaxes UN windows of edges: edge e[j] from a vertex at distance du[j] where act[j].  Each step is
// issued for every window before any result is used -- the (col, w) loads, the d[v] pre-filter loads,
// the atomicMins, the worklist stamps, one reservation per worklist -- so a group of windows costs a few
// round trips, not a few per window (inactive lanes read edge 0, which exists whenever any edge is
// relaxed).  Returns the atomicMins issued by this lane.
// Measured on RMAT-24: 4 windows 166 ms, 8 windows 110 ms, 16 windows 79 ms per solve.
template <int UN>
__device__ __forceinline__ uint32_t nf_relax(const NfArgs& a, const bool (&act)[UN], const uint32_t (&du)[UN],
                                             const uint32_t (&e)[UN]) {
  uint32_t v[UN], nd[UN], dv[UN], old[UN];
  bool cand[UN], pn[UN], pf[UN];
#pragma unroll
  for (int j = 0; j < UN; j++) {
    const uint32_t ej = act[j] ? e[j] : 0u;
    v[j] = a.col[ej];
    nd[j] = a.w[ej];
  }
#pragma unroll
  for (int j = 0; j < UN; j++) {
    nd[j] = nf_add(du[j], nd[j], a.is_float);
    dv[j] = a.dist[act[j] ? v[j] : 0u];
  }
  uint32_t issued = 0;
#pragma unroll
  for (int j = 0; j < UN; j++) {
    cand[j] = act[j] && nd[j] < dv[j];  // plain load as a pre-filter; the atomic decides
    issued += cand[j] ? 1u : 0u;
    old[j] = cand[j] ? atomicMin(a.dist + v[j], nd[j]) : 0u;
  }
  uint32_t sn[UN];
#pragma unroll
  for (int j = 0; j < UN; j++) {
    const bool imp = cand[j] && nd[j] < old[j];
    const bool near = imp && nd[j] < a.thr;
    pn[j] = near;
    pf[j] = imp && !near;
    sn[j] = near ? atomicExch(a.near_stamp + v[j], a.round)
                 : (pf[j] ? atomicExch(a.far_stamp + v[j], a.phase) : 0u);
  }
#pragma unroll
  for (int j = 0; j < UN; j++) {
    pn[j] = pn[j] && sn[j] != a.round;
    pf[j] = pf[j] && sn[j] != a.phase;
  }
  wave_append_n<UN>(pn, v, a.out, a.cnt + C_NEAR);
  wave_append_n<UN>(pf, v, a.far, a.cnt + C_FAR);
  return issued;
}

__device__ __forceinline__ void add_stats(const NfArgs& a, uint32_t edges, uint32_t relax) {
  // wave totals by lane 0 (edges and relax are per-lane counts)
  for (int o = 32; o >= 1; o >>= 1) {
    edges += __shfl_down(edges, o, 64);
    relax += __shfl_down(relax, o, 64);
  }
  if (lane_id() == 0) {
    if (edges) atomicAdd(a.st + S_EDGES, (unsigned long long)edges);
    if (relax) atomicAdd(a.st + S_RELAX, (unsigned long long)relax);
  }
}

// Windows per group: NF_WIDE while that many windows of edges remain (a task is exactly NF_WIDE windows),
// NF_NARROW for the rest of a short-row chunk.
constexpr int NF_WIDE = 16;
constexpr int NF_NARROW = 4;
static_assert(NF_BIG == 64 * NF_WIDE, "a task is one wide group");

__shared__ uint32_t s_pref[NF_WAVES][65];
__shared__ uint32_t s_rs[NF_WAVES][64];
__shared__ uint32_t s_du[NF_WAVES][64];

// One group of UN windows of a k_nf_items chunk, starting at flattened edge k0: owner of each edge by
// bisection over the chunk's prefix sums, then nf_relax.
template <int UN>
__device__ __forceinline__ uint32_t items_group(const NfArgs& a, uint32_t wv, uint32_t lane, uint32_t k0,
                                                uint32_t total, uint32_t& edges) {
  bool act[UN];
  uint32_t du_[UN], e[UN];
#pragma unroll
  for (int j = 0; j < UN; j++) {
    const uint32_t k = k0 + 64 * j + lane;
    act[j] = k < total;
    uint32_t lo = 0;  // last range whose prefix is <= k (its length is then non-zero)
    for (uint32_t step = 32; step >= 1; step >>= 1)
      if (s_pref[wv][lo + step] <= k) lo += step;
    e[j] = act[j] ? s_rs[wv][lo] + (k - s_pref[wv][lo]) : 0u;
    du_[j] = s_du[wv][lo];
    edges += act[j] ? 1u : 0u;
  }
  return nf_relax<UN>(a, act, du_, e);
}

__global__ void __launch_bounds__(64 * NF_WAVES) k_nf_items(NfArgs a) {
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
  const uint32_t gw = blockIdx.x * NF_WAVES + wv, nw = gridDim.x * NF_WAVES;
  uint32_t edges = 0, relax = 0;
  for (uint64_t base = (uint64_t)gw * 64; base < a.n_in; base += (uint64_t)nw * 64) {
    const uint64_t i = base + lane;
    const bool valid = i < a.n_in;
    const uint32_t u = valid ? a.in[i] : 0u;
    const uint32_t du = valid ? a.dist[u] : 0u;
    const uint32_t rs = valid ? a.row[u] : 0u, deg = valid ? a.row[u + 1] - rs : 0u;
    const bool big = deg > NF_BIG;
    if (big) {
      const uint32_t nt = (deg + NF_BIG - 1) / NF_BIG;
      const uint32_t t0 = atomicAdd(a.cnt + C_TASK, nt);
      for (uint32_t c = 0; c < nt; c++) a.tasks[t0 + c] = make_uint2(u, c);
    }
    const uint32_t d = big ? 0u : deg;
    const uint32_t incl = wave_incl_scan(d);
    const uint32_t total = __shfl(incl, 63, 64);
    s_pref[wv][lane] = incl - d;
    s_rs[wv][lane] = rs;
    s_du[wv][lane] = du;
    if (lane == 0) s_pref[wv][64] = total;
    __builtin_amdgcn_wave_barrier();
    uint32_t k0 = 0;
    for (; k0 + 64 * NF_WIDE <= total; k0 += 64 * NF_WIDE) relax += items_group<NF_WIDE>(a, wv, lane, k0, total, edges);
    for (; k0 < total; k0 += 64 * NF_NARROW) relax += items_group<NF_NARROW>(a, wv, lane, k0, total, edges);
    __builtin_amdgcn_wave_barrier();
  }
  add_stats(a, edges, relax);
}

__global__ void __launch_bounds__(64 * NF_WAVES) k_nf_tasks(NfArgs a) {
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
  const uint32_t gw = blockIdx.x * NF_WAVES + wv, nw = gridDim.x * NF_WAVES;
  const uint32_t nt = a.cnt[C_TASK];
  uint32_t edges = 0, relax = 0;
  for (uint32_t t = gw; t < nt; t += nw) {
    const uint2 tk = a.tasks[t];
    const uint32_t u = tk.x;
    const uint32_t du = a.dist[u];
    const uint32_t rs = a.row[u] + tk.y * NF_BIG, re = min(rs + NF_BIG, a.row[u + 1]);
    for (uint32_t e0 = rs; e0 < re; e0 += 64 * NF_WIDE) {
      bool act[NF_WIDE];
      uint32_t du_[NF_WIDE], e[NF_WIDE];
#pragma unroll
      for (int j = 0; j < NF_WIDE; j++) {
        e[j] = e0 + 64 * j + lane;
        act[j] = e[j] < re;
        du_[j] = du;
        edges += act[j] ? 1u : 0u;
      }
      relax += nf_relax<NF_WIDE>(a, act, du_, e);
    }
  }
  add_stats(a, edges, relax);
}

// Smallest distance among far entries at or above the threshold (entries below it are settled).
__global__ void __launch_bounds__(256) k_nf_farmin(NfArgs a, uint32_t nfar) {
  uint32_t mn = INF32;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nfar; i += gridDim.x * blockDim.x) {
    const uint32_t d = a.dist[a.far[i]];
    if (d >= a.thr) mn = min(mn, d);
  }
  for (int o = 32; o >= 1; o >>= 1) mn = min(mn, (uint32_t)__shfl_down(mn, o, 64));
  if (lane_id() == 0 && mn != INF32) atomicMin(a.cnt + C_MIN, mn);
}

// Far entries below the new threshold `nthr` join the near worklist (a.out); the rest go to far_next.
__global__ void __launch_bounds__(256) k_nf_split(NfArgs a, uint32_t nfar, uint32_t nthr, uint32_t* far_next) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < nfar; i0 += stride) {
    const uint32_t i = i0 + threadIdx.x;
    const bool valid = i < nfar;
    const uint32_t v = valid ? a.far[i] : 0u;
    const uint32_t d = valid ? a.dist[v] : 0u;
    const bool live = valid && d >= a.thr;  // below the old threshold: settled and expanded already
    bool to_near = false, to_far = false;
    if (live && d < nthr) to_near = atomicExch(a.near_stamp + v, a.round) != a.round;
    if (live && d >= nthr) to_far = atomicExch(a.far_stamp + v, a.phase + 1) != a.phase + 1;
    wave_append(to_near, v, a.out, a.cnt + C_NEAR);
    wave_append(to_far, v, far_next, a.cnt + C_FARNEXT);
  }
}

__global__ void k_nf_init(uint32_t* dist, uint32_t n, uint32_t src, uint32_t inf, uint32_t* s0, uint32_t* s1,
                          uint32_t* in) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    dist[i] = i == src ? 0u : inf;
    s0[i] = 0;
    s1[i] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) in[0] = src;
}

// Moves the far-next count into the far count and clears the round counters.
__global__ void k_nf_phase(uint32_t* cnt) {
  if (threadIdx.x == 0) {
    cnt[C_FAR] = cnt[C_FARNEXT];
    cnt[C_FARNEXT] = 0;
    cnt[C_TASK] = 0;
    cnt[C_MIN] = INF32;
  }
}

__global__ void k_nf_round(uint32_t* cnt) {
  if (threadIdx.x == 0) {
    cnt[C_NEAR] = 0;
    cnt[C_TASK] = 0;
  }
}

struct Bufs {
  uint32_t *dist = nullptr, *a = nullptr, *b = nullptr, *far0 = nullptr, *far1 = nullptr, *ns = nullptr,
           *fs = nullptr, *cnt = nullptr;
  uint2* tasks = nullptr;
  unsigned long long* st = nullptr;
  uint32_t* hcnt = nullptr;  // pinned copy of the counters
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~Bufs() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (void* p : {(void*)dist, (void*)a, (void*)b, (void*)far0, (void*)far1, (void*)ns, (void*)fs, (void*)cnt,
                    (void*)tasks, (void*)st})
      if (p) (void)hipFree(p);
    if (hcnt) (void)hipHostFree(hcnt);
  }
};

#define NFCHK(expr)                                                                          \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess) {                                                                  \
      set_error_detail(std::string(#expr) + ": " + hipGetErrorString(_e));                   \
      return _e == hipErrorOutOfMemory ? ADDS_ENOMEM : ADDS_EHIP;                            \
    }                                                                                        \
  } while (0)

// threshold k * delta as a distance word, strictly above `above`
uint32_t threshold_word(double k, double delta, uint32_t is_float) {
  const double t = k * delta;
  if (is_float) {
    const float f = (float)t;
    uint32_t b;
    memcpy(&b, &f, 4);
    return std::isinf(f) ? 0x7F800000u : b;
  }
  return t >= 4294967295.0 ? INF32 : (uint32_t)t;
}

}  // namespace
}  // namespace adds

using namespace adds;

extern "C" adds_status adds_sssp_nearfar(adds_graph* g, uint32_t src, void* dist, float delta, uint32_t flags,
                                         adds_stats* stats, void* stream) {
  uint32_t n = 0;
  uint64_t m = 0;
  adds_wtype wt;
  if (!g || !dist || adds_graph_info(g, &n, &m, &wt) != ADDS_OK || src >= n) return ADDS_EINVAL;
  const uint32_t* row;
  const uint32_t* col;
  const void* w;
  adds_graph_device_arrays(g, &row, &col, &w);
  const uint32_t is_float = wt == ADDS_W_F32;
  double dl = delta;
  if (!(dl > 0)) {  // the reference's default: 32 * mean weight / mean degree (nf_int.patch:334)
    const double ave_deg = n ? (double)m / n : 1.0;
    dl = 32.0 * graph_mean_weight(g) / (ave_deg > 0 ? ave_deg : 1.0);
    if (!is_float) dl = std::floor(dl);
    if (!(dl > 0)) dl = is_float ? 1e-30 : 1.0;
  }
  if (!is_float) dl = std::max(1.0, std::floor(dl));
  hipStream_t s = (hipStream_t)stream;
  Bufs B;
  const uint64_t ntask = 2 * m / NF_BIG + 64;
  NFCHK(hipMalloc(&B.dist, 4ull * n));
  NFCHK(hipMalloc(&B.a, 4ull * n + 4));
  NFCHK(hipMalloc(&B.b, 4ull * n + 4));
  NFCHK(hipMalloc(&B.far0, 4ull * n + 4));
  NFCHK(hipMalloc(&B.far1, 4ull * n + 4));
  NFCHK(hipMalloc(&B.ns, 4ull * n));
  NFCHK(hipMalloc(&B.fs, 4ull * n));
  NFCHK(hipMalloc(&B.cnt, 4 * C_WORDS));
  NFCHK(hipMalloc(&B.tasks, 8 * ntask));
  NFCHK(hipMalloc(&B.st, 8 * S_WORDS));
  NFCHK(hipHostMalloc((void**)&B.hcnt, 4 * C_WORDS, hipHostMallocDefault));
  NFCHK(hipMemsetAsync(B.cnt, 0, 4 * C_WORDS, s));
  NFCHK(hipMemsetAsync(B.st, 0, 8 * S_WORDS, s));
  NFCHK(hipEventCreate(&B.e0));
  NFCHK(hipEventCreate(&B.e1));
  NFCHK(hipEventRecord(B.e0, s));
  const uint32_t inf = is_float ? 0x7F800000u : INF32;
  hipLaunchKernelGGL(k_nf_init, dim3(1024), dim3(256), 0, s, B.dist, n, src, inf, B.ns, B.fs, B.a);
  NFCHK(hipGetLastError());

  NfArgs a{};
  a.row = row;
  a.col = col;
  a.w = (const uint32_t*)w;
  a.dist = B.dist;
  a.tasks = B.tasks;
  a.near_stamp = B.ns;
  a.far_stamp = B.fs;
  a.cnt = B.cnt;
  a.st = B.st;
  a.is_float = is_float;
  uint32_t* in = B.a;
  uint32_t* out = B.b;
  uint32_t* far = B.far0;
  uint32_t* far_next = B.far1;
  uint32_t n_in = 1, round = 1, phase = 1, rounds = 0, splits = 0;
  double k = 1;  // threshold = k * delta
  uint32_t thr = threshold_word(k, dl, is_float);
  uint64_t work = 0;
  const uint32_t cu = std::max(1u, adds_device_cu_count());
  for (;;) {
    while (n_in) {
      a.in = in;
      a.n_in = n_in;
      a.out = out;
      a.far = far;
      a.round = ++round;
      a.phase = phase;
      a.thr = thr;
      hipLaunchKernelGGL(k_nf_round, dim3(1), dim3(64), 0, s, B.cnt);
      // Enough wavefronts to keep many loads in flight per CU (both kernels are bound by load latency):
      // up to 16 workgroups of 4 wavefronts per CU.
      const uint32_t blocks = std::min<uint64_t>(16ull * cu, (n_in + 64 * NF_WAVES - 1) / (64 * NF_WAVES));
      hipLaunchKernelGGL(k_nf_items, dim3(blocks), dim3(64 * NF_WAVES), 0, s, a);
      hipLaunchKernelGGL(k_nf_tasks, dim3(16 * cu), dim3(64 * NF_WAVES), 0, s, a);
      NFCHK(hipGetLastError());
      NFCHK(hipMemcpyAsync(B.hcnt, B.cnt, 4 * C_WORDS, hipMemcpyDeviceToHost, s));
      NFCHK(hipStreamSynchronize(s));
      work += n_in;
      rounds++;
      n_in = B.hcnt[C_NEAR];
      std::swap(in, out);
    }
    const uint32_t nfar = B.hcnt[C_FAR];
    if (nfar == 0) break;
    // next threshold: the first multiple of delta above the far pile's smallest live distance
    a.far = far;
    a.thr = thr;
    a.round = ++round;
    a.phase = phase;
    a.out = in;
    NFCHK(hipMemsetAsync(B.cnt + C_MIN, 0xFF, 4, s));
    hipLaunchKernelGGL(k_nf_farmin, dim3(std::min<uint32_t>(4 * cu, (nfar + 255) / 256)), dim3(256), 0, s, a, nfar);
    NFCHK(hipMemcpyAsync(B.hcnt, B.cnt, 4 * C_WORDS, hipMemcpyDeviceToHost, s));
    NFCHK(hipStreamSynchronize(s));
    const uint32_t mn = B.hcnt[C_MIN];
    if (mn == INF32 || mn == inf) break;  // every far entry has settled below the threshold
    double mv;
    if (is_float) {
      float f;
      memcpy(&f, &mn, 4);
      mv = f;
    } else {
      mv = mn;
    }
    k = std::max(k + 1, std::floor(mv / dl) + 1);
    uint32_t nthr = threshold_word(k, dl, is_float);
    while (nthr <= mn && nthr != inf && nthr != INF32) nthr = threshold_word(++k, dl, is_float);
    NFCHK(hipMemsetAsync(B.cnt + C_NEAR, 0, 4, s));
    hipLaunchKernelGGL(k_nf_split, dim3(std::min<uint32_t>(4 * cu, (nfar + 255) / 256)), dim3(256), 0, s, a, nfar,
                       nthr, far_next);
    hipLaunchKernelGGL(k_nf_phase, dim3(1), dim3(64), 0, s, B.cnt);
    NFCHK(hipGetLastError());
    NFCHK(hipMemcpyAsync(B.hcnt, B.cnt, 4 * C_WORDS, hipMemcpyDeviceToHost, s));
    NFCHK(hipStreamSynchronize(s));
    n_in = B.hcnt[C_NEAR];
    std::swap(far, far_next);
    thr = nthr;
    phase++;
    splits++;
    if (n_in == 0 && B.hcnt[C_FAR] == 0) break;
  }
  NFCHK(hipEventRecord(B.e1, s));
  NFCHK(hipEventSynchronize(B.e1));
  float ms = 0;
  NFCHK(hipEventElapsedTime(&ms, B.e0, B.e1));
  unsigned long long hst[S_WORDS];
  NFCHK(hipMemcpy(hst, B.st, 8 * S_WORDS, hipMemcpyDeviceToHost));
  const hipMemcpyKind kind = (flags & ADDS_F_HOST_OUT) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  NFCHK(hipMemcpyAsync(dist, B.dist, 4ull * n, kind, s));
  NFCHK(hipStreamSynchronize(s));
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->ms = ms;
    stats->kernel_ms = ms;
    stats->vertices_expanded = work;
    stats->edges_scanned = hst[S_EDGES];
    stats->relaxations = hst[S_RELAX];
    stats->rebuckets = splits;
    stats->probes = rounds;
  }
  return ADDS_OK;
}
