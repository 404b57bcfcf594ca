// Backward pass of float SSSP: dL/dw from dL/d(dist) through each source's shortest-path tree.
//
// The reference has no predecessor tree (its float "parent" field holds the vertex's own id,
// ads_float/kernel.cu:82, csr_graph.h:34-36), so the tree is derived from converged distances, per
// source s over its reached vertices:
//   round 0: the source is the root; pe(v) = the smallest edge id e = (u -> v) with d[u] < d[v] and
//            fl(d[u] + w_e) == d[v];
//   round r: a vertex still without a parent takes the smallest e = (u -> v) with d[u] == d[v] and
//            fl(d[u] + w_e) == d[v] whose tail was resolved before round r (zero-weight edges and adds
//            that round away are tree edges too; parents stay acyclic).
// The gradient is an exact integer sum: with one power-of-two scale S per batch (chosen on the host from
// the largest |g| and the number of nonzero g so that no sum can overflow), G = rint(g * 2^S) is a
// 64-bit integer and
//   GW[e] = sum over (s, t) of G_s[t] * [e is on t's tree path],   dL/dw_e = ldexpf((float)GW[e], -S),
// exact sums rounded once, so the result does not depend on the order of any addition.
//
// Kernels:
//   k_bwd_parent   round 0, one thread per vertex for all k sources (the vertex's in-edges, from the
//                  transposed CSR, are loaded once and the distance loads of several sources are in flight
//                  together): parent (edge, tail) pairs, per-block counts of unresolved vertices and
//                  nonzero g and the largest |g| (bits), and the list of nonzero g entries (staged per
//                  workgroup in LDS, one reservation per workgroup)
//   k_bwd_reduce   adds the per-block counts
//   k_bwd_round / k_bwd_commit   the equal-distance rounds (rare: zero or absorbed weights), in two
//                  passes so that a round reads only parents chosen in earlier rounds
//   sparse g (few nonzero entries): the tree paths of the nonzero entries share most of their length, so
//   each union edge is updated once instead of once per entry below it:
//     k_bwd_claim    one lane per entry climbs from its vertex and claims vertices (compare-and-swap of
//                    the owner word) until it meets a claimed one: the union of the paths is cut into
//                    disjoint segments, and each segment records the segment it attaches to and where
//     k_bwd_attach   one lane per entry climbs the segment tree and adds its G at every attach vertex
//     k_bwd_segment  one lane per segment walks it bottom-up with a running sum (its own G plus what
//                    attaches at each vertex) and adds it to the edge into each vertex
//     Ownership is decided by races, but every sum is an integer sum: the result is the same for any
//     winner.  Each phase's longest chain is a segment or a segment-tree path, not the tree's depth.
//   dense g: k_bwd_count / k_bwd_climb: child counters, then one climb per leaf; the last child to
//                  arrive at a vertex carries its subtree sum on up
//   k_bwd_convert  dL/dw = ldexpf((float)GW, -S)
#include "device.h"

namespace adds {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t ROOT = 0xFFFFFFFEu;

__device__ __forceinline__ void bwd_mem_wait() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ __forceinline__ void add64(unsigned long long* p, long long v) {
  (void)__hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long fixed(float g, int S) { return (long long)rint(ldexp((double)g, S)); }

__global__ void k_bwd_tweights(const uint32_t* reid, const float* w, float* rw, uint64_t m) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m; p += (uint64_t)gridDim.x * blockDim.x)
    rw[p] = w[reid[p]];
}

// Round 0 for every source of vertex v (threads over v, so the in-edge lists are read once).  Up to
// IN_REG in-edges are held in registers and the distance loads of SRC_BATCH sources are issued together;
// longer in-edge lists take the plain loop.
constexpr int IN_REG = 8;
constexpr int SRC_BATCH = 4;
constexpr uint32_t PARENT_THREADS = 256;
constexpr uint32_t LIST_STAGE = 4096;  // LDS staging of list entries per workgroup (flushed when full)

__device__ __forceinline__ void flush_list(const BwdArgs& b, uint32_t* s_list, uint32_t* s_cnt, uint32_t* s_base) {
  __syncthreads();
  const uint32_t c = min(*s_cnt, LIST_STAGE);
  if (threadIdx.x == 0 && c)
    *s_base = (uint32_t)__hip_atomic_fetch_add(b.ctr + 4, (unsigned long long)c, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < c; i += blockDim.x)
    if (*s_base + i < b.list_cap) b.list[*s_base + i] = s_list[i];
  __syncthreads();
  if (threadIdx.x == 0) *s_cnt = 0;
  __syncthreads();
}

__global__ void __launch_bounds__(PARENT_THREADS) k_bwd_parent(BwdArgs b) {
  __shared__ uint32_t s_list[LIST_STAGE];
  __shared__ uint32_t s_cnt, s_base;
  __shared__ unsigned long long s_part[4][PARENT_THREADS / 64];
  const uint32_t lane = __lane_id(), wv = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  unsigned long long unres = 0, nnz = 0, nonfinite = 0;
  uint32_t maxbits = 0;
  for (uint32_t v0 = blockIdx.x * blockDim.x; v0 < b.n; v0 += gridDim.x * blockDim.x) {
    const uint32_t v = v0 + threadIdx.x;
    const bool have_v = v < b.n;
    const uint32_t vv = have_v ? v : 0u;
    const uint32_t p0 = b.rrow[vv], p1 = have_v ? b.rrow[vv + 1] : p0;
    const bool short_row = p1 - p0 <= IN_REG;
    uint32_t iu[IN_REG], ie[IN_REG];
    float iw[IN_REG];
#pragma unroll
    for (int j = 0; j < IN_REG; j++) {
      const bool ok = short_row && p0 + j < p1;
      iu[j] = ok ? b.rsrc[p0 + j] : vv;
      iw[j] = ok ? b.rw[p0 + j] : __builtin_inff();  // never tight: du + inf != dv for finite dv
      ie[j] = ok ? b.reid[p0 + j] : NONE;
    }
    for (uint32_t s0 = 0; s0 < b.k; s0 += SRC_BATCH) {
      // a full staging buffer is flushed before a batch could overflow it; the fill is read between two
      // barriers so that every wavefront of the workgroup takes the same branch
      __syncthreads();
      const uint32_t staged = s_cnt;
      __syncthreads();
      if (staged + PARENT_THREADS * SRC_BATCH > LIST_STAGE) flush_list(b, s_list, &s_cnt, &s_base);
      float dvb[SRC_BATCH], dub[SRC_BATCH][IN_REG];
#pragma unroll
      for (int i = 0; i < SRC_BATCH; i++) {
        const uint64_t base = (uint64_t)min(s0 + i, b.k - 1) * b.n;
        dvb[i] = b.dist[base + vv];
#pragma unroll
        for (int j = 0; j < IN_REG; j++) dub[i][j] = b.dist[base + iu[j]];
      }
#pragma unroll
      for (int i = 0; i < SRC_BATCH; i++) {
        const uint32_t s = s0 + i;
        if (s >= b.k || !have_v) continue;
        const uint64_t base = (uint64_t)s * b.n;
        const uint64_t idx = base + v;
        const float dv = dvb[i];
        uint32_t pe = NONE, pu = NONE;
        if (dv != __builtin_inff()) {
          if (v == b.srcs[s]) {
            pe = ROOT;
          } else if (short_row) {
            // the smallest tight edge id: in-edges are in ascending edge order, so the first match
#pragma unroll
            for (int j = IN_REG - 1; j >= 0; j--) {
              const float du = dub[i][j];
              if (du < dv && du + iw[j] == dv) {
                pe = ie[j];
                pu = iu[j];
              }
            }
            unres += pe == NONE;
          } else {
            for (uint32_t p = p0; p < p1; p++) {
              const uint32_t u = b.rsrc[p];
              const float du = b.dist[base + u];
              if (du < dv && du + b.rw[p] == dv) {
                pe = b.reid[p];
                pu = u;
                break;
              }
            }
            unres += pe == NONE;
          }
          const float g = b.gdist[idx];
          const uint32_t gb = __float_as_uint(g) & 0x7FFFFFFFu;
          nonfinite += gb >= 0x7F800000u;
          if (gb != 0 && gb < 0x7F800000u) {
            nnz++;
            maxbits = max(maxbits, gb);
            // nonzero g entries for the sparse form (a root's g reaches no edge)
            if (pe != ROOT) s_list[__hip_atomic_fetch_add(&s_cnt, 1u, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_WORKGROUP)] = (uint32_t)idx;
          }
        }
        b.P[idx] = make_uint2(pe, pu);
      }
    }
  }
  flush_list(b, s_list, &s_cnt, &s_base);
  // per-block counts: wave sums in LDS, one record per block (added by k_bwd_reduce: no atomics here)
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    unres += __shfl_xor(unres, d, 64);
    nnz += __shfl_xor(nnz, d, 64);
    nonfinite += __shfl_xor(nonfinite, d, 64);
    maxbits = max(maxbits, (uint32_t)__shfl_xor(maxbits, d, 64));
  }
  if (lane == 0) {
    s_part[0][wv] = unres;
    s_part[1][wv] = nnz;
    s_part[2][wv] = maxbits;
    s_part[3][wv] = nonfinite;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long t = 0;
    for (uint32_t w = 0; w < PARENT_THREADS / 64; w++)
      t = threadIdx.x == 2 ? max(t, s_part[2][w]) : t + s_part[threadIdx.x][w];
    b.part[4 * (uint64_t)blockIdx.x + threadIdx.x] = t;
  }
}

__global__ void k_bwd_reduce(BwdArgs b, uint32_t blocks) {
  __shared__ unsigned long long s[4][256];
  unsigned long long t[4] = {0, 0, 0, 0};
  for (uint32_t i = threadIdx.x; i < blocks; i += blockDim.x) {
    t[0] += b.part[4 * (uint64_t)i];
    t[1] += b.part[4 * (uint64_t)i + 1];
    t[2] = max(t[2], b.part[4 * (uint64_t)i + 2]);
    t[3] += b.part[4 * (uint64_t)i + 3];
  }
  for (int j = 0; j < 4; j++) s[j][threadIdx.x] = t[j];
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long r = 0;
    for (uint32_t i = 0; i < blockDim.x; i++) r = threadIdx.x == 2 ? max(r, s[2][i]) : r + s[threadIdx.x][i];
    b.ctr[threadIdx.x] = r;
  }
}

// One equal-distance round: choices into C (P is only read), then k_bwd_commit.
__global__ void k_bwd_round(BwdArgs b) {
  const uint64_t kn = (uint64_t)b.k * b.n;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < kn; idx += (uint64_t)gridDim.x * blockDim.x) {
    uint2 c = make_uint2(NONE, NONE);
    const float dv = b.dist[idx];
    if (dv != __builtin_inff() && b.P[idx].x == NONE) {
      const uint64_t base = (idx / b.n) * b.n;
      const uint32_t v = (uint32_t)(idx - base);
      for (uint32_t p = b.rrow[v]; p < b.rrow[v + 1]; p++) {
        const uint32_t u = b.rsrc[p];
        const float du = b.dist[base + u];
        if (du == dv && du + b.rw[p] == dv && b.P[base + u].x != NONE) {
          c = make_uint2(b.reid[p], u);
          break;
        }
      }
    }
    b.C[idx] = c;
  }
}

__global__ void k_bwd_commit(BwdArgs b) {
  const uint64_t kn = (uint64_t)b.k * b.n;
  unsigned long long got = 0;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < kn; idx += (uint64_t)gridDim.x * blockDim.x) {
    const uint2 c = b.C[idx];
    if (c.x != NONE) {
      b.P[idx] = c;
      got++;
    }
  }
  if (got) __hip_atomic_fetch_add(b.ctr + 5, got, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- sparse form ----
// Phase A: entry i claims its vertex and the vertices above it until one is already claimed; seg_up[i]
// is the owner of that vertex (NONE when the climb ended at the root), attach[i] the vertex.  The claim
// (compare-and-swap with a result) and the parent record are requested together, one round trip a step.
// Every climb ends: parents strictly decrease (distance, round), and the step count is capped as well.
__global__ void __launch_bounds__(256) k_bwd_claim(BwdArgs b, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint64_t idx = b.list[i];
  const uint64_t base = (idx / b.n) * b.n;
  uint64_t x = idx;
  uint32_t up = NONE, at = NONE;
  for (uint32_t steps = 0; steps <= b.n; steps++) {
    const uint2 pr = b.P[x];
    uint32_t expected = NONE;
    const bool won = __hip_atomic_compare_exchange_strong(b.owner + x, &expected, i, __ATOMIC_RELAXED,
                                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!won) {
      up = expected;
      at = (uint32_t)(x - base);
      break;
    }
    if (pr.x >= ROOT) break;
    x = base + pr.y;
  }
  b.seg_up[i] = up;
  b.attach[i] = at;
}

// Phase B: entry i's G is added at every attach vertex on its way up the segment tree (what hangs below
// a vertex of another segment is the sum of the entries whose segment chain attaches there).
__global__ void __launch_bounds__(256) k_bwd_attach(BwdArgs b, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint64_t idx = b.list[i];
  const uint64_t base = (idx / b.n) * b.n;
  const long long G = fixed(b.gdist[idx], b.S);
  uint32_t c = i;
  for (uint32_t hops = 0; hops < count; hops++) {
    const uint32_t up = b.seg_up[c], at = b.attach[c];
    if (up == NONE) break;
    add64((unsigned long long*)b.Acc + base + at, G);
    c = up;
  }
}

// Phase C: segment i, bottom-up: running sum = its own G + everything attached at or below the current
// vertex = the subtree sum of that vertex, added to the edge into it.
__global__ void __launch_bounds__(256) k_bwd_segment(BwdArgs b, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint64_t idx = b.list[i];
  const uint64_t base = (idx / b.n) * b.n;
  long long run = fixed(b.gdist[idx], b.S);
  uint64_t x = idx;
  for (uint32_t steps = 0; steps <= b.n; steps++) {
    const uint32_t own = b.owner[x];
    const long long a = b.Acc[x];
    const uint2 pr = b.P[x];
    if (own != i) break;
    run += a;
    if (pr.x >= ROOT) break;
    add64(b.GW + pr.x, run);
    x = base + pr.y;
  }
}

// ---- dense form ----
// Pending-children counters (non-root vertices with a parent count at their parent).
__global__ void k_bwd_count(BwdArgs b) {
  const uint64_t kn = (uint64_t)b.k * b.n;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < kn; idx += (uint64_t)gridDim.x * blockDim.x) {
    const uint2 pr = b.P[idx];
    if (pr.x < ROOT)
      __hip_atomic_fetch_add(b.cnt + (idx / b.n) * b.n + pr.y, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// One climb per leaf: a vertex's subtree sum is its own G plus what its children added into Acc; it
// goes onto the vertex's parent edge and into the parent's Acc, then the parent's counter drops (after
// the add has completed), and the child that brings it to zero carries the parent's sum on up.
__global__ void k_bwd_climb(BwdArgs b) {
  const uint64_t kn = (uint64_t)b.k * b.n;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < kn; idx += (uint64_t)gridDim.x * blockDim.x) {
    uint2 pr = b.P[idx];
    if (pr.x >= ROOT || b.cnt[idx] != 0) continue;  // unreached, root, or not a leaf
    const uint64_t base = (idx / b.n) * b.n;
    long long total = fixed(b.gdist[idx], b.S);
    for (;;) {
      add64(b.GW + pr.x, total);
      const uint64_t px = base + pr.y;
      add64((unsigned long long*)b.Acc + px, total);
      const uint2 ppr = b.P[px];
      const long long pg = fixed(b.gdist[px], b.S);
      bwd_mem_wait();  // the Acc add has completed before the count drops
      const uint32_t old = __hip_atomic_fetch_sub(b.cnt + px, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (old != 1 || ppr.x >= ROOT) break;  // other children pending, or the parent is the root
      total = pg + (long long)__hip_atomic_load((unsigned long long*)b.Acc + px, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
      pr = ppr;
    }
  }
}

__global__ void k_bwd_convert(const long long* GW, float* gw, uint64_t m, int S) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x)
    gw[e] = ldexpf((float)GW[e], -S);
}

static uint32_t blocks_for(uint64_t work, uint32_t cap) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((work + 255) / 256, cap));
}

uint32_t bwd_parent_blocks(uint32_t n) { return blocks_for(n, BWD_MAX_PARENT_BLOCKS); }

hipError_t launch_bwd_tweights(const uint32_t* reid, const float* w, float* rw, uint64_t m, hipStream_t s) {
  if (m) hipLaunchKernelGGL(k_bwd_tweights, dim3(blocks_for(m, 8192)), dim3(256), 0, s, reid, w, rw, m);
  return hipGetLastError();
}
hipError_t launch_bwd_parent(const BwdArgs& b, hipStream_t s) {
  const uint32_t blocks = bwd_parent_blocks(b.n);
  hipLaunchKernelGGL(k_bwd_parent, dim3(blocks), dim3(PARENT_THREADS), 0, s, b);
  hipLaunchKernelGGL(k_bwd_reduce, dim3(1), dim3(256), 0, s, b, blocks);
  return hipGetLastError();
}
hipError_t launch_bwd_round(const BwdArgs& b, hipStream_t s) {
  const uint32_t blocks = blocks_for((uint64_t)b.k * b.n, 16384);
  hipLaunchKernelGGL(k_bwd_round, dim3(blocks), dim3(256), 0, s, b);
  hipLaunchKernelGGL(k_bwd_commit, dim3(blocks), dim3(256), 0, s, b);
  return hipGetLastError();
}
hipError_t launch_bwd_sparse(const BwdArgs& b, uint32_t count, hipStream_t s) {
  if (!count) return hipSuccess;
  const uint32_t blocks = (count + 255) / 256;
  hipLaunchKernelGGL(k_bwd_claim, dim3(blocks), dim3(256), 0, s, b, count);
  hipLaunchKernelGGL(k_bwd_attach, dim3(blocks), dim3(256), 0, s, b, count);
  hipLaunchKernelGGL(k_bwd_segment, dim3(blocks), dim3(256), 0, s, b, count);
  return hipGetLastError();
}
hipError_t launch_bwd_climb(const BwdArgs& b, hipStream_t s) {
  const uint32_t blocks = blocks_for((uint64_t)b.k * b.n, 16384);
  hipLaunchKernelGGL(k_bwd_count, dim3(blocks), dim3(256), 0, s, b);
  // one thread per vertex-instance: a climb can be long, so no thread should own several leaves
  hipLaunchKernelGGL(k_bwd_climb, dim3(blocks_for((uint64_t)b.k * b.n, 0x7FFFFFFF)), dim3(256), 0, s, b);
  return hipGetLastError();
}
hipError_t launch_bwd_convert(const BwdArgs& b, hipStream_t s) {
  if (b.m) hipLaunchKernelGGL(k_bwd_convert, dim3(blocks_for(b.m, 16384)), dim3(256), 0, s, (const long long*)b.GW, b.gw,
                              b.m, b.S);
  return hipGetLastError();
}

}  // namespace adds
