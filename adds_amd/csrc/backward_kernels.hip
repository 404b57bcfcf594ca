// Backward pass of float SSSP: dL/dw from dL/d(dist) through each source's shortest-path tree.
//
// The reference has no predecessor tree (its float "parent" field holds the vertex's own id,
// ads_float/kernel.cu:82, csr_graph.h:34-36), so the tree is derived from converged distances:
//   parent edge pe(v) = the smallest edge id e = (u -> v) with fl(d[u] + w_e) == d[v] and d[u] < d[v].
// The strict d[u] < d[v] makes parent pointers decrease d, so they form a forest (vertices without such
// an in-edge, the source among them, are roots).  With d[v] = d[pe.u] + w_pe along the tree:
//   adj[v] = g[v] + sum over children c of adj[c]      (dL/dd[v], total derivative)
//   dL/dw_e = sum over sources s with pe_s(v) == e of adj_s[v]
// Every floating-point sum has a fixed order (children by ascending out-edge id, sources by index), so
// the result is deterministic and equal to a sequential CPU evaluation in the same order.
//
// Kernels: k_bwd_parent picks pe/pu, counts children and marks in the parent's child mask which of its
// out-edges (by position in its row) are tree edges; k_bwd_leaves marks childless vertices; k_bwd_accum
// starts one climb per leaf: it finalises adj of a vertex, then decrements its parent's child counter
// and continues with the parent when it was the last child; k_bwd_gw scatters adjoints onto parent
// edges (each edge has one head vertex, so one thread owns each output word).
#include "device.h"

namespace adds {

constexpr uint32_t NO_EDGE = 0xFFFFFFFFu;
constexpr uint32_t LEAF = 0x80000000u;
// child mask: bit p = out-edge p of the row is a tree edge, for p < 31; bit 31 = some tree edge sits at
// position 31 or beyond (the climb then scans the whole row)
constexpr uint32_t MASK_WIDE = 0x80000000u;

__device__ __forceinline__ uint32_t bald(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_bwd_parent(BwdArgs b, uint32_t* pu) {
  const uint64_t kn = (uint64_t)b.k * b.n;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < kn; idx += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t base = (idx / b.n) * b.n;
    const uint32_t v = (uint32_t)(idx - base);
    const float dv = b.dist[idx];
    uint32_t pe = NO_EDGE, u = NO_EDGE;
    if (dv != __builtin_inff()) {
      for (uint32_t p = b.rrow[v]; p < b.rrow[v + 1]; p++) {
        const uint32_t x = b.rsrc[p];
        const uint32_t e = b.reid[p];
        const float du = b.dist[base + x];
        if (du < dv && du + b.w[e] == dv) {
          pe = e;
          u = x;
          break;
        }
      }
    }
    b.pe[idx] = pe;
    pu[idx] = u;
    if (pe != NO_EDGE) {
      __hip_atomic_fetch_add(b.cnt + base + u, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t pos = pe - b.row_ptr[u];
      __hip_atomic_fetch_or(b.cmask + base + u, pos < 31 ? 1u << pos : MASK_WIDE, __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

__global__ void k_bwd_leaves(BwdArgs b) {
  const uint64_t kn = (uint64_t)b.k * b.n;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < kn; idx += (uint64_t)gridDim.x * blockDim.x)
    if (b.cnt[idx] == 0 && b.dist[idx] != __builtin_inff()) b.cnt[idx] = LEAF;
}

// One climb per leaf.  Each tree level costs four dependent round trips: the children's ids, their
// adjoints, this vertex's adjoint store (with the parent's row, mask, gradient and parent loaded
// alongside, in case this climb continues there), and the parent's child counter.
__global__ void k_bwd_accum(BwdArgs b, const uint32_t* pu) {
  const uint64_t kn = (uint64_t)b.k * b.n;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < kn; idx += (uint64_t)gridDim.x * blockDim.x) {
    if (b.cnt[idx] != LEAF) continue;
    uint64_t x = idx;
    const uint64_t base = (idx / b.n) * b.n;
    uint32_t v = (uint32_t)(x - base);
    uint32_t e0 = b.row_ptr[v], e1 = b.row_ptr[v + 1], mask = bald(b.cmask + x), up = pu[x];
    float gx = b.gdist[x];
    for (;;) {
      float a = gx;
      if (mask & MASK_WIDE) {
        // some child beyond position 30: every out-edge, eight at a time, all loads of a batch in flight
        for (uint32_t eb = e0; eb < e1; eb += 8) {
          uint64_t c[8];
          bool child[8];
          float ca[8];
#pragma unroll
          for (int j = 0; j < 8; j++) c[j] = base + b.col[eb + j < e1 ? eb + j : e0];
#pragma unroll
          for (int j = 0; j < 8; j++) child[j] = eb + j < e1 && b.pe[c[j]] == eb + j;
#pragma unroll
          for (int j = 0; j < 8; j++) ca[j] = __uint_as_float(bald((const uint32_t*)b.adj + (child[j] ? c[j] : x)));
#pragma unroll
          for (int j = 0; j < 8; j++)
            if (child[j]) a += ca[j];
        }
      } else {
        // the children named by the mask, in ascending out-edge order, eight at a time
        uint32_t mm = mask;
        while (mm) {
          uint32_t pos[8];
          bool has[8];
#pragma unroll
          for (int j = 0; j < 8; j++) {
            has[j] = mm != 0;
            pos[j] = has[j] ? (uint32_t)__builtin_ctz(mm) : 0u;
            mm &= mm - 1;  // (a no-op once empty)
          }
          uint64_t c[8];
          float ca[8];
#pragma unroll
          for (int j = 0; j < 8; j++) c[j] = has[j] ? base + b.col[e0 + pos[j]] : x;
#pragma unroll
          for (int j = 0; j < 8; j++) ca[j] = __uint_as_float(bald((const uint32_t*)b.adj + c[j]));
#pragma unroll
          for (int j = 0; j < 8; j++)
            if (has[j]) a += ca[j];
        }
      }
      __hip_atomic_store((uint32_t*)b.adj + x, __float_as_uint(a), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (up == NO_EDGE) break;
      // the parent's row, mask, gradient and parent, in flight with the store
      const uint64_t px = base + up;
      const uint32_t pe0 = b.row_ptr[up], pe1 = b.row_ptr[up + 1], pmask = bald(b.cmask + px), pup = pu[px];
      const float pg = b.gdist[px];
      // adj[x] has completed (s_waitcnt with all counters at zero) before the count drops
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_waitcnt(0);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      const uint32_t old = __hip_atomic_fetch_sub(b.cnt + px, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (old != 1) break;
      // last child done: the parent is ready
      x = px;
      v = up;
      e0 = pe0;
      e1 = pe1;
      mask = pmask;
      up = pup;
      gx = pg;
    }
  }
}

__global__ void k_bwd_gw(BwdArgs b) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < b.n; v += gridDim.x * blockDim.x) {
    for (uint32_t s = 0; s < b.k; s++) {
      const uint64_t idx = (uint64_t)s * b.n + v;
      const uint32_t e = b.pe[idx];
      if (e != NO_EDGE) b.gw[e] += b.adj[idx];
    }
  }
}

hipError_t launch_backward(const BwdArgs& b, hipStream_t s) {
  const uint64_t kn = (uint64_t)b.k * b.n;
  // pu lives after the adjoint block of the same staging buffer (4 words per vertex-instance)
  uint32_t* pu = (uint32_t*)b.adj + kn;
  hipError_t e = launch_fill_u32(b.cnt, kn, 0, s);
  if (e == hipSuccess) e = launch_fill_u32(b.cmask, kn, 0, s);
  if (e == hipSuccess) e = launch_fill_u32((uint32_t*)b.gw, b.m, 0, s);
  if (e != hipSuccess) return e;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((kn + 255) / 256, 16384);
  hipLaunchKernelGGL(k_bwd_parent, dim3(blocks), dim3(256), 0, s, b, pu);
  hipLaunchKernelGGL(k_bwd_leaves, dim3(blocks), dim3(256), 0, s, b);
  // one thread per vertex-instance: a climb can be long, so no thread should own several leaves
  const uint32_t ablocks = (uint32_t)std::min<uint64_t>((kn + 255) / 256, 0x7FFFFFFFull);
  hipLaunchKernelGGL(k_bwd_accum, dim3(ablocks), dim3(256), 0, s, b, (const uint32_t*)pu);
  const uint32_t vblocks = (uint32_t)std::min<uint64_t>((b.n + 255) / 256, 16384);
  hipLaunchKernelGGL(k_bwd_gw, dim3(vblocks), dim3(256), 0, s, b);
  return hipGetLastError();
}

}  // namespace adds
