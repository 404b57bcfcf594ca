// Cross-XCD message-passing litmus test for the memory-ordering recipe the solver relies on.
//
// Workgroup pairs (2p, 2p+1) are meant to run on different XCDs (workgroups are dealt round-robin over the
// 8 XCDs); every workgroup records the XCD it ran on (the XCC_ID hardware register), so the host counts the
// pairs that really crossed XCDs instead of assuming it.
// In every round the producer
//   1. writes a payload word with a relaxed agent-scope store        (like a worklist entry / adjoint),
//   2. waits for its completion (workgroup-scope fence = s_waitcnt vmcnt(0)),
//   3. publishes the round through a flag: a relaxed agent-scope store in variant 0, an agent-scope
//      fetch_add in variant 1 (like a queue tail / idle counter / child counter).
// The consumer spins on the flag with relaxed agent-scope loads, waits, then loads the payload with a
// relaxed agent-scope load.  Seeing the new flag but an old payload is the forbidden outcome.  A third
// variant checks atomicMin visibility: the producer lowers dist[p] with an agent-scope atomic min, waits,
// and publishes; the consumer must read a dist no larger than the published value.
// Every spin is bounded; a bound hit is counted separately (it would mean a lost message).
#include "device.h"

namespace adds {

__device__ __forceinline__ uint32_t lld(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lst(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The solver's completion wait: s_waitcnt with all counters at zero, fenced against compiler motion.
__device__ __forceinline__ void lwait() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// The XCD (XCC) this wavefront runs on: hardware register HW_REG_XCC_ID, bits [3:0].
__device__ __forceinline__ uint32_t xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 0xFu; }

// words per pair: [0] payload, [32] flag, [64] ack, [96] dist   (each on its own 128-byte line)
__global__ void k_litmus(uint32_t* mem, uint32_t rounds, uint32_t variant, unsigned long long* out, uint32_t* xcc) {
  if (threadIdx.x != 0) return;
  xcc[blockIdx.x] = xcc_id();
  const uint32_t pair = blockIdx.x >> 1;
  const bool producer = (blockIdx.x & 1) == 0;
  uint32_t* w = mem + (size_t)pair * 128;
  unsigned long long forbidden = 0, lost = 0;
  for (uint32_t t = 1; t <= rounds; t++) {
    if (producer) {
      if (variant == 2) {
        __hip_atomic_fetch_min(w + 96, 0xFFFFFFFFu - t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        lst(w, t);
      }
      if (variant != 3) lwait();  // variant 3 omits the wait: it shows whether reordering is observable
      if (variant == 1) __hip_atomic_fetch_add(w + 32, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      else lst(w + 32, t);
      uint32_t spins = 0;
      while (lld(w + 64) != t && ++spins < (1u << 22)) __builtin_amdgcn_s_sleep(1);
      if (spins >= (1u << 22)) {
        lost++;
        break;
      }
    } else {
      uint32_t spins = 0;
      while (lld(w + 32) != t && ++spins < (1u << 22)) __builtin_amdgcn_s_sleep(1);
      if (spins >= (1u << 22)) {
        lost++;
        break;
      }
      lwait();
      if (variant == 2) {
        if (lld(w + 96) > 0xFFFFFFFFu - t) forbidden++;
      } else {
        if (lld(w) < t) forbidden++;
      }
      lwait();
      lst(w + 64, t);
    }
  }
  if (forbidden) atomicAdd(out, forbidden);
  if (lost) atomicAdd(out + 1, lost);
}

hipError_t launch_litmus(uint32_t* mem, uint32_t pairs, uint32_t rounds, uint32_t variant, unsigned long long* out,
                         uint32_t* xcc, hipStream_t s) {
  hipLaunchKernelGGL(k_litmus, dim3(2 * pairs), dim3(64), 0, s, mem, rounds, variant, out, xcc);
  return hipGetLastError();
}

// XCD of every workgroup of a launch of the given shape (the solver's: workgroups_per_cu x CUs, 256 threads).
__global__ void k_xcd_map(uint32_t* xcc) {
  if (threadIdx.x == 0) xcc[blockIdx.x] = xcc_id();
}
hipError_t launch_xcd_map(uint32_t* xcc, uint32_t workgroups, uint32_t threads, hipStream_t s) {
  hipLaunchKernelGGL(k_xcd_map, dim3(workgroups), dim3(threads), 0, s, xcc);
  return hipGetLastError();
}

// Random-gather ceiling: the memory pattern of the solver's edge windows without the solver.  Each
// element i of a stream of m (col, w) pairs names a uniformly random vertex col < n (a 64-bit hash of
// i); a thread loads GP_U pairs of its grid-stride sequence, then gathers d[col] for each and compares
// it with w, so GP_U gathers per lane are in flight.  The rate this reaches for a given n is what a
// kernel issuing one d[v] gather per edge can expect from the memory system when the targets have no
// locality.
__global__ void k_gather_probe_init(uint2* cw, uint64_t m, uint32_t n, uint32_t* d) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m; i += stride) {
    uint64_t h = (i + 1) * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 32;
    cw[i] = make_uint2((uint32_t)(h % n), (uint32_t)(h >> 40));
  }
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) d[i] = (uint32_t)i * 2654435761u;
}
constexpr int GP_U = 8;
__global__ void __launch_bounds__(256) k_gather_probe(const uint2* __restrict__ cw, const uint32_t* __restrict__ d,
                                                      uint64_t m, uint32_t* out) {
  const uint64_t T = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint32_t acc = 0;
  for (uint64_t base = t; base < m; base += T * GP_U) {
    uint64_t x[GP_U];
    uint32_t y[GP_U];
#pragma unroll
    for (int j = 0; j < GP_U; j++) {
      const uint64_t i = base + j * T;
      x[j] = i < m ? __builtin_nontemporal_load((const uint64_t*)cw + i) : 0ull;
    }
#pragma unroll
    for (int j = 0; j < GP_U; j++) y[j] = d[(uint32_t)x[j]];
#pragma unroll
    for (int j = 0; j < GP_U; j++) acc += y[j] > (uint32_t)(x[j] >> 32);
  }
  out[t] = acc;
}
hipError_t launch_gather_probe_init(uint2* cw, uint64_t m, uint32_t n, uint32_t* d, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_probe_init, dim3(4096), dim3(256), 0, s, cw, m, n, d);
  return hipGetLastError();
}
hipError_t launch_gather_probe(const uint2* cw, const uint32_t* d, uint64_t m, uint32_t* out, uint32_t workgroups,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_gather_probe, dim3(workgroups), dim3(256), 0, s, cw, d, m, out);
  return hipGetLastError();
}

}  // namespace adds
