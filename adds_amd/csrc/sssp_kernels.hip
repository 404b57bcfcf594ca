// Persistent asynchronous delta-stepping SSSP for gfx950 (MI355X).
//
// What it computes is the reference's relaxation loop (ads_int/kernel.cu:104-123, ads_float/kernel.cu:
// 113-132): pop a vertex u, read d[u], and for every out-edge (u,v,w) with d[v] > d[u]+w apply
// atomicMin(d[v], d[u]+w) and push v into the bucket of its new distance (wl.h:1902-1928, bucket mapping
// wl.h:1958-1972).  How it is organised is new:
//
//  * No manager kernel.  The reference's one-block scheduler (wl.h:192-1527) is replaced by a ring of
//    `nb` delta buckets, each split into `parts` sub-queues (one per XCD: workgroup i runs on XCD
//    i % 8), with agent-scope atomic head (claim) / tail (reserve) counters.  Every wavefront claims up
//    to `chunk` items from the open buckets with one compare-and-swap.
//  * Queue storage is a pool of 2048-entry chunks handed out by an atomic bump pointer.  A queue maps
//    its virtual chunk j to a physical chunk through `tab`; the producer that reserves a chunk's first
//    slot allocates it.  Nothing is reused inside a solve, so no slot can be overwritten before it is
//    read; the cleanup kernel restores the pool for the next solve.
//  * Entries carry the distance they were pushed with.  A popped entry whose vertex already holds a
//    smaller distance is stale (a later push covers it) and is dropped; otherwise the vertex is relaxed
//    with the carried distance, which is a real path length, so correctness never depends on when the
//    consumer observes the producer's atomicMin.
//  * High-degree vertices (the reference's TB-cooperative path, wl.h:1608-1696) are cut into edge tasks
//    of `task_edges` edges on a device-wide task ring that every wavefront serves first.
//  * Edges of the claimed vertices are flattened with a wavefront prefix sum and processed 64 at a
//    time; each lane finds its owning vertex by binary search in LDS (replacing CUB's WarpScan and the
//    byte owner table, kernel.cu:89-102).
//  * Termination: a 64-bit word counts (busy transitions << 32 | idle wavefronts).  A wavefront goes
//    busy before it claims and idle after its pushes have completed.  The solve is over when a
//    wavefront reads "all idle", finds every queue empty, and re-reads the very same word: no claim can
//    have happened in between, and idle wavefronts do not push.
//
// Memory model (gfx950, one agent, 8 XCDs with private L2s): every cross-wavefront word is accessed
// with relaxed agent-scope atomics (RMW atomics are coherent at agent scope; atomic loads/stores are
// issued with sc1 so they bypass the non-coherent part of the hierarchy).  Ordering between different
// words is obtained by waiting for completion of the earlier access (s_waitcnt vmcnt(0), emitted by a
// workgroup-scope fence) before issuing the later one.  Plain loads are used only for CSR arrays and
// for the d[v] pre-filter, where a stale (larger) value only costs an extra atomic.
#include "device.h"

namespace adds {

// ------------------------------------------------------------------------------------------------
// primitives
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ald(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ald64(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ast(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ast64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t afadd(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t afadd64(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t afmin(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool acas(uint32_t* p, uint32_t& expected, uint32_t desired) {
  return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
}
// Wait until every memory access issued so far by this wavefront has completed (and stop the compiler
// from moving accesses across).
__device__ __forceinline__ void mem_wait() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ uint32_t first(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}
__device__ __forceinline__ uint64_t wave_sum64(uint64_t x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

__device__ __forceinline__ uint32_t* qhead(const KArgs& a, uint32_t q) { return a.st + S_QBASE + 2 * q * LINE; }
__device__ __forceinline__ uint32_t* qtail(const KArgs& a, uint32_t q) {
  return a.st + S_QBASE + 2 * q * LINE + LINE;
}

__device__ __forceinline__ void set_error(const KArgs& a, uint32_t code) {
  uint32_t z = 0;
  acas(a.st + S_DONE + 1, z, code);
}
__device__ __forceinline__ bool stopped(const KArgs& a) {
  return (ald(a.st + S_DONE) | ald(a.st + S_DONE + 1)) != 0;
}

// Distance arithmetic.  Int: uint32 add.  Float: IEEE binary32 add (round to nearest even; the build
// disables contraction), compared as uint32 bits, which orders non-negative floats.
template <bool FLT>
__device__ __forceinline__ uint32_t dadd(uint32_t d, uint32_t w) {
  if constexpr (FLT) {
    return __float_as_uint(__uint_as_float(d) + __uint_as_float(w));
  } else {
    return d + w;
  }
}
template <bool FLT>
__device__ __forceinline__ uint32_t bucket_key(uint32_t d, float inv_delta) {
  float x = FLT ? __uint_as_float(d) : (float)d;
  x = fminf(x * inv_delta, 4.0e9f);
  return (uint32_t)x;
}

struct WaveStats {
  uint64_t v[8];
};

// Spin until a pool-chunk mapping or entry is written.  Every wait here is on a wavefront that has
// already reserved the slot and is running, so the loops end; the cap and the stop flag turn a broken
// invariant into an error instead of a hang.
constexpr uint32_t SPIN_CAP = 1u << 21;

__device__ __forceinline__ uint32_t wait_tab(const KArgs& a, const uint32_t* tp) {
  uint32_t ph = ald(tp);
  uint32_t spins = 0;
  while (ph == TAB_UNSET) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 1023) == 0) {
      if (stopped(a)) return TAB_BAD;
      if (spins > SPIN_CAP) {
        set_error(a, DERR_TIMEOUT);
        return TAB_BAD;
      }
    }
    ph = ald(tp);
  }
  return ph;
}
__device__ __forceinline__ uint64_t wait_entry(const KArgs& a, const uint64_t* ep) {
  uint64_t e = ald64(ep);
  uint32_t spins = 0;
  while (e == EMPTY64) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 1023) == 0) {
      if (stopped(a)) return EMPTY64;
      if (spins > SPIN_CAP) {
        set_error(a, DERR_TIMEOUT);
        return EMPTY64;
      }
    }
    e = ald64(ep);
  }
  return e;
}

// Writes entry `e` at position `pos` of queue `q` (allocating the chunk when pos starts one).
__device__ __forceinline__ void write_entry(const KArgs& a, uint32_t q, uint32_t pos, uint64_t e) {
  const uint32_t j = pos >> CH_SHIFT, off = pos & (CH - 1);
  if (j >= a.maxv) {
    set_error(a, DERR_OVERFLOW);
    return;
  }
  uint32_t* tp = a.tab + (size_t)q * a.maxv + j;
  if (off == 0) {
    uint32_t ph = afadd(a.st + S_POOL, 1u);
    if (ph >= a.pool_chunks) {
      set_error(a, DERR_OVERFLOW);
      ph = TAB_BAD;
    }
    ast(tp, ph);
  }
  const uint32_t ph = wait_tab(a, tp);
  if (ph == TAB_BAD) return;
  ast64(a.pool + ((size_t)ph << CH_SHIFT) + off, e);
}

// Pushes (gid, nd) for every lane with f set, one tail reservation per distinct target queue.
template <bool FLT>
__device__ __forceinline__ void push_wave(const KArgs& a, bool f, uint32_t gid, uint32_t nd, uint32_t part,
                                          uint32_t c_now, WaveStats& ws) {
  uint64_t m = ballot(f);
  if (m == 0) return;
  const uint32_t lane = lane_id();
  uint32_t b = bucket_key<FLT>(nd, a.inv_delta);
  b = max(b, c_now);
  b = min(b, c_now + a.nb - 2);
  const uint32_t q = (b & (a.nb - 1)) * a.parts + part;
  if (lane == 0) ws.v[ST_PUSH] += __popcll(m);
  while (m) {
    const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
    const uint32_t lq = rdlane(q, leader);
    const uint64_t peers = ballot(f && q == lq);
    uint32_t base = 0;
    if (lane == leader) base = afadd(qtail(a, lq), (uint32_t)__popcll(peers));
    base = rdlane(base, leader);
    if (f && q == lq) write_entry(a, lq, base + mbcnt(peers), ((uint64_t)nd << 32) | gid);
    m &= ~peers;
  }
}

// Relaxes UNROLL independent edges per lane (one 64-edge window each) so that UNROLL dependent chains
// (col/w load -> d[v] load -> atomicMin) are in flight per wavefront, then pushes the improvements.
constexpr int UNROLL = 4;

template <bool FLT>
__device__ __forceinline__ void relax_group(const KArgs& a, const bool (&act)[UNROLL], const uint32_t (&e)[UNROLL],
                                            const uint32_t (&du)[UNROLL], const uint32_t (&gb)[UNROLL], uint32_t part,
                                            uint32_t c_now, WaveStats& ws) {
  uint32_t tgt[UNROLL], nd[UNROLL], dv[UNROLL];
  bool f[UNROLL];
#pragma unroll
  for (int j = 0; j < UNROLL; j++) {
    tgt[j] = 0;
    nd[j] = 0;
    if (act[j]) {
      tgt[j] = gb[j] + a.col[e[j]];
      nd[j] = dadd<FLT>(du[j], a.w[e[j]]);
    }
  }
#pragma unroll
  for (int j = 0; j < UNROLL; j++) dv[j] = act[j] ? a.dist[tgt[j]] : 0u;  // stale values only cost an atomic
#pragma unroll
  for (int j = 0; j < UNROLL; j++) {
    f[j] = act[j] && nd[j] < dv[j];
    ws.v[ST_EDGES] += act[j] ? 1u : 0u;
  }
  uint32_t old[UNROLL];
#pragma unroll
  for (int j = 0; j < UNROLL; j++) {
    old[j] = 0xFFFFFFFFu;
    if (f[j]) {
      ws.v[ST_RELAX]++;
      old[j] = afmin(a.dist + tgt[j], nd[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < UNROLL; j++) push_wave<FLT>(a, f[j] && old[j] > nd[j], tgt[j], nd[j], part, c_now, ws);
}

// ------------------------------------------------------------------------------------------------
// the persistent solver
// ------------------------------------------------------------------------------------------------
enum : uint32_t { W_NONE = 0, W_TASK = 1, W_ITEMS = 2, W_STOP = 3 };

template <bool FLT, bool REPL>
__global__ void __launch_bounds__(WG_THREADS) k_sssp(KArgs a) {
  __shared__ uint32_t s_pref[WAVES_PER_WG][65];
  __shared__ uint32_t s_rs[WAVES_PER_WG][64];
  __shared__ uint32_t s_du[WAVES_PER_WG][64];
  __shared__ uint32_t s_gb[WAVES_PER_WG][64];

  const uint32_t lane = lane_id();
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t part = blockIdx.x % a.parts;
  const uint32_t nq = a.nb * a.parts;
  WaveStats ws;
#pragma unroll
  for (int i = 0; i < 8; i++) ws.v[i] = 0;
  bool busy = false;  // wave-uniform
  const uint64_t t_start = wall_clock64();
  uint32_t idle_iter = 0, work_iter = 0;

  for (;;) {
    // ---------------- probe the open queues in parallel, then claim (lane 0) ----------------
    // lane l < (concur+1)*parts probes bucket c-1+l/parts, sub-queue (part + l%parts) % parts, so the
    // ballot's bit order is the claim priority; lane 63 probes the hub task ring (served first).
    const uint32_t c = ald(a.st + S_CUR);
    uint32_t ph_ = 0, pt_ = 0, pq = 0;
    bool pv = false;
    {
      const uint32_t nprobe = (a.concur + 1) * a.parts;
      if (lane < nprobe) {
        const uint32_t j = lane / a.parts;  // 0 = behind bucket c-1
        if (c > 0 || j > 0) {
          const uint32_t slot = (c - 1 + j) & (a.nb - 1);
          pq = slot * a.parts + ((part + lane % a.parts) % a.parts);
          ph_ = ald(qhead(a, pq));
          pt_ = ald(qtail(a, pq));
          pv = ph_ < pt_;
        }
      } else if (lane == 63) {
        ph_ = ald(a.st + S_TASK_H);
        pt_ = ald(a.st + S_TASK_T);
        pv = ph_ < pt_;
      }
    }
    uint64_t pm = ballot(pv);
    uint32_t kind = W_NONE, q = 0, h = 0, cnt = 0, cb = 0;
    if (pm) {
      if (!busy) {
        if (lane == 0) afadd64((uint64_t*)(a.st + S_VER), 0xFFFFFFFFull);  // +1 transition, -1 idle
        busy = true;
        mem_wait();
      }
      // task ring first
      if (pm >> 63) {
        uint32_t th = rdlane(ph_, 63), tt = rdlane(pt_, 63);
        if (lane == 0) {
          for (uint32_t tries = 0; th < tt && tries < 16; tries++) {
            if (acas(a.st + S_TASK_H, th, th + 1)) {
              kind = W_TASK;
              h = th;
              break;
            }
            tt = ald(a.st + S_TASK_T);
          }
        }
        kind = rdlane(kind, 0);
        h = rdlane(h, 0);
        pm &= ~(1ull << 63);
      }
      while (kind == W_NONE && pm) {
        const uint32_t L = (uint32_t)__ffsll((long long)pm) - 1;
        uint32_t hh = rdlane(ph_, L), tt2 = rdlane(pt_, L);
        const uint32_t qq = rdlane(pq, L);
        if (lane == 0) {
          for (uint32_t tries = 0; hh < tt2 && tries < 16; tries++) {
            const uint32_t nn = min(a.chunk, tt2 - hh);
            if (acas(qhead(a, qq), hh, hh + nn)) {
              kind = W_ITEMS;
              h = hh;
              cnt = nn;
              break;
            }
            tt2 = ald(qtail(a, qq));
          }
        }
        kind = rdlane(kind, 0);
        if (kind == W_ITEMS) {
          q = qq;
          h = rdlane(h, 0);
          cnt = rdlane(cnt, 0);
          cb = c - 1 + L / a.parts;
        }
        pm &= pm - 1;
      }
    }

    // The watchdog and the host abort word are also polled while busy, so that a solve that never
    // goes idle can still be ended.
    if (kind != W_NONE && (++work_iter & 255) == 0) {
      uint32_t quit = 0;
      if (lane == 0) {
        if (wall_clock64() - t_start > a.timeout_ticks) {
          set_error(a, DERR_TIMEOUT);
          quit = 1;
        } else if (a.host_abort &&
                   __hip_atomic_load(a.host_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0) {
          set_error(a, DERR_ABORT);
          quit = 1;
        } else if (ald(a.st + S_DONE + 1) != 0) {
          quit = 1;
        }
      }
      if (rdlane(quit, 0)) break;
    }

    if (kind == W_TASK) {
      idle_iter = 0;
      // ---------------- hub edge task ----------------
      if (h >= a.task_cap) {  // only after an overflow: the slot was never written
        set_error(a, DERR_TASKS);
        break;
      }
      const uint4* tp = a.tasks + h;
      uint4 t;
      uint32_t spins = 0;
      for (;;) {
        t.x = ald(&tp->x);
        t.y = ald(&tp->y);
        t.z = ald(&tp->z);
        t.w = ald(&tp->w);
        if (t.x != EMPTY32 && t.y != EMPTY32 && t.z != EMPTY32 && t.w != EMPTY32) break;
        __builtin_amdgcn_s_sleep(1);
        if ((++spins & 1023) == 0 && (stopped(a) || spins > SPIN_CAP)) {
          if (spins > SPIN_CAP) set_error(a, DERR_TIMEOUT);
          t.x = EMPTY32;
          break;
        }
      }
      if (t.x == EMPTY32) break;
      const uint32_t gid_u = first(t.x), du = first(t.y), eb = first(t.z), ee = first(t.w);
      const uint32_t dnow = ald(a.dist + gid_u);
      if (dnow < du) {
        if (lane == 0) ws.v[ST_STALE]++;
        continue;
      }
      const uint32_t gbase = REPL ? (gid_u / a.n) * a.n : 0u;
      for (uint32_t base = eb; base < ee; base += 64 * UNROLL) {
        const uint32_t c_now = ald(a.st + S_CUR);
        bool act[UNROLL];
        uint32_t e[UNROLL], dus[UNROLL], gbs[UNROLL];
#pragma unroll
        for (int j = 0; j < UNROLL; j++) {
          e[j] = base + 64 * j + lane;
          act[j] = e[j] < ee;
          dus[j] = du;
          gbs[j] = gbase;
        }
        relax_group<FLT>(a, act, e, dus, gbs, part, c_now, ws);
      }
      continue;
    }

    if (kind == W_ITEMS) {
      idle_iter = 0;
      // ---------------- load the claimed entries ----------------
      bool have = lane < cnt;
      uint32_t gid = 0, du = 0;
      if (have) {
        const uint32_t pos = h + lane;
        const uint32_t ph = (pos >> CH_SHIFT) < a.maxv ? wait_tab(a, a.tab + (size_t)q * a.maxv + (pos >> CH_SHIFT))
                                                       : TAB_BAD;
        uint64_t e = EMPTY64;
        if (ph != TAB_BAD) e = wait_entry(a, a.pool + ((size_t)ph << CH_SHIFT) + (pos & (CH - 1)));
        if (e == EMPTY64) {
          have = false;
          set_error(a, DERR_TIMEOUT);
        }
        gid = (uint32_t)e;
        du = (uint32_t)(e >> 32);
      }
      if (ballot(lane < cnt && !have)) break;  // error path: leave the solve
      const uint32_t c_now0 = ald(a.st + S_CUR);
      // stale / early filtering
      bool ok = false, early = false;
      if (have) {
        const uint32_t dnow = a.dist[gid];
        if (dnow < du) {
          ws.v[ST_STALE]++;
        } else if (bucket_key<FLT>(du, a.inv_delta) > cb) {
          early = true;
          ws.v[ST_REBUCKET]++;
        } else {
          ok = true;
        }
      }
      push_wave<FLT>(a, early, gid, du, part, c_now0, ws);
      uint32_t rs = 0, deg = 0, gbase = 0;
      if (ok) {
        uint32_t v = gid;
        if (REPL) {
          const uint32_t inst = gid / a.n;
          gbase = inst * a.n;
          v = gid - gbase;
        }
        rs = a.row_ptr[v];
        deg = a.row_ptr[v + 1] - rs;
        ws.v[ST_EXPANDED]++;
      }
      // ---------------- hub vertices become edge tasks ----------------
      const bool hub = ok && deg > a.hub_degree;
      uint64_t hm = ballot(hub);
      while (hm) {
        const uint32_t L = (uint32_t)__ffsll((long long)hm) - 1;
        const uint32_t hg = rdlane(gid, L), hd = rdlane(du, L), hrs = rdlane(rs, L), hdeg = rdlane(deg, L);
        const uint32_t ntask = (hdeg + a.task_edges - 1) / a.task_edges;
        uint32_t t0 = 0;
        if (lane == 0) {
          t0 = afadd(a.st + S_TASK_T, ntask);
          ws.v[ST_TASKS] += ntask;
          if (t0 + ntask > a.task_cap) set_error(a, DERR_TASKS);
        }
        t0 = rdlane(t0, 0);
        for (uint32_t i = lane; i < ntask; i += 64) {
          const uint32_t ti = t0 + i;
          if (ti < a.task_cap) {
            const uint32_t e0 = hrs + i * a.task_edges;
            const uint32_t e1 = min(e0 + a.task_edges, hrs + hdeg);
            uint4* tp = a.tasks + ti;
            ast(&tp->x, hg);
            ast(&tp->y, hd);
            ast(&tp->z, e0);
            ast(&tp->w, e1);
          }
        }
        hm &= hm - 1;
      }
      // ---------------- flattened edge windows ----------------
      const uint32_t mydeg = hub ? 0u : deg;
      const uint32_t incl = wave_incl_scan(mydeg);
      const uint32_t total = rdlane(incl, 63);
      if (total) {
        s_pref[wv][lane] = incl - mydeg;
        s_rs[wv][lane] = rs;
        s_du[wv][lane] = du;
        if (REPL) s_gb[wv][lane] = gbase;
        if (lane == 0) s_pref[wv][64] = total;
        wave_sync();
        for (uint32_t base = 0; base < total; base += 64 * UNROLL) {
          const uint32_t c_now = ald(a.st + S_CUR);
          bool act[UNROLL];
          uint32_t e[UNROLL], dus[UNROLL], gbs[UNROLL];
#pragma unroll
          for (int j = 0; j < UNROLL; j++) {
            const uint32_t i = base + 64 * j + lane;
            act[j] = i < total;
            e[j] = 0;
            dus[j] = 0;
            gbs[j] = 0;
            if (act[j]) {
              // owner = last lane L with s_pref[L] <= i (its degree is then non-zero)
              uint32_t lo = 0, hi = 64;
              while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_pref[wv][mid] <= i) lo = mid;
                else hi = mid;
              }
              e[j] = s_rs[wv][lo] + (i - s_pref[wv][lo]);
              dus[j] = s_du[wv][lo];
              gbs[j] = REPL ? s_gb[wv][lo] : 0u;
            }
          }
          relax_group<FLT>(a, act, e, dus, gbs, part, c_now, ws);
        }
        wave_sync();
      }
      if (ald(a.st + S_DONE + 1) != 0) break;
      continue;
    }

    // ---------------- no work visible: go idle, test termination, advance the head ----------------
    if (busy) {
      mem_wait();  // every push of this wavefront has completed
      if (lane == 0) afadd64((uint64_t*)(a.st + S_VER), 1ull);
      busy = false;
      mem_wait();
    }
    if (stopped(a)) break;
    {
      const uint64_t w1 = ald64((const uint64_t*)(a.st + S_VER));
      if ((uint32_t)w1 == a.total_waves) {
        mem_wait();
        bool nonempty = false;
        for (uint32_t qq = lane; qq < nq; qq += 64) nonempty |= ald(qhead(a, qq)) != ald(qtail(a, qq));
        if (lane == 0) nonempty |= ald(a.st + S_TASK_H) != ald(a.st + S_TASK_T);
        const bool any = ballot(nonempty) != 0;
        mem_wait();
        const uint64_t w2 = ald64((const uint64_t*)(a.st + S_VER));
        if (!any && w2 == w1) {
          if (lane == 0) ast(a.st + S_DONE, 1u);
          break;
        }
      }
    }
    // Advance the head bucket to the first non-empty bucket ahead, but only when the behind and head
    // buckets are empty (items still being processed are not waited for: the solver is asynchronous).
    // One wavefront at a time scans (try-lock): idle wavefronts must not flood the queue counters.
    bool advanced = false;
    uint32_t locked = 0;
    if (lane == 0) {
      uint32_t z = 0;
      locked = acas(a.st + S_ADV, z, 1u) ? 1u : 0u;
    }
    if (rdlane(locked, 0)) {
      const uint32_t c2 = ald(a.st + S_CUR);
      uint32_t firstj = 0xFFFFFFFFu;
      bool behind_or_head = false;
      for (uint32_t j0 = 0; j0 < a.nb; j0 += 64) {
        const uint32_t jj = j0 + lane;  // bucket c2-1+jj
        bool ne = false;
        if (jj < a.nb) {
          const uint32_t slot = (c2 - 1 + jj) & (a.nb - 1);
          for (uint32_t pi = 0; pi < a.parts; pi++) {
            const uint32_t qq = slot * a.parts + pi;
            ne |= ald(qhead(a, qq)) < ald(qtail(a, qq));
          }
        }
        uint64_t m = ballot(ne);
        if (j0 == 0) {
          behind_or_head = (m & (c2 > 0 ? 3ull : 2ull)) != 0;
          m &= ~3ull;
        }
        if (m && firstj == 0xFFFFFFFFu) firstj = j0 + (uint32_t)__ffsll((long long)m) - 1;
      }
      if (!behind_or_head && firstj != 0xFFFFFFFFu) {
        if (lane == 0) {
          uint32_t cc = c2;
          acas(a.st + S_CUR, cc, c2 - 1 + firstj);
        }
        advanced = true;
      }
      mem_wait();
      if (lane == 0) ast(a.st + S_ADV, 0u);
    }
    if (advanced) continue;
    // watchdog and host abort
    ++idle_iter;
    if ((idle_iter & 63) == 0) {
      uint32_t quit = 0;
      if (lane == 0) {
        if (wall_clock64() - t_start > a.timeout_ticks) {
          set_error(a, DERR_TIMEOUT);
          quit = 1;
        } else if (a.host_abort &&
                   __hip_atomic_load(a.host_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0) {
          set_error(a, DERR_ABORT);
          quit = 1;
        }
      }
      if (rdlane(quit, 0)) break;
    }
    // back off harder the longer this wavefront stays idle, to keep idle probes off the counters
    if (idle_iter < 8) __builtin_amdgcn_s_sleep(2);
    else if (idle_iter < 64) __builtin_amdgcn_s_sleep(16);
    else __builtin_amdgcn_s_sleep(64);
  }

  // ---------------- exit: publish this wavefront's counters ----------------
  if (busy) {
    mem_wait();
    if (lane == 0) afadd64((uint64_t*)(a.st + S_VER), 1ull);
  }
#pragma unroll
  for (int i = 0; i < 7; i++) {
    const uint64_t s = wave_sum64(ws.v[i]);
    if (lane == 0 && s) afadd64((uint64_t*)(a.st + S_STATS) + i, s);
  }
}

// ------------------------------------------------------------------------------------------------
// init / cleanup / fills
// ------------------------------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint64_t count, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t count, uint64_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Resets the state words and enqueues the sources into bucket 0, sub-queue 0 (single block).
__global__ void k_init_state(KArgs a, const uint32_t* srcs, uint32_t k, int replicated) {
  const uint32_t nw = state_words(a.nb * a.parts);
  for (uint32_t i = threadIdx.x; i < nw; i += blockDim.x) a.st[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    *(uint64_t*)(a.st + S_VER) = (uint64_t)a.total_waves;
    a.st[S_QBASE + LINE] = k;  // tail of queue 0
    a.st[S_POOL] = (k + CH - 1) >> CH_SHIFT;
  }
  for (uint32_t j = threadIdx.x; j < ((k + CH - 1) >> CH_SHIFT); j += blockDim.x) a.tab[j] = j;
  for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
    const uint32_t gid = replicated ? i * a.n + srcs[i] : srcs[i];
    a.dist[gid] = 0;
    a.pool[i] = (uint64_t)gid;  // distance 0 in the high word
  }
}

// Restores the pool, chunk maps and task ring to their empty state, using the previous solve's
// counters (which the next k_init_state clears).
__global__ void k_cleanup(KArgs a) {
  const uint32_t nq = a.nb * a.parts;
  const uint64_t used_chunks = min((uint64_t)a.st[S_POOL], (uint64_t)a.pool_chunks);
  const uint64_t pool_used = used_chunks << CH_SHIFT;
  const uint64_t ntask = min((uint64_t)a.st[S_TASK_T], (uint64_t)a.task_cap);
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nth = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = tid; i < pool_used; i += nth) a.pool[i] = EMPTY64;
  for (uint64_t i = tid; i < ntask; i += nth) a.tasks[i] = make_uint4(EMPTY32, EMPTY32, EMPTY32, EMPTY32);
  for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
    const uint32_t t = a.st[S_QBASE + 2 * q * LINE + LINE];
    const uint32_t used = min((t + CH - 1) >> CH_SHIFT, a.maxv);
    for (uint32_t j = threadIdx.x; j < used; j += blockDim.x) a.tab[(size_t)q * a.maxv + j] = TAB_UNSET;
  }
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t count, uint32_t v, hipStream_t s) {
  if (!count) return hipSuccess;
  uint64_t blocks = std::min<uint64_t>((count + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, s, p, count, v);
  return hipGetLastError();
}
hipError_t launch_fill_u64(uint64_t* p, uint64_t count, uint64_t v, hipStream_t s) {
  if (!count) return hipSuccess;
  uint64_t blocks = std::min<uint64_t>((count + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u64, dim3((uint32_t)blocks), dim3(256), 0, s, p, count, v);
  return hipGetLastError();
}
hipError_t launch_init(const KArgs& a, const uint32_t* srcs, uint32_t k, int replicated, uint32_t inf_bits,
                       hipStream_t s) {
  hipError_t e = launch_fill_u32(a.dist, a.total_n, inf_bits, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_init_state, dim3(1), dim3(256), 0, s, a, srcs, k, replicated);
  return hipGetLastError();
}
hipError_t launch_solve(const KArgs& a, bool is_float, bool replicated, uint32_t grid, hipStream_t s) {
  if (is_float) {
    if (replicated) hipLaunchKernelGGL((k_sssp<true, true>), dim3(grid), dim3(WG_THREADS), 0, s, a);
    else hipLaunchKernelGGL((k_sssp<true, false>), dim3(grid), dim3(WG_THREADS), 0, s, a);
  } else {
    if (replicated) hipLaunchKernelGGL((k_sssp<false, true>), dim3(grid), dim3(WG_THREADS), 0, s, a);
    else hipLaunchKernelGGL((k_sssp<false, false>), dim3(grid), dim3(WG_THREADS), 0, s, a);
  }
  return hipGetLastError();
}
hipError_t launch_cleanup(const KArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_cleanup, dim3(1024), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace adds
