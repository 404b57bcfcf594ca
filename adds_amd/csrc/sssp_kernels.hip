// Persistent asynchronous delta-stepping SSSP for gfx950 (MI355X).
//
// What it computes is the reference's relaxation loop (ads_int/kernel.cu:104-123, ads_float/kernel.cu:
// 113-132): pop a vertex u, read d[u], and for every out-edge (u,v,w) with d[v] > d[u]+w apply
// atomicMin(d[v], d[u]+w) and push v into the bucket of its new distance (wl.h:1902-1928, bucket mapping
// wl.h:1958-1972).  How it is organised is new:
//
//  * No manager kernel.  The reference's one-block scheduler (wl.h:192-1527) is replaced by a ring of
//    `nb` delta buckets, each split into `parts` sub-queues (workgroup i pushes to and first claims from
//    sub-queue i % parts; workgroups are dealt round-robin over the 8 XCDs).  Each queue has three
//    agent-scope counters: tail (positions reserved by producers), published (entries written and not
//    yet claimed) and head (positions claimed).  A consumer takes up to `chunk` from the published
//    count with one fetch_sub (a shortfall is given back) and then advances the head by what it took:
//    no compare-and-swap retry loops.
//  * Light and heavy edges (Meyer & Sanders' delta-stepping).  Each CSR row is stored light-first
//    (w <= light threshold, which is at least delta) with a per-vertex light count.  Expanding a
//    vertex from bucket b relaxes its light edges and files the vertex in b's heavy queue.  The head bucket's heavy queues open once its light
//    phase is done (nothing queued or in flight); heavy edges then land in later buckets only, so each
//    settled vertex relaxes them once.  Without the split, high-degree vertices would re-relax all
//    their edges every time a light in-neighbour improved them within their bucket.
//  * Per-bucket in-flight counters gate the head: like the reference's manager, which advances the
//    head bag only when it is done (wl.h:714-723), the head moves on when the behind and head buckets
//    have nothing queued and nothing in flight, light or heavy; `concur` buckets accept light claims at
//    once (the reference's concurrent bags, wl.h MIN/MAX_CON_BAG).
//  * Queue storage is a pool of 2048-entry chunks handed out by an atomic bump pointer.  A queue maps
//    its virtual chunk j to a physical chunk through `tab`; the producer that reserves a chunk's first
//    slot allocates it.  Nothing is reused inside a solve, so no slot can be overwritten before it is
//    read; the cleanup kernel restores the pool for the next solve.
//  * Entries carry the distance they were pushed with.  A popped entry whose vertex already holds a
//    smaller distance is stale (a later push covers it) and is dropped; otherwise the vertex is relaxed
//    with the carried distance, which is a real path length, so correctness never depends on when the
//    consumer observes the producer's atomicMin.  Pushes are clamped to the window of the bucket being
//    processed, so no hint of the global head is needed on the hot path.
//  * Long edge ranges (the reference's TB-cooperative path for high-degree vertices, wl.h:1608-1696)
//    are not relaxed by the wavefront that meets them (beyond their first chunk): it writes range
//    records and pushes one reference item per `task_edges`-edge chunk into its own queue, all in one
//    reservation, so the chunks spread over every wavefront through the ordinary claims, phases and
//    in-flight accounting of the bucket.
//  * Edges of the claimed vertices are flattened with a wavefront prefix sum and processed in windows
//    of 64 (four or eight in flight per lane); each lane finds its owning vertex by binary search in LDS
//    (replacing CUB's WarpScan and the byte owner table, kernel.cu:89-102).  Improvements over a
//    plain-load pre-filter are collected per wavefront in LDS; one flush per claim applies their
//    atomicMin and queues those that lowered the distance.
//  * Local continuation: a light claim keeps up to 64 improvements for its own bucket and part and
//    processes them as its next claim, so a hop inside a bucket costs no queue round trips.
//  * Tail mode (optional): once few items are left when the head moves, the rest of the solve runs as
//    one bucket, saving the phase changes of a long sparse tail.
//  * Termination: eight 64-bit words, one per group of workgroups, count (busy transitions << 32 |
//    idle wavefronts).  A wavefront goes busy before it claims and idle after its pushes have
//    completed.  The solve is over when a wavefront reads "all idle" in all eight, finds every queue
//    empty, and re-reads the very same words: no claim can have happened in between, and idle
//    wavefronts do not push.
//  * Contention: every word that many wavefronts update lives on a cache line of its own (queue
//    counters, in-flight counters, termination words), and idle wavefronts poll through cached hints.
//
// Memory model (gfx950, one agent, 8 XCDs with private L2s): every word that decides ownership is
// accessed with relaxed agent-scope atomics (RMW atomics are coherent at agent scope; atomic
// loads/stores carry sc1).  Ordering between different words is obtained by waiting for completion of
// the earlier access (an explicit s_waitcnt with all counters at zero, mem_wait below) before issuing
// the later one.  Plain (L2-cached) loads are used for the CSR, the d[v] pre-filter (a stale, larger
// value only costs an extra atomic) and, as hints only, for probing queue counters and idle polls;
// every eighth idle round and the full termination test use atomic loads.
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
// Hint loads: plain cached loads (a workgroup-scope atomic load carries no cache-bypass bits; it
// may return a stale value), never used to decide ownership.  `volatile` would not do: on gfx950 it
// bypasses the caches like a system-scope load.
__device__ __forceinline__ uint32_t hld(const uint32_t* p, bool cached) {
  return cached ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : ald(p);
}
__device__ __forceinline__ uint64_t hld64(const uint64_t* p, bool cached) {
  return cached ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : ald64(p);
}
// Wait until every memory access issued so far by this wavefront has completed (s_waitcnt with all
// counters at zero), with compiler fences so that no access moves across.  A workgroup-scope fence
// would not do: outside tgsplit mode it emits no vmcnt wait at all on gfx950.
__device__ __forceinline__ void mem_wait() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ uint64_t rdlane64(uint64_t v, uint32_t l) {
  return ((uint64_t)rdlane((uint32_t)(v >> 32), l) << 32) | rdlane((uint32_t)v, l);
}
__device__ __forceinline__ uint32_t ffs64(uint64_t m) { return (uint32_t)__ffsll((long long)m) - 1; }

__device__ __forceinline__ uint64_t wave_sum64(uint64_t x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// ---- state addressing ----
__device__ __forceinline__ uint32_t nq_of(const KArgs& a) { return a.nb * a.parts; }
__device__ __forceinline__ uint32_t* qhead(const KArgs& a, uint32_t q) { return a.st + S_QBASE + QLINES * q * LINE; }
template <class A>
__device__ __forceinline__ uint32_t* qtail(const A& a, uint32_t q) {
  return a.st + S_QBASE + QLINES * q * LINE + LINE;
}
template <class A>
__device__ __forceinline__ uint32_t* qavail(const A& a, uint32_t q) {
  return a.st + S_QBASE + QLINES * q * LINE + 2 * LINE;
}
// in-flight counter i of bucket slot `slot`: light claims of part p at i = p, heavy claims at 16 + p
__device__ __forceinline__ uint32_t* inflight(const KArgs& a, uint32_t slot, uint32_t i) {
  return a.st + s_slotlines(nq_of(a)) + slot * SLOT_WORDS + i * INFL_STRIDE;
}
__device__ __forceinline__ uint32_t* phase(const KArgs& a, uint32_t slot) {
  return a.st + s_slotlines(nq_of(a)) + slot * SLOT_WORDS + SLOT_WORDS - LINE;
}
// heavy edges filed in bucket slot `slot` by claims of part p (u64, on the part's light in-flight line)
__device__ __forceinline__ uint64_t* heavy_ctr(const KArgs& a, uint32_t slot, uint32_t p) {
  return (uint64_t*)(inflight(a, slot, p) + INFL_HEAVY);
}
// termination word of group g, and the wavefronts in group g (workgroups b < grid with b % VER_GROUPS == g)
__device__ __forceinline__ uint64_t* verp(const KArgs& a, uint32_t g) { return (uint64_t*)(a.st + S_VER + g * LINE); }
__device__ __forceinline__ uint32_t group_waves(uint32_t grid, uint32_t g) {
  return (grid + VER_GROUPS - 1 - g) / VER_GROUPS * WAVES_PER_WG;
}
template <class A>
__device__ __forceinline__ void set_error(const A& a, uint32_t code) {
  uint32_t z = 0;
  acas(a.st + S_DONE + 1, z, code);
}
template <class A>
__device__ __forceinline__ bool stopped(const A& a) {
  return (ald(a.st + S_DONE) | ald(a.st + S_DONE + 1)) != 0;
}

// Distance arithmetic.  Int: uint32 add saturating at ADDS_SAT_U32 = 2^32-2, so a path sum that does
// not fit never wraps to a small distance and never reads as unreached (INF = 2^32-1); saturation is
// monotone, so the solution is still the unique least fixed point (the near-far solver and the oracle
// saturate the same way).  Float: IEEE binary32 add (round to nearest even; the build disables
// contraction), compared as uint32 bits, which orders non-negative floats.
template <bool FLT>
__device__ __forceinline__ uint32_t dadd(uint32_t d, uint32_t w) {
  if constexpr (FLT) {
    return __float_as_uint(__uint_as_float(d) + __uint_as_float(w));
  } else {
    const uint32_t s = d + w;
    return (s < d || s == 0xFFFFFFFFu) ? 0xFFFFFFFEu : s;
  }
}
template <bool FLT>
__device__ __forceinline__ uint32_t bucket_key(uint32_t d, float inv_delta) {
  float x = FLT ? __uint_as_float(d) : (float)d;
  x = fminf(x * inv_delta, 4.0e9f);
  return (uint32_t)x;
}

struct WaveStats {
  uint64_t v[NSTATS];
};

// Adds the workgroup's counters to its stats copy: every wavefront leaves its sums (lane 0's values plus
// `extra`, which the caller adds for counters it keeps outside ws) in s_stat, then one atomic per nonzero
// counter.  Every wavefront of the workgroup calls it once (it holds a workgroup barrier).
__device__ __forceinline__ void publish_stats(uint32_t* st, uint64_t (*s_stat)[NSTATS], const WaveStats& ws,
                                              const uint64_t* extra) {
  const uint32_t wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NSTATS; i++) {
    const uint64_t x = wave_sum64(ws.v[i]) + (extra ? extra[i] : 0ull);
    if (lane_id() == 0) s_stat[wv][i] = x;
  }
  __syncthreads();
  if (threadIdx.x < NSTATS) {
    uint64_t x = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; w++) x += s_stat[w][threadIdx.x];
    if (x) afadd64((uint64_t*)(st + stat_copy(blockIdx.x % STAT_COPIES)) + threadIdx.x, x);
  }
}

// Spin until a pool-chunk mapping or entry is written.  Every wait here is on a wavefront that has
// already reserved the slot and is running, so the loops end; the cap and the stop flag turn a broken
// invariant into an error instead of a hang.
constexpr uint32_t SPIN_CAP = 1u << 21;

template <class A>
__device__ __forceinline__ uint32_t wait_tab(const A& a, const uint32_t* tp) {
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
template <class A>
__device__ __forceinline__ uint64_t wait_entry(const A& a, const uint64_t* ep) {
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

// Allocates a pool chunk for the chunk-map word `tp` (called by the reserver of the chunk's first slot)
// and returns it.
template <class A>
__device__ __forceinline__ uint32_t alloc_chunk(const A& a, uint32_t* tp) {
  uint32_t ph = afadd(a.st + S_POOL, 1u);
  if (ph >= a.pool_chunks) {
    set_error(a, DERR_OVERFLOW);
    ph = TAB_BAD;
  }
  ast(tp, ph);
  return ph;
}

// Per-wavefront cache of chunk-map entries in LDS, direct-mapped by queue: (queue << 48 | virtual
// chunk << 24 | physical chunk).  A map entry never changes within a solve once written, so a hit is
// always current; the cache is cleared at kernel start.
constexpr uint32_t TC_SIZE = 256;
__device__ __forceinline__ uint64_t tc_key(uint32_t q, uint32_t j) { return ((uint64_t)q << 48) | ((uint64_t)j << 24); }
template <class A>
__device__ __forceinline__ uint32_t chunk_of(const A& a, uint64_t* tc, uint32_t q, uint32_t j) {
  const uint64_t e = tc[q & (TC_SIZE - 1)];
  if ((e & ~0xFFFFFFull) == tc_key(q, j)) return (uint32_t)(e & 0xFFFFFFull);
  const uint32_t ph = wait_tab(a, a.tab + (size_t)q * a.maxv + j);
  if (ph < TAB_BAD) tc[q & (TC_SIZE - 1)] = tc_key(q, j) | ph;
  return ph;
}

// ---- per-wavefront push buffer ----
// Improvements found while relaxing a claim are collected in LDS (entry + target queue) and written
// to the queues in one flush at the end of the claim, or earlier when the buffer is full.  A flush
// groups the entries by queue, reserves every group's positions at once (one lane per group, so the
// reservations' round trips overlap), writes the entries and publishes every group without waiting
// for the stores: a consumer that claims a position whose entry has not landed yet spins on it
// (wait_entry), and every published position has been reserved by a producer that is writing it.
constexpr uint32_t EPL = 4;           // push-buffer entries per lane (measured: 8 or 16 are slower)
constexpr uint32_t PB_CAP = 64 * EPL;  // entries per wavefront, all handled in one flush round
constexpr uint32_t PB_RELAX = 0x80000000u;  // queue-word flag: apply the entry's atomicMin at the flush

// The fields of the solver arguments that a flush needs, passed by value to the out-of-line flush
// (passing the kernel's argument block by reference would copy it to scratch memory).
struct FlushArgs {
  uint32_t* dist;
  uint32_t* st;
  uint64_t* pool;
  uint32_t* tab;
  uint32_t maxv;
  uint32_t pool_chunks;
  uint32_t nowait;
  uint32_t publish;  // 0: positions are reserved and written but not published (a later step publishes)
};
__device__ __forceinline__ FlushArgs flush_args(const KArgs& a, bool publish = true) {
  return FlushArgs{a.dist, a.st, a.pool, a.tab, a.maxv, a.pool_chunks, a.nowait, publish ? 1u : 0u};
}

struct PushBuf {
  uint64_t* e;   // LDS: entries (dist << 32 | vertex-instance)
  uint32_t* q;   // LDS: target queue
  uint64_t* tc;  // LDS: this wavefront's chunk-map cache
  uint32_t n;    // wave-uniform fill
  bool publish = true;  // false in the pull kernels: k_pull_publish publishes every queue at once
};

// The sub-queue part that the items of the replica at base offset gb go to (the pusher's own part unless
// the solve places replicas by part).
__device__ __forceinline__ uint32_t push_part(const KArgs& a, uint32_t gb, uint32_t part) {
  if (!a.inst_part) return part;
  const uint32_t r = a.n_magic ? __umulhi(gb, a.n_magic) : gb;
  return r & (a.parts - 1);
}
// Target queue of a light-queue push: QM_OWN, the bucket of nd; QM_TAIL, the claim's own bucket (the tail
// of a solve is processed as one bucket); QM_ORDER (ordered split first bucket), the claim's own bucket
// while nd lies below the split point, where a claim sends the vertices of many light edges on to their
// own buckets.  Clamped to the window of the bucket being processed (cb): never before it (such items are
// simply early work) and never so far ahead that the ring slot would alias the behind bucket.
constexpr uint32_t QM_OWN = 0, QM_TAIL = 1, QM_ORDER = 2;
template <bool FLT>
__device__ __forceinline__ uint32_t light_queue(const KArgs& a, float idl, uint32_t nd, uint32_t part, uint32_t cb,
                                                uint32_t mode) {
  uint32_t b = bucket_key<FLT>(nd, idl);
  if (mode == QM_TAIL || (mode == QM_ORDER && b <= a.pull_at)) b = cb;
  b = max(b, cb);
  b = min(b, cb + a.nb - 3);
  return (b & (a.nb - 1)) * a.parts + part;
}

// own_q != ~0: up to 64 distance-lowering entries for queue own_q are kept out of the queues and left
// in pbe[0 .. nl) for the caller to process itself (see the local continuation in k_sssp).
// Everything it touches comes in by value and the counts go out by value: a reference to the caller's
// locals would put them in scratch memory whenever the compiler keeps the flush out of line.
struct FlushOut {
  uint32_t nrel, npush, nl;  // atomicMins issued, entries queued, entries kept for the caller
};
__device__ FlushOut flush_buffer(const FlushArgs a, uint64_t* pbe, uint32_t* pbq, uint64_t* tc, uint32_t pbn,
                                 uint32_t own_q) {
  FlushOut out{0, 0, 0};
  if (pbn == 0) return out;
  wave_sync();
  const uint32_t lane = lane_id();
  uint64_t e[EPL];
  uint32_t q[EPL], g[EPL], off[EPL];
  bool rem[EPL];
#pragma unroll
  for (int j = 0; j < (int)EPL; j++) {
    const uint32_t i = j * 64 + lane;
    rem[j] = i < pbn;
    e[j] = rem[j] ? pbe[i] : 0ull;
    q[j] = rem[j] ? pbq[i] : 0u;
    g[j] = 0;
    off[j] = 0;
  }
  // Improvements found by the edge windows carry PB_RELAX: their atomicMin is issued here, all at once,
  // and only those that lowered the distance are pushed.  With nowait the atomics are not waited for and
  // every improvement is pushed: one that lost to a smaller distance becomes a stale entry, dropped
  // when claimed (a round trip saved per flush against a few more queue entries).
  uint32_t old[EPL];
  bool relax[EPL];
#pragma unroll
  for (int j = 0; j < (int)EPL; j++) {
    relax[j] = rem[j] && (q[j] & PB_RELAX) != 0;
    q[j] &= ~PB_RELAX;
    if (a.nowait) {
      if (relax[j]) (void)afmin(a.dist + (uint32_t)e[j], (uint32_t)(e[j] >> 32));
      old[j] = 0xFFFFFFFFu;  // above every pushed distance: nothing is dropped
    } else {
      old[j] = relax[j] ? afmin(a.dist + (uint32_t)e[j], (uint32_t)(e[j] >> 32)) : 0u;
    }
  }
  uint32_t nrel = 0, npush = 0, nl = 0;
#pragma unroll
  for (int j = 0; j < (int)EPL; j++) {
    nrel += (uint32_t)__popcll(ballot(relax[j]));
    if (relax[j] && old[j] <= (uint32_t)(e[j] >> 32)) rem[j] = false;  // another path got there first
  }
  if (own_q != 0xFFFFFFFFu) {
    // the first 64 improvements for own_q stay with this wavefront (the buffer is free again: its
    // contents are in registers)
    wave_sync();
#pragma unroll
    for (int j = 0; j < (int)EPL; j++) {
      const bool mine = rem[j] && relax[j] && q[j] == own_q;
      const uint64_t m = ballot(mine);
      const uint32_t idx = nl + mbcnt(m);
      if (mine && idx < 64) {
        pbe[idx] = e[j];
        rem[j] = false;
      }
      nl = min(nl + (uint32_t)__popcll(m), 64u);
    }
    wave_sync();
  }
#pragma unroll
  for (int j = 0; j < (int)EPL; j++) npush += (uint32_t)__popcll(ballot(rem[j]));
  out.nrel = nrel;
  out.npush = npush;
  out.nl = nl;
  for (;;) {
    // group the remaining entries by queue, at most 64 groups per pass (group k lives in lane k)
    bool pend[EPL];
    uint32_t gq = 0, gc = 0, ng = 0;
#pragma unroll
    for (int j = 0; j < (int)EPL; j++) pend[j] = false;
    while (ng < 64) {
      uint32_t lq = 0xFFFFFFFFu;
#pragma unroll
      for (int j = 0; j < (int)EPL; j++) {
        const uint64_t m = ballot(rem[j]);
        if (lq == 0xFFFFFFFFu && m) lq = rdlane(q[j], ffs64(m));
      }
      if (lq == 0xFFFFFFFFu) break;
      uint32_t cnt = 0;
#pragma unroll
      for (int j = 0; j < (int)EPL; j++) {
        const bool mine = rem[j] && q[j] == lq;
        const uint64_t m = ballot(mine);
        if (mine) {
          g[j] = ng;
          off[j] = cnt + mbcnt(m);
          rem[j] = false;
          pend[j] = true;
        }
        cnt += (uint32_t)__popcll(m);
      }
      if (lane == ng) {
        gq = lq;
        gc = cnt;
      }
      ng++;
    }
    if (ng == 0) break;
    // reserve every group's positions at once and resolve their (at most two) pool chunks
    uint32_t base = 0, p0 = TAB_BAD, p1 = TAB_BAD;
    if (lane < ng) {
      base = afadd(qtail(a, gq), gc);
      const uint32_t j0 = base >> CH_SHIFT, j1 = (base + gc - 1) >> CH_SHIFT;
      if (j1 < a.maxv) {
        // The reserver of a chunk's first slot allocates it; everyone else looks the chunk up.  Every
        // lane's allocations come before any lane's lookup: a lookup may spin until another
        // wavefront's allocation lands, and that wavefront may in turn be spinning on a chunk that a
        // lane of this wavefront allocates, which must therefore not wait behind the spin.
        uint32_t* t0 = a.tab + (size_t)gq * a.maxv + j0;
        uint32_t* t1 = a.tab + (size_t)gq * a.maxv + j1;
        const bool own0 = (base & (CH - 1)) == 0;
        if (own0) p0 = alloc_chunk(a, t0);
        if (j1 != j0) p1 = alloc_chunk(a, t1);
        if (own0 && p0 < TAB_BAD) tc[gq & (TC_SIZE - 1)] = tc_key(gq, j0) | p0;
        if (j1 != j0 && p1 < TAB_BAD) tc[gq & (TC_SIZE - 1)] = tc_key(gq, j1) | p1;
        if (!own0) p0 = chunk_of(a, tc, gq, j0);
        if (j1 == j0) p1 = p0;
      } else {
        set_error(a, DERR_OVERFLOW);
      }
    }
#pragma unroll
    for (int j = 0; j < (int)EPL; j++) {
      const uint32_t gb = __shfl(base, g[j], 64);
      const uint32_t g0 = __shfl(p0, g[j], 64);
      const uint32_t g1 = __shfl(p1, g[j], 64);
      if (pend[j]) {
        const uint32_t pos = gb + off[j];
        const uint32_t ph = (pos >> CH_SHIFT) == (gb >> CH_SHIFT) ? g0 : g1;
        if (ph != TAB_BAD) ast64(a.pool + ((size_t)ph << CH_SHIFT) + (pos & (CH - 1)), e[j]);
      }
    }
    if (lane < ng && a.publish) afadd(qavail(a, gq), gc);  // publish
  }
  wave_sync();
  return out;
}

// Flushes the push buffer (see flush_buffer); returns the number of entries kept for the caller.
__device__ __forceinline__ uint32_t pb_flush(const FlushArgs a, PushBuf& pb, WaveStats& ws,
                                             uint32_t own_q = 0xFFFFFFFFu) {
  const uint64_t t0 = wall_clock64();
  const FlushOut o = flush_buffer(a, pb.e, pb.q, pb.tc, pb.n, own_q);
  pb.n = 0;
  if (lane_id() == 0) {
    ws.v[ST_T_FLUSH] += wall_clock64() - t0;
    ws.v[ST_RELAX] += o.nrel;
    ws.v[ST_PUSH] += o.npush;
    ws.v[ST_LOCAL] += o.nl;
  }
  return o.nl;
}

// Appends (entry, queue) for every lane with f set; flushes first when the buffer would overflow.
__device__ __forceinline__ void pb_add(const KArgs& a, PushBuf& pb, bool f, uint64_t entry, uint32_t q, WaveStats& ws) {
  const uint64_t m = ballot(f);
  if (m == 0) return;
  const uint32_t c = (uint32_t)__popcll(m);
  if (pb.n + c > PB_CAP) pb_flush(flush_args(a, pb.publish), pb, ws);
  if (f) {
    const uint32_t i = pb.n + mbcnt(m);
    pb.e[i] = entry;
    pb.q[i] = q;
  }
  pb.n += c;
}

// Edges are relaxed in groups of UN 64-edge windows per wavefront (one edge per lane and window; UN is
// 4 or 8, a tuning choice), software-pipelined over the groups (see the edge loop in k_sssp) so that
// the two dependent load rounds -- (col, w), then d[v] -- of consecutive groups overlap.
// Every load is issued unconditionally (inactive lanes read index 0, which exists whenever there is an
// edge to relax): branches around loads make the compiler wait for each window's loads in turn.
template <int UN>
struct EdgeGroup {
  bool act[UN];
  uint32_t e[UN], du[UN], gb[UN], cl[UN], wl[UN];
};

// Issues the (col, w) loads of a located group.
template <int UN>
__device__ __forceinline__ void load_cw(const KArgs& a, EdgeGroup<UN>& g) {
#pragma unroll
  for (int j = 0; j < UN; j++) {
    const uint32_t es = g.act[j] ? g.e[j] : 0u;
    // streamed once per expansion: a non-temporal load keeps the L2 for the d[v] gathers
    const uint64_t x = __builtin_nontemporal_load((const uint64_t*)a.cw + es);
    g.cl[j] = (uint32_t)x;
    g.wl[j] = (uint32_t)(x >> 32);
  }
}

// A group whose d[v] gathers are in flight.
template <int UN>
struct Gathered {
  bool act[UN];
  uint32_t tgt[UN], nd[UN], dv[UN], pp[UN];  // pp: the part its improvements go to
};

// Turns a group whose (col, w) loads have been issued into a gathered group: candidate distances and
// the d[v] pre-filter gathers (plain loads: a stale, larger value only costs an atomic at the flush).
template <bool FLT, int UN>
__device__ __forceinline__ void gather(const KArgs& a, const EdgeGroup<UN>& g, Gathered<UN>& r) {
#pragma unroll
  for (int j = 0; j < UN; j++) {
    r.act[j] = g.act[j];
    r.tgt[j] = g.act[j] ? g.gb[j] + g.cl[j] : 0u;
    r.nd[j] = dadd<FLT>(g.du[j], g.wl[j]);
    r.pp[j] = g.gb[j];
  }
#pragma unroll
  for (int j = 0; j < UN; j++) r.dv[j] = a.dist[r.tgt[j]];
}

// Pair windows (the 8-window form): each lane relaxes the two edges of one 16-byte-aligned pair of the
// light-first rows, loaded with one 16-byte load, so a window covers 128 edges for one (col, w) load
// instruction and two gathers (64-edge windows need two load instructions per 128 edges): 4 pair windows
// per group relax as many edges as 8 plain ones with a quarter fewer vector memory instructions
// (measured: C4a -3%, C3 float -1.5%).  The flattened order counts the pairs that overlap each range; a
// pair's edges outside the range are masked.  The row array carries one pad entry, so the last pair is
// always in bounds.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <int UN>
struct PairGroup {
  bool act[UN];
  uint32_t p[UN], du[UN], gb[UN], lo[UN], hi[UN];  // pair index, source distance, instance base, range
  uint32_t c0[UN], w0[UN], c1[UN], w1[UN];
};
template <int UN>
struct PairGathered {
  bool a0[UN], a1[UN];
  uint32_t t0[UN], t1[UN], n0[UN], n1[UN], d0[UN], d1[UN], pp[UN];  // pp: the pair's replica base
};
template <int UN>
__device__ __forceinline__ void load_pairs(const KArgs& a, PairGroup<UN>& g) {
#pragma unroll
  for (int j = 0; j < UN; j++) {
    const uint32_t ps = g.act[j] ? g.p[j] : 0u;
    const u32x4 x = __builtin_nontemporal_load((const u32x4*)a.cw + ps);
    g.c0[j] = x.x;
    g.w0[j] = x.y;
    g.c1[j] = x.z;
    g.w1[j] = x.w;
  }
}
template <bool FLT, int UN>
__device__ __forceinline__ void gather_pairs(const KArgs& a, const PairGroup<UN>& g, PairGathered<UN>& r) {
#pragma unroll
  for (int j = 0; j < UN; j++) {
    r.a0[j] = g.act[j] && 2 * g.p[j] >= g.lo[j];
    r.a1[j] = g.act[j] && 2 * g.p[j] + 1 < g.hi[j];
    r.t0[j] = r.a0[j] ? g.gb[j] + g.c0[j] : 0u;
    r.t1[j] = r.a1[j] ? g.gb[j] + g.c1[j] : 0u;
    r.n0[j] = dadd<FLT>(g.du[j], g.w0[j]);
    r.n1[j] = dadd<FLT>(g.du[j], g.w1[j]);
    r.pp[j] = g.gb[j];
  }
#pragma unroll
  for (int j = 0; j < UN; j++) {
    r.d0[j] = a.dist[r.t0[j]];
    r.d1[j] = a.dist[r.t1[j]];
  }
}
template <bool FLT, bool REPL, int UN>
__device__ __forceinline__ void finish_pairs(const KArgs& a, float idl, const PairGathered<UN>& r, uint32_t part,
                                             uint32_t cb, bool heavy, uint32_t qmode, PushBuf& pb, WaveStats& ws) {
  bool f0[UN], f1[UN];
  uint32_t nact = 0;
#pragma unroll
  for (int j = 0; j < UN; j++) {
    f0[j] = r.a0[j] && r.n0[j] < r.d0[j];
    f1[j] = r.a1[j] && r.n1[j] < r.d1[j];
    nact += (uint32_t)__popcll(ballot(r.a0[j])) + (uint32_t)__popcll(ballot(r.a1[j]));
  }
  if (lane_id() == 0) {
    ws.v[ST_EDGES] += nact;
    if (heavy) ws.v[ST_HEDGES] += nact;
  }
#pragma unroll
  for (int j = 0; j < UN; j++) {
    const uint32_t pp = REPL ? push_part(a, r.pp[j], part) : part;
    pb_add(a, pb, f0[j], ((uint64_t)r.n0[j] << 32) | r.t0[j], light_queue<FLT>(a, idl, r.n0[j], pp, cb, qmode) | PB_RELAX, ws);
    pb_add(a, pb, f1[j], ((uint64_t)r.n1[j] << 32) | r.t1[j], light_queue<FLT>(a, idl, r.n1[j], pp, cb, qmode) | PB_RELAX, ws);
  }
}

// Consumes a gathered group: improvements over the pre-filter value go to the push buffer, whose
// flush applies the atomicMin.  No store or atomic is issued in the edge loop, so the compiler's
// waits there cover loads only and the pipeline keeps two groups' loads in flight.
template <bool FLT, bool REPL, int UN>
__device__ __forceinline__ void finish(const KArgs& a, float idl, const Gathered<UN>& r, uint32_t part, uint32_t cb,
                                       bool heavy, uint32_t qmode, PushBuf& pb, WaveStats& ws) {
  bool f[UN];
  uint32_t nact = 0;
#pragma unroll
  for (int j = 0; j < UN; j++) {
    f[j] = r.act[j] && r.nd[j] < r.dv[j];
    nact += (uint32_t)__popcll(ballot(r.act[j]));
  }
  if (lane_id() == 0) {
    ws.v[ST_EDGES] += nact;
    if (heavy) ws.v[ST_HEDGES] += nact;
  }
#pragma unroll
  for (int j = 0; j < UN; j++)
    pb_add(a, pb, f[j], ((uint64_t)r.nd[j] << 32) | r.tgt[j],
           light_queue<FLT>(a, idl, r.nd[j], REPL ? push_part(a, r.pp[j], part) : part, cb, qmode) | PB_RELAX, ws);
}

// Appends a bucket-timeline record when tracing is on (called by one lane).
__device__ __forceinline__ void trace_event(const KArgs& a, uint64_t kind, uint32_t bucket, uint32_t prev) {
  if (!a.trace) return;
  const uint64_t i = afadd64(a.trace, 1ull);
  if (i >= a.trace_cap) return;
  uint64_t* r = a.trace + 4 + 4 * i;
  r[0] = wall_clock64();
  r[1] = (kind << 56) | bucket;
  r[2] = prev;
  r[3] = ald(a.st + S_POOL);
}

// Activity histogram (tracing only): when this wavefront's clock enters a new bin, lane 0 adds what its
// counters gained since the last flush to the bin it leaves.  hl (LDS, per wavefront): current bin,
// then the counters at the last flush.
__device__ __forceinline__ void trace_hist(const KArgs& a, uint64_t* hl, uint64_t rel, const uint64_t* tk,
                                           uint64_t edges, uint64_t expanded, uint64_t flush, bool last) {
  const uint64_t bin = min(rel / a.trace_slice, (uint64_t)TRACE_HBINS - 1);
  if (lane_id() == 0 && (bin != hl[0] || last)) {
    const uint32_t copy = (blockIdx.x * WAVES_PER_WG + (threadIdx.x >> 6)) % TRACE_HCOPIES;
    uint64_t* h = a.trace + 4 + 4 * (uint64_t)a.trace_cap + (copy * (uint64_t)TRACE_HBINS + hl[0]) * 8;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      afadd64(h + i, tk[i] - hl[1 + i]);
      hl[1 + i] = tk[i];
    }
    afadd64(h + 4, edges - hl[5]);
    hl[5] = edges;
    afadd64(h + 5, expanded - hl[6]);
    hl[6] = expanded;
    afadd64(h + 6, flush - hl[7]);
    hl[7] = flush;
    hl[0] = bin;
  }
}

constexpr int IPL = 4;  // claimed items per lane: a claim takes at most 64 * IPL items
static_assert(64 * IPL == 256, "the pull's staging counts are per 256 vertex-instances (api.hip)");

// Pushes the reference items of a claim's split ranges into queue q in one go: one reservation for
// all of them, the pool chunks that the reservation starts allocated in one step (consecutive physical
// chunks), every lane writing a share, one publication.  (Through the push buffer, a claim holding
// a few hubs spent tens of flushes on this.)  Item k (= lane * IPL + j) owns reference positions
// [pref[k], pref[k] + nref) of the reservation; rec is its first range record.  pref and rk are this
// wavefront's LDS scratch arrays (free at this point).
__device__ __forceinline__ uint32_t push_refs(const KArgs& a, uint64_t* tc, uint32_t q, uint32_t* pref, uint32_t* rk,
                                           const bool* split, const uint32_t* rec, const uint32_t* nref) {
  const uint32_t lane = lane_id();
  uint32_t mine = 0;
#pragma unroll
  for (int j = 0; j < IPL; j++) mine += split[j] ? nref[j] : 0u;
  const uint32_t incl = wave_incl_scan(mine);
  const uint32_t total = rdlane(incl, 63);
  uint32_t run = incl - mine;
#pragma unroll
  for (int j = 0; j < IPL; j++) {
    pref[lane * IPL + j] = run;
    rk[lane * IPL + j] = split[j] ? rec[j] : 0u;
    run += split[j] ? nref[j] : 0u;
  }
  if (lane == 0) pref[64 * IPL] = total;
  wave_sync();
  // reserve, then allocate the chunks whose first slot is ours (before any lookup, as in the flush),
  // then look up the chunk the reservation starts in if another producer owns it
  uint32_t base = 0, ph0 = TAB_BAD, pfirst = TAB_BAD, first_owned = 0, j1 = 0;
  bool okr = false;
  if (lane == 0) {
    base = afadd(qtail(a, q), total);
    const uint32_t j0 = base >> CH_SHIFT;
    j1 = (base + total - 1) >> CH_SHIFT;
    first_owned = j0 + ((base & (CH - 1)) != 0 ? 1u : 0u);
    okr = j1 < a.maxv;
    if (!okr) set_error(a, DERR_OVERFLOW);
    if (okr && first_owned <= j1) {
      const uint32_t n_own = j1 - first_owned + 1;
      ph0 = afadd(a.st + S_POOL, n_own);
      if (ph0 + n_own > a.pool_chunks) {
        set_error(a, DERR_OVERFLOW);
        okr = false;
      }
    }
  }
  okr = rdlane(okr ? 1u : 0u, 0) != 0;
  base = rdlane(base, 0);
  ph0 = rdlane(ph0, 0);
  first_owned = rdlane(first_owned, 0);
  j1 = rdlane(j1, 0);
  if (okr) {
    for (uint32_t jj = first_owned + lane; jj <= j1; jj += 64)
      ast(a.tab + (size_t)q * a.maxv + jj, ph0 + (jj - first_owned));
    if (lane == 0 && first_owned != (base >> CH_SHIFT)) pfirst = chunk_of(a, tc, q, base >> CH_SHIFT);
    pfirst = rdlane(pfirst, 0);
    if (first_owned != (base >> CH_SHIFT) && pfirst == TAB_BAD) okr = false;
  }
  if (okr) {
    for (uint32_t r0 = 0; r0 < total; r0 += 64) {
      const uint32_t r = r0 + lane;
      // owning item: the last k with pref[k] <= r (an upward bisection of fixed depth)
      uint32_t lo = 0;
#pragma unroll
      for (uint32_t step = 32 * IPL; step >= 1; step >>= 1)
        if (pref[lo + step] <= r) lo += step;
      if (r < total) {
        const uint64_t e = ((uint64_t)REF_TAG << 32) | ((rk[lo] << REF_CHUNK_BITS) + (r - pref[lo]));
        const uint32_t pos = base + r, jc = pos >> CH_SHIFT;
        const uint32_t ph = jc >= first_owned ? ph0 + (jc - first_owned) : pfirst;
        ast64(a.pool + ((size_t)ph << CH_SHIFT) + (pos & (CH - 1)), e);
      }
    }
  }
  // publish (also after an error: the positions are reserved; the solve reports the error)
  if (lane == 0) afadd(qavail(a, q), total);
  wave_sync();
  return total;
}

// Claims up to `chunk` items of queue qq (bucket bb): take from the published count (fetch_sub; a
// shortfall is given back), then fetch_add the head by what was taken, and count the claim in flight
// for the bucket.  Published items are at positions the head has not passed, so every claimed position
// has been reserved by a producer.  Lane 0 works; the outcome is broadcast.
__device__ __forceinline__ bool try_claim(const KArgs& a, uint32_t qq, uint32_t bb, uint32_t nq, uint32_t& h,
                                          uint32_t& cnt, WaveStats& ws) {
  uint32_t ok = 0, hh = 0, cc = 0;
  if (lane_id() == 0) {
    // guided size: a share of what is published (at least claim_min), at most the queue's chunk
    const int32_t chunk = (int32_t)(qq >= nq ? a.chunk_heavy : a.chunk);
    // The in-flight count lands before the head moves past the claimed items (a round trip; a separate
    // counter of claims in progress, updated without waiting, cost more: C4a +2%), so the head advance
    // never sees the items gone while the claim is counted nowhere.
    uint32_t* fl = inflight(a, bb & (a.nb - 1), (qq % a.parts) + (qq >= nq ? MAX_PARTS : 0u));
    const int32_t old = (int32_t)afadd(qavail(a, qq), 0u - (uint32_t)chunk);
    const int32_t avail = max(old, 0);
    const int32_t share = max((int32_t)a.claim_min, avail / (int32_t)a.claim_div);
    const int32_t take = min(min(avail, chunk), share);
    if (take < chunk) afadd(qavail(a, qq), (uint32_t)(chunk - take));
    if (take > 0) {
      afadd(fl, (uint32_t)take);
      mem_wait();  // counted in flight before the head moves past the items
      hh = afadd(qhead(a, qq), (uint32_t)take);
      cc = (uint32_t)take;
      ok = 1;
    } else {
      ws.v[ST_CASFAIL]++;
    }
  }
  if (!rdlane(ok, 0)) return false;
  h = rdlane(hh, 0);
  cnt = rdlane(cc, 0);
  return true;
}

// ------------------------------------------------------------------------------------------------
// the persistent solver
// ------------------------------------------------------------------------------------------------
// An idle wavefront's every 32nd probe uses atomic (cache-bypassing) loads, the others cached ones: stale
// counts only delay finding work (a lost claim forces a fresh probe next time), and the fresh loads are
// same-address traffic on the counters the working wavefronts update (measured against every 8th: C2
// -11%, C3 -4%, C1 -2%; every 64th slowed C1 by a fifth).
#ifndef ADDS_PROBE_PERIOD
#define ADDS_PROBE_PERIOD 32
#endif
// Idle back-off (s_sleep units of 64 clocks; 127 is the largest): an idle wavefront sleeps FIRST for its
// first FIRST_ROUNDS rounds, then MID until round 64, then MAX.  Long sleeps keep idle probes off the queue counters that
// the working wavefronts update (measured: 16 / 64 -> 127 / 127 took C1 -13%, C2 -7%, C3 -2%, C5 -3%).
#ifndef ADDS_SLEEP_FIRST
#define ADDS_SLEEP_FIRST 2
#endif
#ifndef ADDS_SLEEP_FIRST_ROUNDS
#define ADDS_SLEEP_FIRST_ROUNDS 8
#endif
#ifndef ADDS_SLEEP_MID
#define ADDS_SLEEP_MID 127
#endif
#ifndef ADDS_SLEEP_MAX
#define ADDS_SLEEP_MAX 127
#endif
template <bool FLT, bool REPL, int UN>
__global__ void __launch_bounds__(WG_THREADS) k_sssp(KArgs a) {
  __shared__ uint32_t s_pref[WAVES_PER_WG][64 * IPL + 1];
  __shared__ uint32_t s_rs[WAVES_PER_WG][64 * IPL];
  __shared__ uint32_t s_du[WAVES_PER_WG][64 * IPL];
  __shared__ uint32_t s_gb[WAVES_PER_WG][64 * IPL];
  __shared__ uint32_t s_re[WAVES_PER_WG][UN == 8 ? 64 * IPL : 1];
  constexpr bool PAIRS = UN == 8;  // the 8-window form relaxes 4 pair windows per group
  __shared__ uint64_t s_pbe[WAVES_PER_WG][PB_CAP];
  __shared__ uint32_t s_pbq[WAVES_PER_WG][PB_CAP];
  __shared__ uint64_t s_tc[WAVES_PER_WG][TC_SIZE];
  __shared__ uint64_t s_hl[WAVES_PER_WG][8];
  __shared__ uint64_t s_stat[WAVES_PER_WG][NSTATS];

  const uint32_t lane = lane_id();
  const uint32_t wv = threadIdx.x >> 6;
  // Launch l > 0 continues the solve after pull phase l, and only if launch l - 1 requested it (and
  // the solve has not ended or failed); it clears the pull phases' work counters for the next one.
  if (a.launch_idx > 0) {
    if (ald(a.st + PULL_GEN) != a.launch_idx || stopped(a)) return;
    if (blockIdx.x == 0 && threadIdx.x < 6) ast(a.st + S_PULL2 + threadIdx.x, 0u);
  }
  // bucket width: after a pull phase that re-bucketed the pending work (pull_delta), the coarser one;
  // and with resume_all_light every edge is light from then on (a light claim relaxes the whole row and
  // files nothing heavy)
  const bool rebucketed = a.launch_idx > 0 && a.inv_delta_resume > 0.f && ald(a.st + PULL_REBUCKET) != 0;
  const float idl = rebucketed ? a.inv_delta_resume : a.inv_delta;
  const bool all_light = rebucketed && a.resume_all_light;
  // Split first bucket (ADDS_F_PULL_SPLIT): buckets 0 .. pull_at are narrower than the light threshold,
  // so no heavy edge can lower a distance inside them; they run their light phases only (nothing is
  // filed heavy), and at the end of bucket pull_at one pull phase relaxes the heavy edges of all of them.
  // Until then the head stops at pull_at and the solve cannot end.
  const bool defer = a.launch_idx == 0 && a.pull_at != PULL_AT_OFF && a.pull_slots != 0 && ald(a.pull_ok) != 0;
  // Ordered split (order_degree): below the split point only vertices of at least order_degree light
  // edges are expanded in bucket order.  Every push below it goes to the pusher's own bucket; a claim
  // expands the few-edge vertices it finds there at once and sends the many-edge ones on to their own
  // buckets (as early items).  An item's bucket never exceeds its distance's, so bucket pull_at's light
  // phase still ends with every vertex below the split point settled.
  const bool order = defer && a.order_degree != 0;
  PushBuf pbuf{s_pbe[wv], s_pbq[wv], s_tc[wv], 0u};
  for (uint32_t i = lane; i < TC_SIZE; i += 64) s_tc[wv][i] = ~0ull;  // no queue has id 0xFFFF
  wave_sync();
  const uint32_t part = blockIdx.x % a.parts;
  const uint32_t wave_gid = blockIdx.x * WAVES_PER_WG + wv;
  const uint32_t vgrp = blockIdx.x % VER_GROUPS;  // termination group
  const uint32_t nq = nq_of(a);
  const uint32_t nprobe = (a.concur + 3) * a.parts;  // probed queues, one per lane
  WaveStats ws;
#pragma unroll
  for (int i = 0; i < NSTATS; i++) ws.v[i] = 0;
  bool busy = false;  // wave-uniform
  const uint64_t t_start = wall_clock64();
  // histogram time origin (tracing only): the solve's start -- the first launch's, or the time of the
  // timeline's first record for the launches that continue after a pull phase
  const uint64_t t_hist = (a.trace && a.launch_idx > 0) ? __hip_atomic_load(a.trace + 4, __ATOMIC_RELAXED,
                                                                           __HIP_MEMORY_SCOPE_AGENT)
                                                        : t_start;
  uint32_t idle_iter = 0, work_iter = 0;
  bool missed = false;  // the last claim attempt lost: probe with atomic loads next time
  uint64_t tk[4] = {0, 0, 0, 0};  // activity ticks: probe, items, edges, idle (wave-uniform)
  uint32_t sticky_q = 0xFFFFFFFFu, sticky_b = 0;  // queue and bucket of the last full claim (see below)
  // Local continuation: improvements a light claim finds for its own bucket and part (at most 64, one
  // per lane) are processed by the same wavefront as its next claim instead of going through the
  // queue: no reservation, publication, probe or claim round trips for a hop inside a bucket.  They stay
  // counted in flight for that bucket's light phase until this wavefront has processed them.
  uint32_t nlocal = 0, local_b = 0;
  uint64_t local_e = 0;
  if (wave_gid == 0 && lane == 0) trace_event(a, 0, 0, 0);
  if (lane < 8) s_hl[wv][lane] = 0;
  wave_sync();

  for (;;) {
    const uint64_t t0 = wall_clock64();
    if (a.trace) trace_hist(a, s_hl[wv], t0 - t_hist, tk, ws.v[ST_EDGES], ws.v[ST_EXPANDED], ws.v[ST_T_FLUSH], false);
    // ---------------- probe the open queues in parallel ----------------
    // `parts` lanes each: light behind, light head, heavy behind, heavy head (only once the head's
    // phase word is set), light of the concur-1 buckets ahead.  Bit order is the claim priority; each
    // wavefront starts at its own part.
    bool got = false, from_local = false;
    uint32_t q = 0, h = 0, cnt = 0, cb = 0;
    // A pull phase has been requested: this launch ends once every wavefront has finished what it holds
    // (claims in progress complete and push their improvements; local items are processed first).
    if (a.pull_slots && !nlocal &&
        hld(a.st + PULL_GEN, a.cached_probe && (idle_iter & 7) != 7) > a.launch_idx)
      break;
    if (nlocal) {
      got = true;
      from_local = true;
      cb = local_b;
      q = (cb & (a.nb - 1)) * a.parts + part;
      cnt = nlocal;
      nlocal = 0;
    }
    // Sticky claim: after a full claim from the behind or head bucket, the same queue is claimed again
    // directly (no probe) while that bucket is still behind or head.
    if (!got && sticky_q != 0xFFFFFFFFu) {
      const uint32_t c0 = ald(a.st + S_CUR);
      if (sticky_b + 1 >= c0 && sticky_b <= c0) {
        got = try_claim(a, sticky_q, sticky_b, nq, h, cnt, ws);
        if (got) {
          q = sticky_q;
          cb = sticky_b;
        }
      }
      if (!got) sticky_q = 0xFFFFFFFFu;
    }
    const bool cached = a.cached_probe && (idle_iter & (ADDS_PROBE_PERIOD - 1)) != ADDS_PROBE_PERIOD - 1 && !missed;
    missed = false;
    const uint32_t c = got ? 0u : hld(a.st + S_CUR, cached);
    const bool head_heavy_open = !got && hld(phase(a, c & (a.nb - 1)), cached) != 0;
    uint32_t pq = 0, pbk = 0;
    bool pv = false, pf = false;
    {
      const uint32_t grp = lane / a.parts;
      uint32_t j = 0xFFFFFFFFu;  // bucket c-1+j
      bool heavy = false;
      if (grp < 2) {
        j = grp;
      } else if (grp < 4) {
        j = grp - 2;
        heavy = true;
        if (j == 1 && !head_heavy_open) j = 0xFFFFFFFFu;
      } else if (lane < nprobe) {
        j = grp - 2;
      }
      if (!got && j != 0xFFFFFFFFu && (c > 0 || j > 0)) {
        pbk = c - 1 + j;
        pq = (heavy ? nq : 0u) + (pbk & (a.nb - 1)) * a.parts + ((part + lane % a.parts) % a.parts);
        const int32_t av = (int32_t)hld(qavail(a, pq), cached);
        pv = av > 0;
        pf = av >= (int32_t)a.chunk;
      }
    }
    const uint64_t pm = got ? 0ull : ballot(pv);
    // (full_first: queues that hold a full claim are tried before the priority order, so that when the
    // head bucket is down to its last few items the wavefronts do not all contend for them while the
    // open buckets ahead hold plenty)
    const uint64_t pmf = (a.full_first && pm) ? ballot(pf) : 0ull;
    if (lane == 0 && !got) ws.v[ST_PROBE]++;
    if (pm) {
      if (!busy) {
        if (lane == 0) afadd64(verp(a, vgrp), 0xFFFFFFFFull);  // +1 transition, -1 idle
        busy = true;
        mem_wait();
      }
      // Take up to `chunk` from the published count (fetch_sub; a shortfall is given back), then
      // fetch_add the head by what was taken.  Published items are at positions the head has not
      // passed, so every claimed position has been reserved by a producer.
      uint64_t qf = pmf, qm = pm & ~pmf;
      for (int tries = 0; tries < 3 && !got && (qf | qm); tries++) {
        uint32_t L;
        if (qf) {
          L = ffs64(qf);
          qf &= qf - 1;
        } else {
          L = ffs64(qm);
          qm &= qm - 1;
        }
        const uint32_t qq = rdlane(pq, L);
        const uint32_t bb = rdlane(pbk, L);
        if (try_claim(a, qq, bb, nq, h, cnt, ws)) {
          got = true;
          q = qq;
          cb = bb;
          if (cnt == (qq >= nq ? a.chunk_heavy : a.chunk) && bb <= c && a.sticky) {
            sticky_q = qq;
            sticky_b = bb;
          }
        } else {
          missed = true;
        }
      }
    }
    const uint64_t t1 = wall_clock64();
    tk[0] += t1 - t0;

    if (got) {
      idle_iter = 0;
      // The watchdog and the host abort word are also polled while busy, so that a solve that never
      // goes idle can still be ended.
      if ((++work_iter & 255) == 0) {
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
      const bool heavy_claim = q >= nq;
      // Tail: once set (see the head advance), pushes go to the claim's own bucket and nothing is
      // early, so the few items left are processed as one bucket instead of one bucket -- two phase
      // changes -- per distance step.  Read as a hint: a wavefront that sees it late only pushes some
      // items to their own buckets, which is also correct.
      const bool tail_mode = a.tail_items != 0 && hld(a.st + S_TAIL, true) != 0;
      const uint32_t qmode = tail_mode ? QM_TAIL : order ? QM_ORDER : QM_OWN;
      // ---------------- load the claimed entries (IPL per lane: item k = j * 64 + lane) ----------------
      // Positions h .. h+cnt-1 span at most two pool chunks (cnt <= 256 < CH).
      uint32_t p0 = TAB_BAD, p1 = TAB_BAD;
      if (!from_local) {
        const uint32_t j0 = h >> CH_SHIFT, j1 = (h + cnt - 1) >> CH_SHIFT;
        if (lane == 0 && j1 < a.maxv) {
          p0 = chunk_of(a, pbuf.tc, q, j0);
          p1 = j1 != j0 ? chunk_of(a, pbuf.tc, q, j1) : p0;
        }
        p0 = rdlane(p0, 0);
        p1 = rdlane(p1, 0);
        if (p0 == TAB_BAD || p1 == TAB_BAD) {
          if (lane == 0) set_error(a, DERR_TIMEOUT);
          break;  // error path: leave the solve
        }
      }
      bool have[IPL];
      uint32_t gid[IPL], du[IPL];
      if (from_local) {
#pragma unroll
        for (int j = 0; j < IPL; j++) {
          have[j] = j == 0 && lane < cnt;
          gid[j] = (uint32_t)local_e;
          du[j] = (uint32_t)(local_e >> 32);
        }
      } else {
        uint64_t ev[IPL];
        const uint64_t* ep[IPL];
#pragma unroll
        for (int j = 0; j < IPL; j++) {
          const uint32_t k = j * 64 + lane;
          have[j] = k < cnt;
          const uint32_t pos = h + (have[j] ? k : 0u);
          const uint32_t ph = (pos >> CH_SHIFT) == (h >> CH_SHIFT) ? p0 : p1;
          ep[j] = a.pool + ((size_t)ph << CH_SHIFT) + (pos & (CH - 1));
          ev[j] = ald64(ep[j]);
        }
        bool lost = false;
#pragma unroll
        for (int j = 0; j < IPL; j++) {
          if (have[j] && ev[j] == EMPTY64) {  // the producer's store has not landed yet
            ev[j] = wait_entry(a, ep[j]);
            lost |= ev[j] == EMPTY64;
          }
          gid[j] = (uint32_t)ev[j];
          du[j] = (uint32_t)(ev[j] >> 32);
        }
        if (ballot(lost)) {
          if (lane == 0) set_error(a, DERR_TIMEOUT);
          break;  // error path: leave the solve
        }
      }
      // ---------------- resolve each item to an edge range ----------------
      // A vertex item: stale if the vertex already holds a smaller distance, early if its distance
      // belongs to a later bucket (moved there), otherwise its light range (light claim, and the vertex
      // is filed in this bucket's heavy queue) or heavy range (heavy claim).  A reference item: one
      // chunk of a recorded range, stale if the recorded vertex has moved on.
      uint32_t vgid[IPL], rs[IPL], deg[IPL], gbase[IPL];
      bool ok[IPL], is_ref[IPL];
#pragma unroll
      for (int j = 0; j < IPL; j++) {
        is_ref[j] = have[j] && du[j] == REF_TAG;
        vgid[j] = have[j] ? gid[j] : 0u;
        rs[j] = 0;
        deg[j] = 0;
        if (is_ref[j]) {
          const uint32_t* r = a.hrec + (size_t)(gid[j] >> REF_CHUNK_BITS) * HREC_WORDS;
          const uint32_t k = gid[j] & (REF_MAX_CHUNKS - 1);
          vgid[j] = ald(r + 0);
          du[j] = ald(r + 1);
          const uint32_t r0 = ald(r + 2), te = ald(r + 3), r1 = ald(r + 4);
          rs[j] = r0 + k * te;
          deg[j] = min(te, r1 - rs[j]);
        }
      }
      // Coherent loads of the distances: a stale (larger) value would defeat the filter.  Each item's row
      // (start, end, light edges) is one 16-byte record (rinfo): one random load instead of three
      // (measured: C3 int -2%, float -4%).  A partial claim is latency-bound: the records are loaded
      // alongside the distances (unconditionally, at vertex 0 for empty slots) so that all are in
      // flight together.  A full claim (its queue had plenty of work) loads the records of live items
      // only, after the distances: a second round trip, but a stale item -- often most of them --
      // then costs one random load instead of two.
      uint32_t dnow[IPL], r0v[IPL], r1v[IPL], lcv[IPL];
#pragma unroll
      for (int j = 0; j < IPL; j++) gbase[j] = REPL ? (vgid[j] / a.n) * a.n : 0u;
      if (!from_local && cnt == (heavy_claim ? a.chunk_heavy : a.chunk)) {
#pragma unroll
        for (int j = 0; j < IPL; j++) dnow[j] = ald(a.dist + vgid[j]);
#pragma unroll
        for (int j = 0; j < IPL; j++) {
          const uint4 ri = a.rinfo[(have[j] && dnow[j] >= du[j]) ? vgid[j] - gbase[j] : 0u];
          r0v[j] = ri.x;
          r1v[j] = ri.y;
          lcv[j] = ri.z;
        }
      } else {
#pragma unroll
        for (int j = 0; j < IPL; j++) {
          dnow[j] = ald(a.dist + vgid[j]);
          const uint4 ri = a.rinfo[vgid[j] - gbase[j]];
          r0v[j] = ri.x;
          r1v[j] = ri.y;
          lcv[j] = ri.z;
        }
      }
      uint32_t nstale = 0, nreb = 0, nexp = 0, nlate = 0;
      bool early[IPL], has_heavy[IPL];
#pragma unroll
      for (int j = 0; j < IPL; j++) {
        const bool stale = have[j] && dnow[j] < du[j];
        // Early items go to their own bucket -- when it is inside the ring's window.  One beyond it was
        // clamped to the window's end when pushed and would be clamped there again at every head move (a
        // distance 2^31 past the head with delta 1 would take tens of millions of moves), so it is
        // relaxed now: the order of relaxations only affects the work, never the result.
        const uint32_t key = bucket_key<FLT>(du[j], idl);
        early[j] = have[j] && !stale && !is_ref[j] && !heavy_claim && !tail_mode && key > cb &&
                   key <= cb + a.nb - 3 && (!order || key > a.pull_at || lcv[j] >= a.order_degree);
        ok[j] = have[j] && !stale && !early[j];
        has_heavy[j] = false;
        if (!is_ref[j]) {
          const uint32_t lc = all_light ? r1v[j] - r0v[j] : lcv[j];
          rs[j] = heavy_claim ? r0v[j] + lc : r0v[j];
          deg[j] = ok[j] ? (heavy_claim ? r1v[j] - rs[j] : lc) : 0u;
          has_heavy[j] = r1v[j] - r0v[j] > lc;
        }
        if (!ok[j]) deg[j] = 0;
        nstale += (uint32_t)__popcll(ballot(stale));
        nreb += (uint32_t)__popcll(ballot(early[j]));
      }
      uint64_t filed_heavy = 0;  // heavy edges of the vertices filed below (the pull-phase cost model)
#pragma unroll
      for (int j = 0; j < IPL; j++) {
        const bool file_heavy = ok[j] && !heavy_claim && has_heavy[j] && !(defer && cb <= a.pull_at);
        filed_heavy += file_heavy ? r1v[j] - r0v[j] - lcv[j] : 0u;  // (never filed when all_light)
        nexp += (uint32_t)__popcll(ballot(ok[j] && !is_ref[j] && !heavy_claim));
        nlate += (uint32_t)__popcll(ballot(ok[j] && bucket_key<FLT>(du[j], idl) + 1 < cb));
        const uint32_t pp = REPL ? push_part(a, gbase[j], part) : part;
        pb_add(a, pbuf, early[j], ((uint64_t)du[j] << 32) | gid[j], light_queue<FLT>(a, idl, du[j], pp, cb, QM_OWN), ws);
        pb_add(a, pbuf, file_heavy, ((uint64_t)du[j] << 32) | vgid[j], nq + (cb & (a.nb - 1)) * a.parts + pp, ws);
      }
      if (a.pull_slots) {
        filed_heavy = wave_sum64(filed_heavy);
        if (lane == 0 && filed_heavy) afadd64(heavy_ctr(a, cb & (a.nb - 1), part), filed_heavy);
      }
      if (lane == 0) {
        ws.v[ST_STALE] += nstale;
        ws.v[ST_REBUCKET] += nreb;
        ws.v[ST_EXPANDED] += nexp;
        ws.v[ST_LATE] += nlate;
      }
      // ---------------- long ranges become range records + reference items ----------------
      // The references go to the claim's own queue (this bucket, same phase), so every wavefront can
      // take a chunk of task_edges edges.  A range of more than REF_MAX_CHUNKS chunks gets consecutive
      // records of REF_MAX_CHUNKS chunks each, so that reference ids run on across them: id = first
      // record << REF_CHUNK_BITS + chunk.  The records have completed before any reference is written
      // (a consumer reads the record as soon as it sees the reference).
      bool split[IPL];
      uint32_t rec[IPL], nref[IPL];
      bool any_split = false;
#pragma unroll
      for (int j = 0; j < IPL; j++) {
        split[j] = ok[j] && !is_ref[j] && deg[j] > a.hub_degree;
        rec[j] = 0;
        nref[j] = 0;
        if (split[j]) {
          // the claim keeps the range's first chunk for itself (so that a claim of long ranges still
          // has edges to relax at once); the rest goes out as references
          const uint32_t te = a.task_edges, keep = min(te, a.hub_degree);
          nref[j] = (deg[j] - keep + te - 1) / te;
          const uint32_t nrec = (nref[j] + REF_MAX_CHUNKS - 1) / REF_MAX_CHUNKS;
          rec[j] = afadd(a.st + S_HREC, nrec);
          if (rec[j] + nrec > a.hrec_cap) {
            set_error(a, DERR_OVERFLOW);
            split[j] = false;  // relaxed here; the solve reports the overflow and is rerun larger
            nref[j] = 0;
          } else {
            for (uint32_t i = 0; i < nrec; i++) {
              uint32_t* r = a.hrec + (size_t)(rec[j] + i) * HREC_WORDS;
              const uint32_t s0 = rs[j] + keep + i * (REF_MAX_CHUNKS * te);
              ast(r + 0, vgid[j]);
              ast(r + 1, du[j]);
              ast(r + 2, s0);
              ast(r + 3, te);
              ast(r + 4, min(s0 + REF_MAX_CHUNKS * te, rs[j] + deg[j]));
            }
            ws.v[ST_TASKS] += nref[j];
          }
        }
        any_split |= ballot(split[j]) != 0;
      }
      if (any_split) {
        mem_wait();
        const uint32_t nr = push_refs(a, pbuf.tc, q, s_pref[wv], s_rs[wv], split, rec, nref);
        if (lane == 0) ws.v[ST_PUSH] += nr;
#pragma unroll
        for (int j = 0; j < IPL; j++)
          if (split[j]) deg[j] = min(a.task_edges, a.hub_degree);
      }
      const uint64_t t2 = wall_clock64();
      tk[1] += t2 - t1;
      // ---------------- flattened edge windows ----------------
      // Range k = lane * IPL + j occupies [s_pref[k], s_pref[k+1]) of the flattened order.
      uint32_t mysum = 0, unit[IPL];  // flattened units of each range: edges, or pairs in pair mode
#pragma unroll
      for (int j = 0; j < IPL; j++) {
        unit[j] = PAIRS ? (deg[j] ? ((rs[j] + deg[j] + 1) >> 1) - (rs[j] >> 1) : 0u) : deg[j];
        mysum += unit[j];
      }
      const uint32_t incl = wave_incl_scan(mysum);
      const uint32_t total = rdlane(incl, 63);
      if (total) {
        uint32_t run = incl - mysum;
#pragma unroll
        for (int j = 0; j < IPL; j++) {
          const uint32_t k = lane * IPL + j;
          s_pref[wv][k] = run;
          s_rs[wv][k] = rs[j];
          s_du[wv][k] = du[j];
          if (REPL) s_gb[wv][k] = gbase[j];
          if (PAIRS) s_re[wv][k] = rs[j] + deg[j];
          run += unit[j];
        }
        if (lane == 0) s_pref[wv][64 * IPL] = total;
        wave_sync();
        // owner of flattened edge i = last range k with s_pref[k] <= i (its length is then non-zero):
        // an upward bisection of fixed depth, interleaved over the windows
        auto locate = [&](uint32_t base, EdgeGroup<UN>& g) {
          uint32_t idx[UN], lo[UN];
#pragma unroll
          for (int j = 0; j < UN; j++) {
            idx[j] = base + 64 * j + lane;
            g.act[j] = idx[j] < total;
            lo[j] = 0;
          }
#pragma unroll
          for (uint32_t step = 32 * IPL; step >= 1; step >>= 1) {
#pragma unroll
            for (int j = 0; j < UN; j++) {
              const uint32_t mid = lo[j] + step;
              if (s_pref[wv][mid] <= idx[j]) lo[j] = mid;
            }
          }
#pragma unroll
          for (int j = 0; j < UN; j++) {
            g.e[j] = s_rs[wv][lo[j]] + (idx[j] - s_pref[wv][lo[j]]);
            g.du[j] = s_du[wv][lo[j]];
            g.gb[j] = REPL ? s_gb[wv][lo[j]] : 0u;
          }
        };
        if constexpr (PAIRS) {
          constexpr int UP = UN / 2;
          auto locate_p = [&](uint32_t base, PairGroup<UP>& g) {
            uint32_t idx[UP], lo[UP];
#pragma unroll
            for (int j = 0; j < UP; j++) {
              idx[j] = base + 64 * j + lane;
              g.act[j] = idx[j] < total;
              lo[j] = 0;
            }
#pragma unroll
            for (uint32_t step = 32 * IPL; step >= 1; step >>= 1) {
#pragma unroll
              for (int j = 0; j < UP; j++) {
                const uint32_t mid = lo[j] + step;
                if (s_pref[wv][mid] <= idx[j]) lo[j] = mid;
              }
            }
#pragma unroll
            for (int j = 0; j < UP; j++) {
              const uint32_t r0 = s_rs[wv][lo[j]];
              g.p[j] = (r0 >> 1) + (idx[j] - s_pref[wv][lo[j]]);
              g.lo[j] = r0;
              g.hi[j] = s_re[wv][lo[j]];
              g.du[j] = s_du[wv][lo[j]];
              g.gb[j] = REPL ? s_gb[wv][lo[j]] : 0u;
            }
          };
          PairGroup<UP> ga, gb;
          PairGathered<UP> ra, rb;
          locate_p(0, ga);
          load_pairs(a, ga);
          locate_p(64 * UP, gb);
          load_pairs(a, gb);
          gather_pairs<FLT, UP>(a, ga, ra);
          for (uint32_t base = 0; base < total; base += 64 * UP) {
            locate_p(base + 128 * UP, ga);
            load_pairs(a, ga);
            gather_pairs<FLT, UP>(a, gb, rb);
            finish_pairs<FLT, REPL, UP>(a, idl, ra, part, cb, heavy_claim, qmode, pbuf, ws);
            ra = rb;
            gb = ga;
          }
        } else {
        // Three-stage pipeline over groups of 64 * UN edges: while group g is consumed, the
        // gathers of g+1 and the (col, w) loads of g+2 are in flight.  Loads are issued in the order
        // they are waited for, and past the end every window is inactive (its loads read index 0).
        EdgeGroup<UN> ga, gb;
        Gathered<UN> ra, rb;
        locate(0, ga);
        load_cw(a, ga);
        locate(64 * UN, gb);
        load_cw(a, gb);
        gather<FLT, UN>(a, ga, ra);
        for (uint32_t base = 0; base < total; base += 64 * UN) {
          locate(base + 128 * UN, ga);
          load_cw(a, ga);
          gather<FLT, UN>(a, gb, rb);
          finish<FLT, REPL, UN>(a, idl, ra, part, cb, heavy_claim, qmode, pbuf, ws);
          ra = rb;
          gb = ga;
        }
        }
        wave_sync();
      }
      // The claim stops counting as in flight.  Its pushes need not have completed: their tail
      // reservations have (atomics with a result), so the queues already read as non-empty to the head
      // advance and to the termination test.  Improvements kept for local continuation stay counted,
      // under this wavefront's own part of the bucket.
      // (local_cont 2: only after claims of at most a quarter chunk -- the sparse stretches, where a hop's
      // queue round trips are the cost -- while full claims leave their improvements to every wavefront)
      const bool keep_local = a.local_cont == 1 || (a.local_cont == 2 && cnt * 4 <= a.chunk);
      const uint32_t own_q = (keep_local && !heavy_claim) ? (cb & (a.nb - 1)) * a.parts + part : 0xFFFFFFFFu;
      const uint32_t nl = pb_flush(flush_args(a), pbuf, ws, own_q);
      if (nl) {
        local_e = lane < nl ? pbuf.e[lane] : 0ull;
        wave_sync();
        nlocal = nl;
        local_b = cb;
      }
      if (lane == 0) {
        const uint32_t sl = cb & (a.nb - 1);
        const uint32_t ic = (q % a.parts) + (heavy_claim ? MAX_PARTS : 0u);
        if (ic == part) {
          afadd(inflight(a, sl, ic), nl - cnt);
        } else {
          if (nl) {
            afadd(inflight(a, sl, part), nl);
            mem_wait();  // counted under the new part before the claim's count is released
          }
          afadd(inflight(a, sl, ic), 0u - cnt);
        }
      }
      const uint64_t t3 = wall_clock64();
      tk[2] += t3 - t2;
      if (a.trace && total > 16384 && lane == 0) trace_event(a, 4, total, (uint32_t)(t3 - t1));
      if (hld(a.st + S_DONE + 1, true) != 0) break;
      continue;
    }

    // ---------------- no work visible: go idle, test termination, move the head ----------------
    if (busy) {
      mem_wait();  // every push of this wavefront has completed
      if (lane == 0) afadd64(verp(a, vgrp), 1ull);
      busy = false;
      mem_wait();
    }
    // (the stop words with cached loads except every eighth idle round, like the probes)
    {
      const bool fresh = !a.cached_probe || (idle_iter & 7) == 7;
      if ((hld(a.st + S_DONE, !fresh) | hld(a.st + S_DONE + 1, !fresh)) != 0) break;
    }
    // Termination: every group's word reads all idle, every queue is empty, and no group's word has
    // changed since (a wavefront goes busy before it claims, and idle only after its pushes have
    // completed).  The own group's word is checked first, so idle wavefronts mostly read one line.
    if (!defer &&
        (uint32_t)hld64(verp(a, vgrp), a.cached_probe && (idle_iter & 7) != 7) == group_waves(gridDim.x, vgrp)) {
      const uint64_t w1 = lane < VER_GROUPS ? ald64(verp(a, lane)) : 0ull;
      const uint64_t idle = wave_sum64(w1 & 0xFFFFFFFFull);
      if (idle == a.total_waves) {
        mem_wait();
        bool nonempty = false;
        for (uint32_t qq = lane; qq < 2 * nq; qq += 64) nonempty |= ald(qhead(a, qq)) != ald(qtail(a, qq));
        const bool any = ballot(nonempty) != 0;
        mem_wait();
        const uint64_t w2 = lane < VER_GROUPS ? ald64(verp(a, lane)) : 0ull;
        if (!any && ballot(w2 != w1) == 0) {
          if (lane == 0) {
            ast(a.st + S_DONE, 1u);
            trace_event(a, 3, (uint32_t)ald(a.st + S_CUR), 0);
          }
          break;
        }
      }
    }
    // One wavefront at a time (try-lock) inspects the head: idle wavefronts must not flood the counters.
    // (1) When the head bucket's light phase is done (light queues empty, no light claim in flight),
    //     its heavy queues open.  (2) When the behind and head buckets are done -- nothing queued and
    //     nothing in flight, light or heavy -- the head moves to the first bucket ahead that has
    //     anything queued or in flight, like the reference's manager (wl.h:714-723).  Phase words of
    //     slots leaving the window are cleared.
    bool advanced = false, opened = false;
    uint32_t locked = 0;
    // test (a cached hint, fresh every eighth idle round: a stale copy must not keep the lock looking
    // taken), then test-and-set
    if (lane == 0 && hld(a.st + S_ADV, a.cached_probe && (idle_iter & 7) != 7) == 0) {
      uint32_t z = 0;
      locked = acas(a.st + S_ADV, z, 1u) ? 1u : 0u;
    }
    const uint32_t pgen = (rdlane(locked, 0) && a.pull_slots) ? rdlane(lane == 0 ? ald(a.st + PULL_GEN) : 0u, 0) : 0u;
    if (pgen > a.launch_idx) {
      // a pull phase is pending: the head stays where it is until the next launch
      if (lane == 0) ast(a.st + S_ADV, 0u);
    } else if (rdlane(locked, 0)) {
      const uint32_t c2 = ald(a.st + S_CUR);
      const uint32_t hs = c2 & (a.nb - 1), bs = (c2 - 1) & (a.nb - 1);
      // Head/behind state, one lane per part, in three rounds: (1) each queue's head, then its tail;
      // (2) the in-flight counters; (3) the tails again.  A claim is counted in flight before the head
      // passes its items, and stops counting only after its pushes have reserved their tail positions.
      // So a queue that reads empty in (1), idle in (2) and unchanged in (3) had no claim of its bucket
      // running at any time in between (a single pass can miss a claim that starts or ends between its
      // reads, and open a heavy phase or move the head too early).
      bool head_light = false, head_heavy = false, behind_any = false;
      if (lane < a.parts) {
        const bool beh = c2 > 0;
        const uint32_t ql[4] = {hs * a.parts + lane, nq + hs * a.parts + lane, bs * a.parts + lane,
                                nq + bs * a.parts + lane};
        const uint32_t fi[4] = {lane, MAX_PARTS + lane, lane, MAX_PARTS + lane};
        const uint32_t fs[4] = {hs, hs, bs, bs};
        uint32_t h0[4], t0[4], f[4], t1[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const bool use = i < 2 || beh;
          h0[i] = use ? ald(qhead(a, ql[i])) : 0u;
          t0[i] = use ? ald(qtail(a, ql[i])) : 0u;
        }
        mem_wait();
#pragma unroll
        for (int i = 0; i < 4; i++)
          f[i] = (i < 2 || beh) ? ald(inflight(a, fs[i], fi[i])) : 0u;
        mem_wait();
#pragma unroll
        for (int i = 0; i < 4; i++) t1[i] = (i < 2 || beh) ? ald(qtail(a, ql[i])) : 0u;
        bool busy_q[4];
#pragma unroll
        for (int i = 0; i < 4; i++) busy_q[i] = h0[i] < t0[i] || f[i] != 0 || t1[i] != t0[i];
        head_light = busy_q[0];
        head_heavy = busy_q[1];
        behind_any = beh && (busy_q[2] || busy_q[3]);
      }
      const bool hl = ballot(head_light) != 0, hh = ballot(head_heavy) != 0, ba = ballot(behind_any) != 0;
      // Early pull: once at most pull_early items are left at or below the split point, after at least
      // four times as many were seen (the start of a solve has few too), the pull for the split point is
      // asked for now; what is still queued is carried over by the pull kernels.
      bool early = false;
      if (defer && a.pull_early != 0 && c2 < a.pull_at) {
        uint64_t pend = 0;  // (the buckets up to the split point; those beyond it hold unsettled vertices)
        for (uint32_t sl = lane; sl < a.nb; sl += 64) {
          const uint32_t b = c2 - 1 + ((sl - (c2 - 1)) & (a.nb - 1));  // the window's bucket in slot sl
          if (b > a.pull_at) continue;
          for (uint32_t pi = 0; pi < a.parts; pi++) {
            const uint32_t ql = sl * a.parts + pi, qh = nq + ql;
            const uint32_t hl0 = ald(qhead(a, ql)), tl0 = ald(qtail(a, ql));
            const uint32_t hh0 = ald(qhead(a, qh)), th0 = ald(qtail(a, qh));
            pend += (uint64_t)(tl0 - hl0) + (th0 - hh0) + ald(inflight(a, sl, pi)) + ald(inflight(a, sl, MAX_PARTS + pi));
          }
        }
        pend = wave_sum64(pend);
        uint32_t go = 0;
        if (lane == 0) {
          const uint32_t p32 = (uint32_t)min(pend, (uint64_t)0xFFFFFFFFu);
          const uint32_t peak = max(ald(a.st + S_TAIL + 1), p32);
          ast(a.st + S_TAIL + 1, peak);
          if (pend <= a.pull_early && peak >= 4ull * a.pull_early) {
            ast(a.st + PULL_BUCKET, a.pull_at);
            ast(a.st + PULL_IDL, __float_as_uint(idl));
            ast64((uint64_t*)(a.st + PULL_H), 0ull);
            ast(a.st + PULL_CARRY, 1u);
            mem_wait();
            ast(a.st + PULL_GEN, pgen + 1);
            trace_event(a, 5, a.pull_at, 0);
            go = 1;
          }
        }
        early = rdlane(go, 0) != 0;
      }
      if (early) {
        // this launch ends: its wavefronts leave once they hold nothing
      } else if (!hl && rdlane(lane == 0 ? ald(phase(a, hs)) : 1u, 0) == 0) {
        // The head bucket's light phase is done.  Its heavy edges are relaxed from the sending side
        // (open the heavy queues), or -- when the queues hold many heavy edges and a pull slot is left
        // -- from the receiving side by a pull phase: request it and let this launch end.
        bool req = false;
        uint64_t hv = 0;
        if (defer) {
          req = c2 >= a.pull_at;  // the deferred heavy edges: pulled at bucket pull_at
        } else if (a.launch_idx < a.pull_slots && pgen == a.launch_idx &&
            rdlane(lane == 0 ? (ald(a.st + S_TAIL) == 0 && ald(a.pull_ok) != 0 ? 1u : 0u) : 0u, 0) != 0) {
          hv = wave_sum64(lane < a.parts ? ald64(heavy_ctr(a, hs, lane)) : 0ull);
          req = a.pull_eager ? hv > 0 : hv >= a.pull_min;
        }
        if (req && ba) {
          // a claim of the behind bucket can still push into this one: decide on a later pass
        } else if (req) {
          if (lane == 0) {
            ast(a.st + PULL_BUCKET, c2);
            ast(a.st + PULL_IDL, __float_as_uint(idl));
            ast64((uint64_t*)(a.st + PULL_H), hv);
            mem_wait();
            ast(a.st + PULL_GEN, pgen + 1);
            trace_event(a, 5, c2, (uint32_t)min(hv, (uint64_t)0xFFFFFFFFu));
          }
        } else {
          if (lane == 0) {
            ast(phase(a, hs), 1u);
            trace_event(a, 1, c2, c2);
          }
          opened = true;
        }
      }
      if (!early && !hl && !hh && !ba && !(defer && c2 >= a.pull_at)) {
        uint32_t firstj = 0xFFFFFFFFu;
        uint64_t pend = 0;  // items queued or in flight in the window (this lane's buckets)
        for (uint32_t j0 = 0; j0 < a.nb; j0 += 64) {
          const uint32_t jj = j0 + lane;  // bucket c2-1+jj
          bool ne = false;
          if (jj >= 2 && jj < a.nb) {
            const uint32_t slot = (c2 - 1 + jj) & (a.nb - 1);
            for (uint32_t pi = 0; pi < a.parts; pi++) {
              const uint32_t ql = slot * a.parts + pi, qh = nq + ql;
              const uint32_t hl = ald(qhead(a, ql)), tl = ald(qtail(a, ql));  // head first: tl >= hl
              const uint32_t hh2 = ald(qhead(a, qh)), th = ald(qtail(a, qh));
              const uint32_t fl = ald(inflight(a, slot, pi)), fh = ald(inflight(a, slot, MAX_PARTS + pi));
              ne |= hl < tl || hh2 < th || fl != 0 || fh != 0;
              pend += (uint64_t)(tl - hl) + (th - hh2) + fl + fh;
            }
          }
          const uint64_t m = ballot(ne);
          if (m && firstj == 0xFFFFFFFFu) firstj = j0 + ffs64(m);
        }
        pend = wave_sum64(pend);
        // deferred heavy edges: the head stops at pull_at, empty or not
        if (defer) firstj = min(firstj, a.pull_at + 1 - c2);
        if (firstj != 0xFFFFFFFFu) {
          const uint32_t c3 = c2 - 1 + firstj;
          // buckets c2-1 .. c3-2 leave the window; their slots come back later as fresh buckets
          for (uint32_t b = c2 - 1 + lane; b + 1 < c3 && b - (c2 - 1) < a.nb; b += 64)
            if (c2 > 0 || b != c2 - 1) {
              ast(phase(a, b & (a.nb - 1)), 0u);
              if (a.pull_slots)
                for (uint32_t pi = 0; pi < a.parts; pi++) ast64(heavy_ctr(a, b & (a.nb - 1), pi), 0ull);
            }
          // Few items left, and well past the peak of the solve: from now on the solve runs as one
          // bucket, whose heavy queues are open at once (a heavy edge relaxed before its vertex has
          // settled is relaxed again later, at negligible cost this late).  The peak test keeps the
          // start of a solve, whose frontier is small too, ordered: run as one bucket from there, a
          // grid solve does many times the work.  Head moves are serialised (S_ADV), so the peak
          // word is read and written by one wavefront at a time.
          if (lane == 0 && a.tail_items != 0 && !defer) {
            const uint32_t p32 = (uint32_t)min(pend, (uint64_t)0xFFFFFFFFu);
            const uint32_t peak = max(ald(a.st + S_TAIL + 1), p32);
            ast(a.st + S_TAIL + 1, peak);
            if (ald(a.st + S_TAIL) != 0 || (pend <= a.tail_items && pend * 4 <= peak)) {
              ast(a.st + S_TAIL, 1u);
              ast(phase(a, c3 & (a.nb - 1)), 1u);
            }
          }
          mem_wait();
          if (lane == 0) {
            uint32_t cc = c2;
            acas(a.st + S_CUR, cc, c3);
            trace_event(a, 2, c3, c2);
          }
          advanced = true;
        }
      }
      mem_wait();
      if (lane == 0) ast(a.st + S_ADV, 0u);
    }
    if (advanced || opened) {
      // this wavefront changed the head's state: look for the work it opened at once, with a fresh
      // (atomic) probe, without backing off
      idle_iter = 0;
      missed = true;
      tk[3] += wall_clock64() - t1;
      continue;
    }
    // watchdog and host abort
    ++idle_iter;
    if (lane == 0) ws.v[ST_IDLE]++;
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
    if (idle_iter < ADDS_SLEEP_FIRST_ROUNDS) __builtin_amdgcn_s_sleep(ADDS_SLEEP_FIRST);
    else if (idle_iter < 64) __builtin_amdgcn_s_sleep(ADDS_SLEEP_MID);
    else __builtin_amdgcn_s_sleep(ADDS_SLEEP_MAX);
    tk[3] += wall_clock64() - t1;
  }

  // ---------------- exit: publish this wavefront's counters ----------------
  if (a.trace)
    trace_hist(a, s_hl[wv], wall_clock64() - t_hist, tk, ws.v[ST_EDGES], ws.v[ST_EXPANDED], ws.v[ST_T_FLUSH], true);
  if (busy) {
    mem_wait();
    if (lane == 0) afadd64(verp(a, vgrp), 1ull);
  }
  {
    uint64_t extra[NSTATS];
#pragma unroll
    for (int i = 0; i < NSTATS; i++) extra[i] = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) extra[ST_T_PROBE + i] = tk[i];  // wave-uniform
    publish_stats(a.st, s_stat, ws, extra);
  }
}

// ------------------------------------------------------------------------------------------------
// pull phases: a bucket's heavy edges relaxed from the receiving side
// ------------------------------------------------------------------------------------------------
// Delta-stepping relaxes the heavy edges of a settled bucket c once, from their sources (the push
// direction above).  A heavy edge can only lower the distance of a vertex that is not yet settled
// (d[u] >= c * delta and w > delta give d[u] + w > (c + 1) * delta > d[v] for every v of buckets <= c),
// so the same relaxations are covered by letting every unsettled vertex v take min(d[v], d[u] + w) over
// its heavy in-edges (u, w).  On a symmetric graph those are the heavy part of v's own light-first row.
// When bucket c holds most of a power-law graph's edges and the unsettled vertices few (on RMAT-24 with
// a hub source: 931M heavy edges leave bucket 0, 29M heavy edges enter the vertices still unsettled),
// this scans a few percent of the edges, and every random gather is one the result needs.  The vertices
// it lowers are queued in their buckets like any improvement; distances and their uniqueness are those
// of the push direction (each is a real path length, and every relaxation the push direction would make
// is made or cannot lower anything).
//
// k_pull_count sums the heavy in-degrees of the unsettled vertices; k_pull_decide (one workgroup) takes
// the cheaper direction (opening the heavy queues for the next launch when pushing costs less); k_pull_a
// relaxes rows of up to PULL_LONG heavy edges, 256 vertex-instances per wavefront, handing longer rows to
// k_pull_b as tasks of at most PULL_TASK edges, each relaxed by one wavefront and applied with an
// atomicMin.
//
// Re-bucketing (pull_delta): what is left after such a pull is the sparse rest of the graph, whose many
// narrow buckets each cost a few claim round trips for little work.  The first pull may therefore move
// the solve to a wider bucket: k_pull_decide empties every queue, clears the bucket ring's phase words
// and heavy-edge counters and sets the head to the first wide bucket; k_pull_a then queues every
// unsettled vertex that has a distance (the dropped items were all of such vertices: every vertex of a
// bucket up to c is settled, and c's own heavy items are covered by the pull) at the wide width, and the
// launches that follow use it.

template <class A>
__device__ __forceinline__ bool pull_requested(const A& a, uint32_t slot) {
  return ald(a.st + PULL_GEN) == slot && !stopped(a);
}
// the bucket width of the launch that requested the pull (its bucket index c is in these units)
__device__ __forceinline__ float pull_req_idl(const KArgs& a) { return __uint_as_float(ald(a.st + PULL_IDL)); }
template <bool FLT>
__host__ __device__ constexpr uint32_t inf_word() {
  return FLT ? 0x7F800000u : 0xFFFFFFFFu;
}
// not settled at bucket c: unreached, or its distance lies beyond bucket c
template <bool FLT>
__device__ __forceinline__ bool unsettled(uint32_t d, uint32_t c, float inv_delta) {
  return d == inf_word<FLT>() || bucket_key<FLT>(d, inv_delta) > c;
}
// A path through an unreached vertex is no path (the saturating integer add would turn it into 2^32-2).
template <bool FLT>
__device__ __forceinline__ uint32_t pull_cand(uint32_t du, uint32_t w) {
  return du == inf_word<FLT>() ? du : dadd<FLT>(du, w);
}
// Pushing costs one gathered edge per queued heavy edge (h); pulling costs its in-edges (a gathered edge
// each, on short rows: counted twice), a row record per unsettled vertex and a distance per vertex.
__device__ __forceinline__ bool pull_pays(uint64_t h, uint64_t in_edges, uint64_t unsettled_v, uint64_t total_n) {
  return 2 * in_edges + unsettled_v / 2 + total_n / 16 <= h;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, d, 64));
  return x;
}

constexpr uint32_t PULL_COUNT_THREADS = 1024;  // k_pull_count: few large workgroups, two atomics each
template <bool FLT, bool REPL>
__global__ void __launch_bounds__(PULL_COUNT_THREADS) k_pull_count(KArgs a, uint32_t slot) {
  __shared__ uint64_t s_red[PULL_COUNT_THREADS / 64][2];
  if (!pull_requested(a, slot) || (slot == 1 && a.pull_at != PULL_AT_OFF)) return;  // forced: no estimate
  const uint32_t c = ald(a.st + PULL_BUCKET);
  const float idl = pull_req_idl(a);
  uint64_t in_edges = 0, uns = 0;
  for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < a.total_n; x += (uint64_t)gridDim.x * blockDim.x) {
    if (unsettled<FLT>(a.dist[x], c, idl)) {
      const uint4 ri = a.rinfo[REPL ? (uint32_t)x % a.n : (uint32_t)x];
      in_edges += ri.y - ri.x - ri.z;
      uns++;
    }
  }
  in_edges = wave_sum64(in_edges);
  uns = wave_sum64(uns);
  if (lane_id() == 0) {
    s_red[threadIdx.x >> 6][0] = in_edges;
    s_red[threadIdx.x >> 6][1] = uns;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    uint64_t x = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; w++) x += s_red[w][threadIdx.x];
    if (x) afadd64((uint64_t*)(a.st + S_PULL2 + 2 + 2 * threadIdx.x), x);
  }
}

constexpr int PULL_UN = 8;  // 64-edge windows per wavefront in flight in the pull kernels
constexpr uint32_t PULL_HEAD = 4;  // heavy edges per vertex scanned one at a time, in band order
// a weight or distance word as the float that heavy_band takes (int words converted like the split does)
template <bool FLT>
__device__ __forceinline__ float wfloat(uint32_t x) {
  return FLT ? __uint_as_float(x) : (float)x;
}

// One workgroup: pull or push, and the queue and ring state for the next launch.
__global__ void __launch_bounds__(1024) k_pull_decide(KArgs a, uint32_t slot) {
  if (!pull_requested(a, slot)) return;
  const uint32_t c = ald(a.st + PULL_BUCKET);
  // the first pull of a split first bucket is forced: its heavy edges were never queued for pushing
  const bool pull = a.pull_eager || (slot == 1 && a.pull_at != PULL_AT_OFF) || pull_pays(ald64((const uint64_t*)(a.st + PULL_H)),
                                              ald64((const uint64_t*)(a.st + S_PULL2 + 2)),
                                              ald64((const uint64_t*)(a.st + S_PULL2 + 4)), a.total_n);
  const bool rebucket = pull && a.inv_delta_resume > 0.f && ald(a.st + PULL_DONE) == 0;
  const uint32_t nq = nq_of(a), cs = c & (a.nb - 1);
  if (rebucket) {
    // every queue emptied (positions stay reserved, so the chunk maps and the cleanup are unaffected),
    // every slot's phase word and heavy-edge counters cleared; in-flight counts are zero (no launch runs).
    // An early request leaves items queued: each queue's range is kept for k_pull_carry.
    const bool carry = ald(a.st + PULL_CARRY) != 0 && a.carry;
    for (uint32_t q = threadIdx.x; q < 2 * nq; q += blockDim.x) {
      const uint32_t t = ald(qtail(a, q));
      if (carry) a.carry[q] = make_uint2(ald(qhead(a, q)), t);
      ast(qhead(a, q), t);
      ast(qavail(a, q), 0u);
    }
    if (threadIdx.x == 0) ast(a.st + S_TAIL + 1, 0u);  // the peak of the launches that follow starts afresh
    for (uint32_t i = threadIdx.x; i < a.nb * a.parts; i += blockDim.x) {
      const uint32_t sl = i / a.parts, p = i % a.parts;
      ast64(heavy_ctr(a, sl, p), 0ull);
      if (p == 0) ast(phase(a, sl), 0u);
    }
  } else if (threadIdx.x == 0 && pull) {
    // the bucket's heavy queues are done with: their edges are pulled
    for (uint32_t p = 0; p < a.parts; p++) {
      const uint32_t q = nq + cs * a.parts + p;
      ast(qhead(a, q), ald(qtail(a, q)));
      ast(qavail(a, q), 0u);
    }
  }
  // every thread has read PULL_DONE (rebucket above) before it is raised
  __syncthreads();
  if (threadIdx.x == 0) {
    ast(a.st + PULL_OUTCOME, pull ? 1u : 2u);
    if (pull) ast(a.st + PULL_DONE, ald(a.st + PULL_DONE) + 1);
    if (rebucket) {
      // the first wide bucket at or below every unsettled distance ((c + 1) * delta and up)
      const float ratio = a.inv_delta_resume / pull_req_idl(a);
      const uint32_t nh = (uint32_t)fmaxf(floorf((float)(c + 1) * ratio) - 1.f, 0.f);
      ast(a.st + PULL_NEWHEAD, nh);
      ast(a.st + S_CUR, nh);
      ast(a.st + PULL_REBUCKET, 1u);
    } else {
      ast(phase(a, cs), 1u);  // pushed: the next launch relaxes the heavy queues; pulled: they are empty
    }
    trace_event(a, 6, c, pull ? (rebucket ? 3u : 1u) : 2u);
  }
}

template <bool FLT, bool REPL>
__global__ void __launch_bounds__(256) k_pull_a(KArgs a, uint32_t slot) {
  __shared__ uint32_t s_pref[WAVES_PER_WG][64 * IPL + 1];
  __shared__ uint32_t s_rs[WAVES_PER_WG][64 * IPL];
  __shared__ uint32_t s_best[WAVES_PER_WG][64 * IPL];
  __shared__ uint32_t s_gb[WAVES_PER_WG][REPL ? 64 * IPL : 1];
  __shared__ uint64_t s_pbe[WAVES_PER_WG][PB_CAP];
  __shared__ uint32_t s_pbq[WAVES_PER_WG][PB_CAP];
  __shared__ uint64_t s_tc[WAVES_PER_WG][TC_SIZE];
  __shared__ uint64_t s_stat[WAVES_PER_WG][NSTATS];
  __shared__ uint32_t s_chunk[WAVES_PER_WG][64];
  if (!pull_requested(a, slot) || ald(a.st + PULL_OUTCOME) != 1u) return;
  const uint32_t c = ald(a.st + PULL_BUCKET);
  const float idl = pull_req_idl(a);
  // pushes: at the requester's width behind head c, or (re-bucketing) at the wide width behind its new
  // head, and then every unsettled vertex with a distance is queued, improved or not
  const bool rebucket = ald(a.st + PULL_REBUCKET) != 0;
  const float pidl = rebucket ? a.inv_delta_resume : idl;
  const uint32_t pc = rebucket ? ald(a.st + PULL_NEWHEAD) : c;
  const uint32_t lane = lane_id();
  const uint32_t wv = threadIdx.x >> 6;
  PushBuf pb{s_pbe[wv], s_pbq[wv], s_tc[wv], 0u, false};
  for (uint32_t i = lane; i < TC_SIZE; i += 64) s_tc[wv][i] = ~0ull;
  wave_sync();
  const uint32_t part = blockIdx.x % a.parts;
  WaveStats ws;
#pragma unroll
  for (int i = 0; i < NSTATS; i++) ws.v[i] = 0;
  constexpr uint32_t INF = inf_word<FLT>();
  // chunks of 256 vertex-instances dealt round-robin over the wavefronts (no shared counter: one address
  // taking an atomic per chunk from every wavefront costs more than the scan itself)
  const uint32_t nwaves = gridDim.x * WAVES_PER_WG;
  const uint32_t wave_gid = blockIdx.x * WAVES_PER_WG + wv;
  // Bulk queueing (re-bucketing pulls): the vertices queued at the new head in this wavefront's own part
  // -- nearly all of them -- are staged per chunk and reserved and written at the end in one run, instead
  // of a flush (a reservation per queue and a chunk-map round trip) per 256 entries.  A wavefront whose
  // chunks could need more than 62 pool chunks queues through the push buffer.
  const uint64_t nch = (a.total_n + 64 * IPL - 1) / (64 * IPL);
  const uint64_t my_chunks = wave_gid < nch ? (nch - wave_gid + nwaves - 1) / nwaves : 0;
  const bool bulk = rebucket && a.pstage && my_chunks * 64 * IPL <= 62ull * CH;
  const uint32_t qbulk = (pc & (a.nb - 1)) * a.parts + part;
  uint32_t stot = 0;  // entries staged by this wavefront
  for (uint64_t ch = wave_gid; ch * (64 * IPL) < a.total_n; ch += nwaves) {
    const uint32_t base = (uint32_t)(ch * (64 * IPL));
    // items: vertex-instance x = base + j * 64 + lane (coalesced); range index k = lane * IPL + j
    uint32_t x[IPL], dv[IPL], rs[IPL], deg[IPL], gb[IPL], best[IPL], he[IPL];
    bool uns[IPL];
    // distances and the first two heavy edges together (coalesced); most vertices need nothing else
    uint4 ph[IPL];
#pragma unroll
    for (int j = 0; j < IPL; j++) {
      x[j] = base + j * 64 + lane;
      const bool in = x[j] < a.total_n;
      gb[j] = REPL ? x[j] - x[j] % a.n : 0u;
      dv[j] = in ? a.dist[x[j]] : 0u;
      ph[j] = a.phead[in ? x[j] - gb[j] : 0u];
    }
    {
      // the second edge is gathered only if its band is not above the best value after the first: an
      // extra round trip, but most second gathers (each a cache line) are saved
      bool g1[IPL], g2[IPL];
      uint32_t d1[IPL], d2[IPL];
#pragma unroll
      for (int j = 0; j < IPL; j++) {
        uns[j] = x[j] < a.total_n && unsettled<FLT>(dv[j], c, idl);
        best[j] = dv[j];
        const uint32_t bd = heavy_band(wfloat<FLT>(dv[j]), a.band_lo, a.band_inv);
        g1[j] = uns[j] && ph[j].x != PHEAD_NONE && heavy_band(wfloat<FLT>(ph[j].y), a.band_lo, a.band_inv) <= bd;
        d1[j] = g1[j] ? a.dist[gb[j] + ph[j].x] : INF;
      }
#pragma unroll
      for (int j = 0; j < IPL; j++) {
        if (g1[j]) best[j] = min(best[j], pull_cand<FLT>(d1[j], ph[j].y));
        g2[j] = uns[j] && ph[j].z != PHEAD_NONE &&
                heavy_band(wfloat<FLT>(ph[j].w), a.band_lo, a.band_inv) <=
                    heavy_band(wfloat<FLT>(best[j]), a.band_lo, a.band_inv);
        d2[j] = g2[j] ? a.dist[gb[j] + ph[j].z] : INF;
      }
      uint32_t nscan = 0;
      bool more[IPL];
#pragma unroll
      for (int j = 0; j < IPL; j++) {
        if (g2[j]) best[j] = min(best[j], pull_cand<FLT>(d2[j], ph[j].w));
        nscan += (uint32_t)g1[j] + (uint32_t)g2[j];
        // a third heavy edge (band at least the second's) can only matter if the second one's band is not
        // above the best value's
        more[j] = g2[j] && heavy_band(wfloat<FLT>(ph[j].w), a.band_lo, a.band_inv) <=
                               heavy_band(wfloat<FLT>(best[j]), a.band_lo, a.band_inv);
      }
      nscan = (uint32_t)wave_sum64(nscan);
      if (lane == 0) {
        ws.v[ST_EDGES] += nscan;
        ws.v[ST_PULL_EDGES] += nscan;
      }
#pragma unroll
      for (int j = 0; j < IPL; j++) {
        const uint4 ri = more[j] ? a.rinfo[x[j] - gb[j]] : make_uint4(0u, 0u, 0u, 0u);
        rs[j] = ri.x + ri.z + 2;
        he[j] = more[j] ? ri.y : rs[j];
      }
    }
    // The rest of the heavy part from its third edge, in band order, one edge per vertex per step: an edge
    // of a higher band than the best value so far weighs more than it (and so does every edge after it),
    // which ends the vertex's scan.  Most vertices end within PULL_HEAD more edges; what is left of the
    // others is relaxed in full below.  The next edge's (col, w) is loaded alongside each step's gather.
    {
      bool act[IPL];
      uint64_t cwn[IPL];
#pragma unroll
      for (int j = 0; j < IPL; j++) {
        act[j] = rs[j] < he[j];
        cwn[j] = act[j] ? __builtin_nontemporal_load((const uint64_t*)a.cw + rs[j]) : 0ull;
      }
      uint32_t nscan = 0;
      for (uint32_t step = 0; step < PULL_HEAD; step++) {
        bool any = false;
#pragma unroll
        for (int j = 0; j < IPL; j++) any |= act[j];
        if (!ballot(any)) break;
        uint32_t u[IPL], ww[IPL], du[IPL];
        uint64_t cw2[IPL];
#pragma unroll
        for (int j = 0; j < IPL; j++) {
          u[j] = (uint32_t)cwn[j];
          ww[j] = (uint32_t)(cwn[j] >> 32);
          if (act[j] && heavy_band(wfloat<FLT>(ww[j]), a.band_lo, a.band_inv) >
                            heavy_band(wfloat<FLT>(best[j]), a.band_lo, a.band_inv)) {
            act[j] = false;
            he[j] = rs[j];  // nothing left to relax
          }
        }
#pragma unroll
        for (int j = 0; j < IPL; j++) {
          du[j] = act[j] ? a.dist[gb[j] + u[j]] : INF;
          const bool more = act[j] && rs[j] + 1 < he[j] && step + 1 < PULL_HEAD;
          cw2[j] = more ? __builtin_nontemporal_load((const uint64_t*)a.cw + rs[j] + 1) : 0ull;
        }
#pragma unroll
        for (int j = 0; j < IPL; j++) {
          if (act[j]) {
            best[j] = min(best[j], pull_cand<FLT>(du[j], ww[j]));
            rs[j]++;
            nscan++;
            act[j] = rs[j] < he[j];
          }
          cwn[j] = cw2[j];
        }
      }
      nscan = (uint32_t)wave_sum64(nscan);
      if (lane == 0) {
        ws.v[ST_EDGES] += nscan;
        ws.v[ST_PULL_EDGES] += nscan;
      }
    }
#pragma unroll
    for (int j = 0; j < IPL; j++) deg[j] = he[j] - rs[j];
    // rows longer than PULL_LONG become tasks for k_pull_b (all of this claim's, or, when the task list
    // is full, none: then they are relaxed here and the reserved positions hold empty tasks)
    {
      uint32_t nt[IPL], mine = 0;
#pragma unroll
      for (int j = 0; j < IPL; j++) {
        nt[j] = deg[j] > PULL_LONG ? (deg[j] + PULL_TASK - 1) / PULL_TASK : 0u;
        mine += nt[j];
      }
      const uint32_t incl = wave_incl_scan(mine);
      const uint32_t total = rdlane(incl, 63);
      if (total) {
        uint32_t tb = 0;
        if (lane == 0) tb = afadd(a.st + S_PULL2, total);
        tb = rdlane(tb, 0);
        const bool fits = (uint64_t)tb + total <= a.ptask_cap;
        uint32_t pos = tb + incl - mine;
#pragma unroll
        for (int j = 0; j < IPL; j++) {
          for (uint32_t t = 0; t < nt[j]; t++, pos++) {
            if (pos >= a.ptask_cap) break;
            const uint32_t e0 = rs[j] + t * PULL_TASK;
            a.ptask[pos] = fits ? make_uint4(x[j], e0, min(e0 + PULL_TASK, rs[j] + deg[j]), 0u)
                                : make_uint4(0u, 0u, 0u, 0u);
          }
          if (fits && nt[j]) deg[j] = 0;
        }
      }
    }
    // flatten the remaining rows (range k = lane * IPL + j owns [s_pref[k], s_pref[k+1]))
    uint32_t mysum = 0;
#pragma unroll
    for (int j = 0; j < IPL; j++) mysum += deg[j];
    const uint32_t incl = wave_incl_scan(mysum);
    const uint32_t total = rdlane(incl, 63);
    uint32_t run = incl - mysum;
#pragma unroll
    for (int j = 0; j < IPL; j++) {
      const uint32_t k = lane * IPL + j;
      s_pref[wv][k] = run;
      s_rs[wv][k] = rs[j];
      s_best[wv][k] = best[j];
      if (REPL) s_gb[wv][k] = gb[j];
      run += deg[j];
    }
    if (lane == 0) s_pref[wv][64 * IPL] = total;
    wave_sync();
    for (uint32_t eb = 0; eb < total; eb += 64 * PULL_UN) {
      uint32_t idx[PULL_UN], lo[PULL_UN], u[PULL_UN], w[PULL_UN], du[PULL_UN];
#pragma unroll
      for (int j = 0; j < PULL_UN; j++) {
        idx[j] = eb + 64 * j + lane;
        lo[j] = 0;
      }
#pragma unroll
      for (uint32_t step = 32 * IPL; step >= 1; step >>= 1) {
#pragma unroll
        for (int j = 0; j < PULL_UN; j++)
          if (s_pref[wv][lo[j] + step] <= idx[j]) lo[j] += step;
      }
#pragma unroll
      for (int j = 0; j < PULL_UN; j++) {
        const uint32_t e = idx[j] < total ? s_rs[wv][lo[j]] + (idx[j] - s_pref[wv][lo[j]]) : 0u;
        const uint64_t cw = __builtin_nontemporal_load((const uint64_t*)a.cw + e);
        u[j] = (uint32_t)cw;
        w[j] = (uint32_t)(cw >> 32);
      }
#pragma unroll
      for (int j = 0; j < PULL_UN; j++) du[j] = a.dist[(REPL ? s_gb[wv][lo[j]] : 0u) + (idx[j] < total ? u[j] : 0u)];
#pragma unroll
      for (int j = 0; j < PULL_UN; j++) {
        const uint32_t cand = idx[j] < total ? pull_cand<FLT>(du[j], w[j]) : INF;
        if (cand < INF)
          __hip_atomic_fetch_min(&s_best[wv][lo[j]], cand, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    wave_sync();
    uint32_t nimp = 0, sc = 0;
#pragma unroll
    for (int j = 0; j < IPL; j++) {
      const uint32_t bj = s_best[wv][lane * IPL + j];
      const bool imp = uns[j] && bj < dv[j];
      if (imp) ast(a.dist + x[j], bj);
      nimp += (uint32_t)__popcll(ballot(imp));
      const uint32_t pp = REPL ? push_part(a, gb[j], part) : part;
      const bool q = imp || (rebucket && uns[j] && bj < INF);
      const uint32_t lq = light_queue<FLT>(a, pidl, bj, pp, pc, QM_OWN);
      const bool st = bulk && q && lq == qbulk;
      const uint64_t sm = ballot(st);
      if (st) a.pstage[(size_t)base + sc + mbcnt(sm)] = ((uint64_t)bj << 32) | x[j];
      sc += (uint32_t)__popcll(sm);
      pb_add(a, pb, q && !st, ((uint64_t)bj << 32) | x[j], lq, ws);
    }
    wave_sync();
    if (lane == 0) {
      ws.v[ST_EDGES] += total;
      ws.v[ST_PULL_EDGES] += total;
      ws.v[ST_PULL_IMPROVED] += nimp;
      if (bulk) a.pcnt[ch] = sc;
    }
    stot += sc;
  }
  pb_flush(flush_args(a, false), pb, ws);
  if (bulk && stot) {
    // reserve this wavefront's run, allocate the pool chunks whose first slot it holds (all of them
    // before any lookup: see flush_buffer), look up the one it shares with the run before it, then copy
    uint32_t rb = 0;
    if (lane == 0) rb = afadd(qtail(a, qbulk), stot);
    rb = rdlane(rb, 0);
    const uint32_t j0 = rb >> CH_SHIFT, j1 = (rb + stot - 1) >> CH_SHIFT;  // j1 - j0 < 64 (62 chunks + 2)
    const uint32_t jl = j0 + lane;
    uint32_t ph = TAB_BAD;
    if (jl <= j1) {
      if (jl >= a.maxv) {
        set_error(a, DERR_OVERFLOW);
      } else if ((jl << CH_SHIFT) >= rb) {
        ph = alloc_chunk(a, a.tab + (size_t)qbulk * a.maxv + jl);
      }
    }
    if (jl <= j1 && jl < a.maxv && (jl << CH_SHIFT) < rb) ph = wait_tab(a, a.tab + (size_t)qbulk * a.maxv + jl);
    s_chunk[wv][lane] = ph;
    // the staged entries and counts were written by other lanes of this wavefront
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    wave_sync();
    uint32_t off = 0;
    for (uint64_t ch = wave_gid; ch * (64 * IPL) < a.total_n; ch += nwaves) {
      const uint32_t cnt = a.pcnt[ch];
      for (uint32_t i = lane; i < cnt; i += 64) {
        const uint32_t pos = rb + off + i;
        const uint32_t pch = s_chunk[wv][(pos >> CH_SHIFT) - j0];
        if (pch < TAB_BAD) ast64(a.pool + ((size_t)pch << CH_SHIFT) + (pos & (CH - 1)), a.pstage[ch * (64 * IPL) + i]);
      }
      off += cnt;
    }
    if (lane == 0) ws.v[ST_PUSH] += stot;
  }
  publish_stats(a.st, s_stat, ws, nullptr);
}

template <bool FLT, bool REPL>
__global__ void __launch_bounds__(256) k_pull_b(KArgs a, uint32_t slot) {
  __shared__ uint64_t s_pbe[WAVES_PER_WG][PB_CAP];
  __shared__ uint32_t s_pbq[WAVES_PER_WG][PB_CAP];
  __shared__ uint64_t s_tc[WAVES_PER_WG][TC_SIZE];
  __shared__ uint64_t s_stat[WAVES_PER_WG][NSTATS];
  if (!pull_requested(a, slot) || ald(a.st + PULL_OUTCOME) != 1u) return;
  const bool rebucket = ald(a.st + PULL_REBUCKET) != 0;
  const float pidl = rebucket ? a.inv_delta_resume : pull_req_idl(a);
  const uint32_t pc = rebucket ? ald(a.st + PULL_NEWHEAD) : ald(a.st + PULL_BUCKET);
  const uint32_t ntask = min(ald(a.st + S_PULL2), a.ptask_cap);
  const uint32_t lane = lane_id();
  const uint32_t wv = threadIdx.x >> 6;
  PushBuf pb{s_pbe[wv], s_pbq[wv], s_tc[wv], 0u, false};
  for (uint32_t i = lane; i < TC_SIZE; i += 64) s_tc[wv][i] = ~0ull;
  wave_sync();
  const uint32_t part = blockIdx.x % a.parts;
  WaveStats ws;
#pragma unroll
  for (int i = 0; i < NSTATS; i++) ws.v[i] = 0;
  constexpr uint32_t INF = inf_word<FLT>();
  // tasks dealt round-robin over the wavefronts (a shared claim counter would take one same-address
  // atomic per wavefront even when there are few tasks)
  for (uint32_t t = blockIdx.x * WAVES_PER_WG + wv; t < ntask; t += gridDim.x * WAVES_PER_WG) {
    const uint4 tk = a.ptask[t];
    if (tk.y >= tk.z) continue;  // an empty task (the list was full)
    const uint32_t gb = REPL ? tk.x - tk.x % a.n : 0u;
    // The range is in band order: an edge of a higher band than this lane's bound (the vertex's distance
    // when the task starts, or a smaller candidate found since) cannot lower it, nor can any edge after it,
    // so the lane skips its gather, and the task ends once no lane of a window group has an edge left.
    const uint32_t d0 = a.dist[tk.x];
    uint32_t best = INF, scanned = 0;
    for (uint32_t e0 = tk.y; e0 < tk.z; e0 += 64 * PULL_UN) {
      uint32_t u[PULL_UN], w[PULL_UN], du[PULL_UN];
      bool act[PULL_UN];
      const uint32_t lim = heavy_band(wfloat<FLT>(min(best, d0)), a.band_lo, a.band_inv);
      bool any = false;
#pragma unroll
      for (int j = 0; j < PULL_UN; j++) {
        const uint32_t e = e0 + 64 * j + lane;
        act[j] = e < tk.z;
        const uint64_t cw = __builtin_nontemporal_load((const uint64_t*)a.cw + (act[j] ? e : tk.y));
        u[j] = (uint32_t)cw;
        w[j] = (uint32_t)(cw >> 32);
        act[j] = act[j] && heavy_band(wfloat<FLT>(w[j]), a.band_lo, a.band_inv) <= lim;
        any |= act[j];
      }
      if (!ballot(any)) break;
#pragma unroll
      for (int j = 0; j < PULL_UN; j++) du[j] = act[j] ? a.dist[gb + u[j]] : INF;
#pragma unroll
      for (int j = 0; j < PULL_UN; j++)
        if (act[j]) {
          best = min(best, pull_cand<FLT>(du[j], w[j]));
          scanned++;
        }
    }
    best = wave_min_u32(best);
    scanned = (uint32_t)wave_sum64(scanned);
    bool imp = false;
    if (lane == 0 && best < INF) {
      const uint32_t old = afmin(a.dist + tk.x, best);
      imp = best < old;
      ws.v[ST_RELAX]++;
      if (imp) ws.v[ST_PULL_IMPROVED]++;
    }
    if (lane == 0) {
      ws.v[ST_EDGES] += scanned;
      ws.v[ST_PULL_EDGES] += scanned;
    }
    const uint32_t pp = REPL ? push_part(a, gb, part) : part;
    pb_add(a, pb, imp, ((uint64_t)best << 32) | tk.x, light_queue<FLT>(a, pidl, best, pp, pc, QM_OWN), ws);
  }
  pb_flush(flush_args(a, false), pb, ws);
  publish_stats(a.st, s_stat, ws, nullptr);
}

// Early pull: the light-queue items still queued at the request whose vertex lies at or below the split
// point, and still holds the item's distance, are queued again at the wide width (the pull saw their
// distance, so their heavy edges are covered; their light edges are relaxed when they are expanded).
// Items of vertices beyond the split point are not needed: k_pull_a queues every such vertex that has a
// distance.  A reference item (one chunk of a long row) is carried as it is.
template <bool FLT, bool REPL>
__global__ void __launch_bounds__(256) k_pull_carry(KArgs a, uint32_t slot) {
  __shared__ uint64_t s_pbe[WAVES_PER_WG][PB_CAP];
  __shared__ uint32_t s_pbq[WAVES_PER_WG][PB_CAP];
  __shared__ uint64_t s_tc[WAVES_PER_WG][TC_SIZE];
  __shared__ uint64_t s_stat[WAVES_PER_WG][NSTATS];
  if (!pull_requested(a, slot) || ald(a.st + PULL_OUTCOME) != 1u || ald(a.st + PULL_CARRY) == 0u ||
      ald(a.st + PULL_REBUCKET) == 0u || !a.carry)
    return;
  const uint32_t c = ald(a.st + PULL_BUCKET);
  const float idl = pull_req_idl(a);
  const float pidl = a.inv_delta_resume;
  const uint32_t pc = ald(a.st + PULL_NEWHEAD);
  const uint32_t lane = lane_id();
  const uint32_t wv = threadIdx.x >> 6;
  PushBuf pb{s_pbe[wv], s_pbq[wv], s_tc[wv], 0u, false};
  for (uint32_t i = lane; i < TC_SIZE; i += 64) s_tc[wv][i] = ~0ull;
  wave_sync();
  const uint32_t part = blockIdx.x % a.parts;
  WaveStats ws;
#pragma unroll
  for (int i = 0; i < NSTATS; i++) ws.v[i] = 0;
  const uint32_t nq = nq_of(a);
  const uint32_t nwaves = gridDim.x * WAVES_PER_WG, wave_gid = blockIdx.x * WAVES_PER_WG + wv;
  // 64-position windows of every light queue's range, dealt round-robin over the wavefronts.  The ranges
  // are read 64 queues per load and only the non-empty ones visited (a few dozen of the nb * parts).
  uint32_t widx = 0;
  for (uint32_t q0 = 0; q0 < nq; q0 += 64) {
    const uint2 htl = q0 + lane < nq ? a.carry[q0 + lane] : make_uint2(0u, 0u);
    uint64_t ne = ballot(htl.y > htl.x);
    for (; ne; ne &= ne - 1) {
      const uint32_t ql = (uint32_t)__builtin_ctzll(ne), q = q0 + ql;
      const uint2 ht = make_uint2(rdlane(htl.x, ql), rdlane(htl.y, ql));
      const uint32_t nwin = (ht.y - ht.x + 63) / 64;
      uint32_t first = (wave_gid + nwaves - widx % nwaves) % nwaves;  // this wavefront's first window here
      for (uint32_t w = first; w < nwin; w += nwaves) {
        const uint32_t pos = ht.x + w * 64 + lane;
        const bool in = pos < ht.y;
        uint64_t e = EMPTY64;
        if (in) {
          const uint32_t ph = a.tab[(size_t)q * a.maxv + (pos >> CH_SHIFT)];
          if (ph < TAB_BAD) e = a.pool[((size_t)ph << CH_SHIFT) + (pos & (CH - 1))];
        }
        bool keep = false;
        uint32_t qd = 0, gbv = 0;
        if (e != EMPTY64) {
          const uint32_t gid = (uint32_t)e, du = (uint32_t)(e >> 32);
          uint32_t vtx = gid;
          qd = du;
          if (du == REF_TAG) {
            const uint32_t* r = a.hrec + (size_t)(gid >> REF_CHUNK_BITS) * HREC_WORDS;
            vtx = ald(r + 0);
            qd = ald(r + 1);
          }
          keep = a.dist[vtx] >= qd && !unsettled<FLT>(qd, c, idl);
          gbv = REPL ? vtx - vtx % a.n : 0u;
        }
        const uint32_t pp = REPL ? push_part(a, gbv, part) : part;
        pb_add(a, pb, keep, e, light_queue<FLT>(a, pidl, qd, pp, pc, QM_OWN), ws);
      }
      widx += nwin;
    }
  }
  pb_flush(flush_args(a, false), pb, ws);
  publish_stats(a.st, s_stat, ws, nullptr);
}

// The pull kernels reserve and write their queue entries without publishing them (one same-address
// atomic per queue and flush less: on RMAT-24 the re-bucketing pull queues 7M vertices into a few
// dozen queues, whose counters then serialised the kernel).  No claim runs between the persistent
// launches, so afterwards every queue's unclaimed positions [head, tail) are written, and all of them are
// published at once.
__global__ void __launch_bounds__(1024) k_pull_publish(KArgs a, uint32_t slot) {
  if (!pull_requested(a, slot) || ald(a.st + PULL_OUTCOME) != 1u) return;
  const uint32_t nq = nq_of(a);
  for (uint32_t q = threadIdx.x; q < 2 * nq; q += blockDim.x) ast(qavail(a, q), ald(qtail(a, q)) - ald(qhead(a, q)));
}

// ------------------------------------------------------------------------------------------------
// init / cleanup / fills / light-heavy split
// ------------------------------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint64_t count, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t count, uint64_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Resets the state words and enqueues the sources into bucket 0, light sub-queue 0 (single block).
// One workgroup of 1024 threads.  The state words (a whole number of 128-byte lines, 16-byte aligned in
// every workspace) are cleared with 16-byte stores: about 470 KB on RMAT-24 at 64 buckets of 4 parts.
__global__ void __launch_bounds__(1024) k_init_state(KArgs a, const uint32_t* srcs, uint32_t k, int replicated) {
  const uint32_t nw = state_words(a.nb * a.parts, a.nb);
  static_assert(LINE % 4 == 0, "state lines are whole uint4s");
  uint4* st4 = reinterpret_cast<uint4*>(a.st);
  for (uint32_t i = threadIdx.x; i < nw / 4; i += blockDim.x) st4[i] = make_uint4(0u, 0u, 0u, 0u);
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t g = 0; g < VER_GROUPS; g++)
      *verp(a, g) = group_waves(a.total_waves / WAVES_PER_WG, g);  // everyone idle
    a.st[S_QBASE + LINE] = k;      // tail of queue 0
    a.st[S_QBASE + 2 * LINE] = k;  // all of it published
    a.st[S_POOL] = (k + CH - 1) >> CH_SHIFT;
  }
  for (uint32_t j = threadIdx.x; j < ((k + CH - 1) >> CH_SHIFT); j += blockDim.x) a.tab[j] = j;
  for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
    const uint32_t gid = replicated ? i * a.n + srcs[i] : srcs[i];
    a.dist[gid] = 0;
    a.pool[i] = (uint64_t)gid;  // distance 0 in the high word
  }
}

// Restores the pool and the chunk maps to their empty state, using the previous solve's counters
// (which the next k_init_state clears).
__global__ void k_cleanup(KArgs a) {
  const uint32_t nqt = 2 * a.nb * a.parts;
  const uint64_t used_chunks = min((uint64_t)a.st[S_POOL], (uint64_t)a.pool_chunks);
  const uint64_t pool_used = used_chunks << CH_SHIFT;
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nth = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = tid; i < pool_used; i += nth) a.pool[i] = EMPTY64;
  for (uint32_t q = blockIdx.x; q < nqt; q += gridDim.x) {
    const uint32_t t = a.st[S_QBASE + QLINES * q * LINE + LINE];
    const uint32_t used = min((t + CH - 1) >> CH_SHIFT, a.maxv);
    for (uint32_t j = threadIdx.x; j < used; j += blockDim.x) a.tab[(size_t)q * a.maxv + j] = TAB_UNSET;
  }
}

// Per-row partition of (col, w) into classes: the light edges (w <= delta) first, then the heavy edges
// band by band in increasing weight (heavy_band), so that a pull phase reading a row's heavy part can
// stop at the first band above its best candidate.  The order inside a class is not kept (no reader
// depends on it).  Rows of up to SPLIT_BIG edges: k_split, 64 rows per wavefront.  Longer rows are
// appended to a list (aux[0] = count, aux[SPLIT_AUX_ROWS..] = rows) and partitioned by k_split_big, one
// 1024-thread workgroup per row, so that the long rows of a power-law graph (on RMAT-24, rows above 1024
// edges hold over half the edges) do not serialise on one wavefront.

__device__ __forceinline__ bool is_light(uint32_t x, float delta, uint32_t idelta, int is_float) {
  return is_float ? __uint_as_float(x) <= delta : x <= idelta;
}
__device__ __forceinline__ uint32_t int_delta(float delta) {
  return delta >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)delta;  // int weights: w <= floor(delta)
}
constexpr uint32_t SPLIT_NCLS = 1 + SPLIT_BANDS;  // light, then the heavy bands
__device__ __forceinline__ uint32_t edge_class(uint32_t x, float delta, uint32_t idelta, int is_float, float blo,
                                               float binv) {
  if (is_light(x, delta, idelta, is_float)) return 0u;
  return 1u + heavy_band(is_float ? __uint_as_float(x) : (float)x, blo, binv);
}

// Rows of up to SPLIT_BIG edges, 64 consecutive rows per wavefront at a time: their edges are
// flattened (row k owns [off[k], off[k+1]) of the order) and read in 64-edge windows, four in flight.
// Pass 1 counts each row's edges per class in LDS; one lane per row turns its counts into the classes'
// start positions; pass 2 sends every edge to its class's next position (LDS atomics).  When the 64 rows
// hold more than SPLIT_CAP edges they are taken as consecutive sub-groups of rows that each fit (a row of
// at most SPLIT_BIG <= SPLIT_CAP edges always does).
constexpr uint32_t SPLIT_CAP = 16384;
static_assert(SPLIT_BIG <= SPLIT_CAP, "a row must fit one sub-group");
__global__ void __launch_bounds__(256) k_split(const uint32_t* row_ptr, const uint32_t* col, const uint32_t* w,
                                               uint32_t n, float delta, float blo, float binv, int is_float,
                                               uint2* cw2, uint32_t* lcnt, uint32_t* aux) {
  __shared__ uint32_t s_off[WAVES_PER_WG][65];
  __shared__ uint32_t s_rs[WAVES_PER_WG][64];
  __shared__ uint32_t s_cnt[WAVES_PER_WG][64 * SPLIT_NCLS];  // per row and class: count, then next position
  const uint32_t lane = lane_id();
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t idelta = int_delta(delta);
  uint32_t* off = s_off[wv];
  uint32_t* cnt = s_cnt[wv];
  auto locate = [&](uint32_t i) -> uint32_t {  // last row k with off[k] <= i
    uint32_t lo = 0;
#pragma unroll
    for (uint32_t step = 32; step >= 1; step >>= 1)
      if (off[lo + step] <= i) lo += step;
    return lo;
  };
  // groups of 64 rows are handed out by a counter (aux[1]), so that the dense groups of a power-law
  // graph's low ids do not pile up on the wavefronts that a static stride would give them to
  for (;;) {
    uint32_t grp = 0;
    if (lane == 0) grp = afadd(aux + 1, 1u);
    const uint64_t base = (uint64_t)rdlane(grp, 0) * 64;
    if (base >= n) break;
    const uint32_t v = (uint32_t)base + lane;
    const bool valid = v < n;
    const uint32_t rs = valid ? row_ptr[v] : 0u, re = valid ? row_ptr[v + 1] : 0u;
    const bool big = valid && re - rs > SPLIT_BIG;
    const uint64_t mb = ballot(big);
    if (mb) {
      uint32_t at = 0;
      if (lane == 0) at = afadd(aux, (uint32_t)__popcll(mb));
      at = rdlane(at, 0);
      if (big) aux[SPLIT_AUX_ROWS + at + mbcnt(mb)] = v;
    }
    const uint32_t d = (valid && !big) ? re - rs : 0u;
    const uint32_t incl = wave_incl_scan(d), excl = incl - d;
    s_rs[wv][lane] = rs;
    for (uint32_t first = 0; first < 64;) {
      const uint32_t e0 = rdlane(excl, first);
      const uint64_t fit = ballot(lane >= first && incl - e0 <= SPLIT_CAP);  // holds bit `first`
      const uint32_t last = 63u - (uint32_t)__clzll((long long)fit);
      const uint32_t total = rdlane(incl, last) - e0;
      const bool mine = lane >= first && lane <= last;
      off[lane] = lane < first ? 0u : (mine ? excl - e0 : total);
      if (lane == 0) off[64] = total;
      for (uint32_t i = lane; i < 64 * SPLIT_NCLS; i += 64) cnt[i] = 0u;
      wave_sync();
      // pass 1: class counts
      for (uint32_t w0 = 0; w0 < total; w0 += 256) {
        uint32_t x[4], k[4];
        bool in[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const uint32_t i = w0 + q * 64 + lane;
          in[q] = i < total;
          k[q] = locate(in[q] ? i : 0u);
          x[q] = in[q] ? w[s_rs[wv][k[q]] + (i - off[k[q]])] : 0u;
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (in[q]) atomicAdd(&cnt[k[q] * SPLIT_NCLS + edge_class(x[q], delta, idelta, is_float, blo, binv)], 1u);
      }
      wave_sync();
      // each class's first position in its row: the light edges at the row start, then band after band
      if (mine) {
        uint32_t pos = rs, lc = 0;
        for (uint32_t c = 0; c < SPLIT_NCLS; c++) {
          const uint32_t t = cnt[lane * SPLIT_NCLS + c];
          cnt[lane * SPLIT_NCLS + c] = pos;
          pos += t;
          if (c == 0) lc = t;
        }
        if (valid && !big) lcnt[v] = lc;
      }
      wave_sync();
      // pass 2: every edge to its place
      for (uint32_t w0 = 0; w0 < total; w0 += 256) {
        uint32_t c[4], x[4], k[4];
        bool in[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const uint32_t i = w0 + q * 64 + lane;
          in[q] = i < total;
          k[q] = locate(in[q] ? i : 0u);
          const uint32_t e = s_rs[wv][k[q]] + ((in[q] ? i : 0u) - off[k[q]]);
          c[q] = in[q] ? col[e] : 0u;
          x[q] = in[q] ? w[e] : 0u;
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (in[q]) {
            const uint32_t dst =
                atomicAdd(&cnt[k[q] * SPLIT_NCLS + edge_class(x[q], delta, idelta, is_float, blo, binv)], 1u);
            cw2[dst] = make_uint2(c[q], x[q]);
          }
      }
      wave_sync();
      first = last + 1;
    }
  }
}

// One workgroup of 1024 threads per long row: class counts (LDS atomics), their start positions, then
// every edge to its class's next position.
__global__ void __launch_bounds__(1024) k_split_big(const uint32_t* row_ptr, const uint32_t* col, const uint32_t* w,
                                                    float delta, float blo, float binv, int is_float, uint2* cw2,
                                                    uint32_t* lcnt, uint32_t* aux) {
  __shared__ uint32_t s_cnt[SPLIT_NCLS];
  __shared__ uint32_t s_row;
  const uint32_t t = threadIdx.x;
  const uint32_t idelta = int_delta(delta);
  const uint32_t nbig = aux[0];
  for (;;) {
    // rows are handed out by a counter (aux[2]): one very long row does not hold up the rows behind it
    if (t == 0) s_row = afadd(aux + 2, 1u);
    if (t < SPLIT_NCLS) s_cnt[t] = 0u;
    __syncthreads();
    const uint32_t i = s_row;
    __syncthreads();
    if (i >= nbig) break;
    const uint32_t v = aux[SPLIT_AUX_ROWS + i];
    const uint32_t rs = row_ptr[v], re = row_ptr[v + 1];
    for (uint32_t e = rs + t; e < re; e += 4096) {
      uint32_t x[4];
#pragma unroll
      for (int q = 0; q < 4; q++) x[q] = e + q * 1024 < re ? w[e + q * 1024] : 0u;
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (e + q * 1024 < re) atomicAdd(&s_cnt[edge_class(x[q], delta, idelta, is_float, blo, binv)], 1u);
    }
    __syncthreads();
    if (t == 0) {
      uint32_t pos = rs;
      for (uint32_t c = 0; c < SPLIT_NCLS; c++) {
        const uint32_t k = s_cnt[c];
        s_cnt[c] = pos;
        pos += k;
        if (c == 0) lcnt[v] = k;
      }
    }
    __syncthreads();
    for (uint32_t e = rs + t; e < re; e += 4096) {
      uint32_t x[4], c[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const bool in = e + q * 1024 < re;
        x[q] = in ? w[e + q * 1024] : 0u;
        c[q] = in ? col[e + q * 1024] : 0u;
      }
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (e + q * 1024 < re) {
          const uint32_t dst = atomicAdd(&s_cnt[edge_class(x[q], delta, idelta, is_float, blo, binv)], 1u);
          cw2[dst] = make_uint2(c[q], x[q]);
        }
    }
    __syncthreads();  // the positions are reset for the next row
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
  hipLaunchKernelGGL(k_init_state, dim3(1), dim3(1024), 0, s, a, srcs, k, replicated);
  return hipGetLastError();
}
template <int UN>
static void launch_solve_un(const KArgs& a, bool is_float, bool replicated, uint32_t grid, hipStream_t s) {
  if (is_float) {
    if (replicated) hipLaunchKernelGGL((k_sssp<true, true, UN>), dim3(grid), dim3(WG_THREADS), 0, s, a);
    else hipLaunchKernelGGL((k_sssp<true, false, UN>), dim3(grid), dim3(WG_THREADS), 0, s, a);
  } else {
    if (replicated) hipLaunchKernelGGL((k_sssp<false, true, UN>), dim3(grid), dim3(WG_THREADS), 0, s, a);
    else hipLaunchKernelGGL((k_sssp<false, false, UN>), dim3(grid), dim3(WG_THREADS), 0, s, a);
  }
}
hipError_t launch_solve(const KArgs& a, bool is_float, bool replicated, uint32_t grid, hipStream_t s) {
  if (a.windows == 8) launch_solve_un<8>(a, is_float, replicated, grid, s);
  else launch_solve_un<4>(a, is_float, replicated, grid, s);
  return hipGetLastError();
}
template <bool FLT, bool REPL>
static void launch_pull_t(const KArgs& a, uint32_t slot, uint32_t grid, hipStream_t s) {
  const uint32_t cblocks = std::max<uint32_t>(1, std::min<uint32_t>((a.total_n + PULL_COUNT_THREADS - 1) /
                                                                   PULL_COUNT_THREADS, 256));
  hipLaunchKernelGGL((k_pull_count<FLT, REPL>), dim3(cblocks), dim3(PULL_COUNT_THREADS), 0, s, a, slot);
  hipLaunchKernelGGL(k_pull_decide, dim3(1), dim3(1024), 0, s, a, slot);  // (single workgroups: one pass over the queues)
  hipLaunchKernelGGL((k_pull_carry<FLT, REPL>), dim3(grid), dim3(256), 0, s, a, slot);
  hipLaunchKernelGGL((k_pull_a<FLT, REPL>), dim3(grid), dim3(256), 0, s, a, slot);
  hipLaunchKernelGGL((k_pull_b<FLT, REPL>), dim3(grid), dim3(256), 0, s, a, slot);
  hipLaunchKernelGGL(k_pull_publish, dim3(1), dim3(1024), 0, s, a, slot);
}
hipError_t launch_pull(const KArgs& a, bool is_float, bool replicated, uint32_t slot, uint32_t grid, hipStream_t s) {
  if (is_float) {
    if (replicated) launch_pull_t<true, true>(a, slot, grid, s);
    else launch_pull_t<true, false>(a, slot, grid, s);
  } else {
    if (replicated) launch_pull_t<false, true>(a, slot, grid, s);
    else launch_pull_t<false, false>(a, slot, grid, s);
  }
  return hipGetLastError();
}
hipError_t launch_cleanup(const KArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_cleanup, dim3(1024), dim3(256), 0, s, a);
  return hipGetLastError();
}
// Per-vertex row record (row start, row end, light edges, 0) for the solver's item loads, and the first
// two heavy edges of each split row (read by the pull phases with the distances, so that most vertices
// need no random row access).
__global__ void k_rinfo(const uint32_t* row_ptr, const uint32_t* lcnt, uint32_t n, const uint2* cw2, uint4* rinfo,
                        uint4* phead) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r0 = row_ptr[v], r1 = row_ptr[v + 1], lc = lcnt[v];
    rinfo[v] = make_uint4(r0, r1, lc, 0u);
    if (phead) {
      const uint32_t h = r0 + lc;
      const uint2 e1 = h < r1 ? cw2[h] : make_uint2(PHEAD_NONE, 0u);
      const uint2 e2 = h + 1 < r1 ? cw2[h + 1] : make_uint2(PHEAD_NONE, 0u);
      phead[v] = make_uint4(e1.x, e1.y, e2.x, e2.y);
    }
  }
}

hipError_t launch_split(const uint32_t* row_ptr, const uint32_t* col, const uint32_t* w, uint32_t n, float delta,
                        float band_lo, float band_inv, bool is_float, uint2* cw2, uint32_t* lcnt, uint32_t* aux,
                        uint4* rinfo, uint4* phead, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipError_t e = launch_fill_u32(aux, SPLIT_AUX_ROWS, 0, s);  // long-row count and the two work counters
  if (e != hipSuccess) return e;
  const uint32_t blocks = std::min<uint32_t>((n + 255) / 256, 4096);  // groups of 64 rows, taken dynamically
  hipLaunchKernelGGL(k_split, dim3(blocks), dim3(256), 0, s, row_ptr, col, w, n, delta, band_lo, band_inv,
                     is_float ? 1 : 0, cw2, lcnt, aux);
  hipLaunchKernelGGL(k_split_big, dim3(1024), dim3(1024), 0, s, row_ptr, col, w, delta, band_lo, band_inv,
                     is_float ? 1 : 0, cw2, lcnt, aux);
  if (rinfo)
    hipLaunchKernelGGL(k_rinfo, dim3(std::min<uint32_t>((n + 255) / 256, 8192)), dim3(256), 0, s, row_ptr, lcnt, n,
                       (const uint2*)cw2, rinfo, phead);
  return hipGetLastError();
}

}  // namespace adds
