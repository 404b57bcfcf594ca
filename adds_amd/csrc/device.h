// Shared declarations between the solver kernels (sssp_kernels.hip) and the C ABI (api.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace adds {

// ---- worklist geometry ----
constexpr uint32_t CH_SHIFT = 11;  // pool chunk = 2048 entries (16 KiB)
constexpr uint32_t CH = 1u << CH_SHIFT;
constexpr uint64_t EMPTY64 = ~0ull;      // unwritten pool entry
constexpr uint32_t TAB_UNSET = 0xFFFFFFFFu;
constexpr uint32_t TAB_BAD = 0xFFFFFFFEu;
constexpr uint32_t WG_THREADS = 256;  // 4 wavefronts of 64
constexpr uint32_t WAVES_PER_WG = WG_THREADS / 64;
constexpr uint32_t MAX_PROBE_QUEUES = 64;  // (concurrency + 3) * parts queues are probed, one per lane
// Edge-range references: a vertex whose edge range (light or heavy part) is longer than hub_degree is
// expanded as range records + queue entries that each name one chunk of the range.  A reference entry
// has REF_TAG in its distance word (no pushed distance is ever 0xFFFFFFFF) and (record << REF_CHUNK_BITS
// | chunk) in its vertex word.  Record r: hrec[8r .. 8r+4] = vertex-instance, distance, first edge, edges
// per chunk, end edge.
constexpr uint32_t REF_TAG = 0xFFFFFFFFu;
constexpr uint32_t REF_CHUNK_BITS = 9;
constexpr uint32_t REF_MAX_CHUNKS = 1u << REF_CHUNK_BITS;
constexpr uint32_t REF_MAX_RECORDS = (1u << (32 - REF_CHUNK_BITS)) - 1;  // keeps every entry != EMPTY64
constexpr uint32_t HREC_WORDS = 8;

// ---- state words (uint32 indices); every hot counter owns a 128-byte line (32 words) ----
constexpr uint32_t LINE = 32;
constexpr uint32_t S_CUR = 0 * LINE;        // absolute index of the head bucket
constexpr uint32_t S_TAIL = 1 * LINE;       // [0] 1 once the solve is in its tail (see tail_items),
                                            // [1] most items pending seen at a head move,
                                            // [2] pull requests made (PULL_GEN), [3] bucket of the last
                                            // request, [4..5] u64 heavy edges queued in it, [6] pulls
                                            // performed, [7] the last request's outcome (1 pull, 2 push),
                                            // [8] the requester's 1/delta (float bits), [9] 1 once a pull
                                            // has re-bucketed the solve (pull_delta), [10] its new head,
                                            // [11] 1 when the request left queued items to carry over
constexpr uint32_t S_DONE = 2 * LINE;       // [0] done flag, [1] error code
constexpr uint32_t S_POOL = 3 * LINE;       // next free pool chunk
// (lines 4 and 5 are spare)
constexpr uint32_t S_ADV = 6 * LINE;        // try-lock serialising head-advance scans
constexpr uint32_t S_HREC = 7 * LINE;       // next free edge-range record
// Termination words, one line each for VER_GROUPS groups of workgroups (workgroup b is in group
// b % VER_GROUPS, which on MI355X is its XCD): u64 (busy transitions << 32) | idle wavefronts.
constexpr uint32_t VER_GROUPS = 8;
constexpr uint32_t S_VER = 8 * LINE;
// Pull-phase work counters (k_pull_*), on S_PULL2 (S_PULL is a spare line): the long-row task count
// [0], [1] spare, u64 heavy in-edges of unsettled vertices [2..3] and u64 unsettled vertices [4..5].
constexpr uint32_t S_PULL = (8 + VER_GROUPS) * LINE;
constexpr uint32_t S_PULL2 = S_PULL + LINE;
constexpr uint32_t PULL_GEN = S_TAIL + 2, PULL_BUCKET = S_TAIL + 3, PULL_H = S_TAIL + 4, PULL_DONE = S_TAIL + 6,
                   PULL_OUTCOME = S_TAIL + 7, PULL_IDL = S_TAIL + 8, PULL_REBUCKET = S_TAIL + 9,
                   PULL_NEWHEAD = S_TAIL + 10, PULL_CARRY = S_TAIL + 11;
// Statistics: STAT_COPIES copies of u64 stats[NSTATS] (two lines each).  Workgroup b adds its counters,
// summed over its wavefronts in LDS, to copy b % STAT_COPIES, and the host adds the copies up: atomics on
// one address serialise (about 12 ns each), and one per wavefront and counter on a single copy cost a
// persistent launch hundreds of microseconds at its end.
constexpr uint32_t STAT_COPIES = 8;
constexpr uint32_t S_STATC = (10 + VER_GROUPS) * LINE;
__host__ __device__ constexpr uint32_t stat_copy(uint32_t c) { return S_STATC + c * 2 * LINE; }
// Queues: nq = nb * parts light queues (index slot * parts + part) followed by nq heavy queues (index
// nq + slot * parts + part).  Queue q: head (claimed) at S_QBASE + 3q*LINE, tail (reserved) at +LINE,
// published-unclaimed count at +2*LINE.  Then the bucket slots (SLOT_WORDS below).
constexpr uint32_t S_QBASE = S_STATC + STAT_COPIES * 2 * LINE;
constexpr uint32_t HOST_STATE_WORDS = S_QBASE;  // what the host copies back after a solve: every word before the queues
constexpr uint32_t QLINES = 3;
constexpr uint32_t s_slotlines(uint32_t nq) { return S_QBASE + QLINES * 2 * nq * LINE; }
// Per bucket slot: the in-flight claim counters (light claims of part p at index p, heavy at 16 + p),
// each on a cache line of its own -- every claim updates one twice, and packed into one line they were
// a point of contention for the whole GPU -- then the phase word (1 = the heavy queues of the head
// bucket are open) on a line of its own.  Words 2..3 of light line p: u64 heavy edges of the vertices
// that claims of part p filed in the bucket's heavy queues (the cost of its push-direction heavy phase).
constexpr uint32_t INFL_HEAVY = 2;
constexpr uint32_t INFL_STRIDE = LINE;
constexpr uint32_t SLOT_WORDS = (2 * 16 + 1) * LINE;
constexpr uint32_t state_words(uint32_t nq, uint32_t nb) { return s_slotlines(nq) + nb * SLOT_WORDS; }
constexpr uint32_t MAX_PARTS = 16;
constexpr uint32_t TRACE_HBINS = 2048;
constexpr uint32_t TRACE_HCOPIES = 16;  // histogram copies (wavefront id modulo 16), summed by the host

enum StatIdx {
  ST_EXPANDED = 0, ST_EDGES, ST_RELAX, ST_PUSH, ST_STALE, ST_TASKS, ST_REBUCKET,
  ST_PROBE, ST_CASFAIL, ST_IDLE, ST_LATE, ST_LOCAL, ST_HEDGES,
  ST_T_FLUSH,  // ticks in push-buffer flushes (part of the item and edge time below)
  // wavefront time by activity (wall-clock ticks): probing/claiming, item handling, edge windows,
  // idle (termination test, head advance, back-off)
  ST_T_PROBE, ST_T_ITEMS, ST_T_EDGES, ST_T_IDLE,
  ST_PULL_EDGES,     // in-edges scanned by pull phases (also counted in ST_EDGES)
  ST_PULL_IMPROVED,  // vertices whose distance a pull phase lowered
  NSTATS
};

static_assert(NSTATS <= 32, "the stats block is two state lines of u64");

enum DevErr : uint32_t { DERR_NONE = 0, DERR_OVERFLOW = 1, DERR_TIMEOUT = 2, DERR_ABORT = 3 };

struct KArgs {
  const uint2* cw;       // per edge (destination, weight bits), each row light-first (w <= delta), one pad
  const uint4* rinfo;    // per vertex (row start, row end, light edges at the front, 0): one load per item
  uint32_t* dist;     // distance words, total_n of them
  uint32_t* st;       // state words
  uint64_t* pool;     // pool_chunks * CH entries: (dist << 32) | vertex-instance id
  uint32_t* tab;      // per queue: maxv virtual-chunk -> physical-chunk map
  uint32_t* hrec;     // edge-range records, hrec_cap of them
  const uint32_t* host_abort;  // host-visible abort word (may be null)
  uint32_t n;         // vertices of the graph (per replica)
  uint32_t total_n;   // replicas * n
  // Replicated solves with ADDS_F_REPLICA_PARTS: the items of replica r go to sub-queue part r % parts
  // (inst_part != 0), so that with parts = 8 each replica's items are mostly taken by one XCD;
  // r = mulhi(base, n_magic) for a replica's base offset r * n (n_magic = ceil(2^32 / n), exact for
  // multiples of n below 2^32).
  uint32_t inst_part;
  uint32_t n_magic;
  uint32_t pool_chunks;
  uint32_t maxv;      // virtual chunks per queue
  uint32_t nb;        // bucket ring length (power of two)
  uint32_t parts;     // sub-queues per bucket
  float inv_delta;    // 1 / bucket width
  uint32_t chunk;     // most items claimed at once by a wavefront (<= 256)
  uint32_t chunk_heavy;  // the same for heavy queues
  uint32_t claim_min; // a claim takes max(claim_min, published / claim_div) items, at most chunk
  uint32_t claim_div;
  uint32_t hrec_cap;
  uint32_t hub_degree;  // ranges longer than this are expanded through range references
  uint32_t task_edges;  // edges per reference chunk
  uint32_t windows;     // 64-edge windows per group of the edge pipeline (4 or 8)
  uint32_t tail_items;  // when the head moves and at most this many items are left in the window (and at
                        // most a quarter of the peak), the rest of the solve runs as one bucket (0: never)
  uint32_t concur;
  uint32_t total_waves;
  uint64_t timeout_ticks;  // wall_clock64 ticks
  uint32_t cached_probe;   // probe queue counters with L2-cached loads (claims stay atomic)
  uint32_t sticky;         // after a full claim from the behind/head bucket, claim the same queue again
  uint32_t full_first;     // claims try queues that hold a full chunk first (ADDS_F_FULL_FIRST)
  uint32_t local_cont;     // a light claim keeps up to 64 same-bucket improvements for itself (2: only a
                           // claim of at most chunk / 4 items)
  uint32_t nowait;         // flushes do not wait for their atomicMins (every improvement is queued)
  // Optional bucket timeline (null = off): trace[0] = record count, then records of 4 words from
  // trace + 4: wall clock, (kind << 56 | bucket), previous head, pool chunks used.  Kinds: 1 = the
  // head bucket's heavy phase opened, 2 = the head moved to `bucket`, 0 = start, 3 = done, 5 = pull phase
  // requested (previous head: heavy edges queued), 6 = pull phase decided (previous head: 1 pulled,
  // 2 pushed, 3 pulled and re-bucketed).
  // After the records: an activity histogram of TRACE_HBINS bins of trace_slice ticks, 8 words per bin
  // (wavefront ticks probing, on items, on edges, idle; edges relaxed; items expanded; ticks in
  // push-buffer flushes; 1 spare), in
  // TRACE_HCOPIES copies.  Kind 4 records mark claims with more than 16384 edges (bucket field = edges,
  // previous-head field = ticks the claim took).
  uint64_t* trace;
  uint32_t trace_cap;    // records
  uint32_t trace_slice;  // wall-clock ticks per histogram bin
  // Pull phases (symmetric graphs): launch l of the persistent kernel (l = 0 .. pull_slots) may request
  // pull l + 1 when the head bucket's light phase is done and its heavy queues hold at least pull_min
  // heavy edges; its wavefronts then leave, k_pull_* relax the bucket's heavy edges from the receiving
  // side (or open the push-direction phase when that is cheaper), and launch l + 1 continues the solve.
  uint32_t launch_idx;
  uint32_t pull_slots;
  uint64_t pull_min;
  uint32_t pull_eager;   // request at every opportunity and always pull (testing and tuning)
  const uint32_t* pull_ok;  // the graph's device word: nonzero while it is declared symmetric
  float inv_delta_resume;   // 1 / pull_delta: the bucket width after the first pull (0: unchanged)
  uint32_t resume_all_light;  // after a re-bucketing pull every edge is light (ADDS_F_PULL_ALL_LIGHT)
  // ADDS_F_PULL_SPLIT: launch 0 runs buckets 0 .. pull_at (narrower than the light threshold) through
  // their light phases only and then asks for one pull phase at bucket pull_at, which relaxes the heavy
  // edges of all of them (PULL_AT_OFF: off)
  uint32_t pull_at;
  // ... and below pull_at only the vertices of at least order_degree light edges wait for their bucket:
  // the others are queued in the bucket of the claim that finds them (0: every vertex waits)
  uint32_t order_degree;
  // the heavy part of every row is in band order (heavy_band below, with these two values); phead[v]:
  // its first two heavy edges (col, w, col, w), PHEAD_NONE as the col of an absent one
  float band_lo, band_inv;
  const uint4* phead;
  // early pull (split first bucket, every edge light after a re-bucketing pull): launch 0 asks for the
  // pull once at most pull_early items are queued or in flight, after four times as many were seen; the
  // items still queued then are carried over into the wide buckets.  carry: per queue (head, tail) at the
  // request, 2 * nb * parts entries (after the pull tasks).
  uint32_t pull_early;
  uint2* carry;
  // re-bucketing pull: k_pull_a stages the vertices it queues at the new head per 256-vertex chunk
  // (pstage[chunk * 256 ..], pcnt[chunk]) and each wavefront then reserves and writes all of its own at
  // once (null: queued through the push buffer)
  uint64_t* pstage;
  uint32_t* pcnt;
  uint32_t ptask_cap;
  uint4* ptask;          // long-row tasks of a pull phase: (vertex-instance, first edge, end edge, 0)
};

// Pull phases: rows with more heavy edges than PULL_LONG are scanned by k_pull_b in tasks of at most
// PULL_TASK edges (one wavefront each), the rest by k_pull_a, 256 vertex-instances per wavefront.
constexpr uint32_t PULL_LONG = 128;
constexpr uint32_t PULL_AT_OFF = 0xFFFFFFFFu;
constexpr uint32_t PHEAD_NONE = 0xFFFFFFFFu;  // no vertex id (vertex-instances stay below 2^32 - 1)
constexpr uint32_t PULL_TASK = 1024;
constexpr uint32_t CARRY_UINT4 = MAX_PROBE_QUEUES * 32;  // room for 2 * nb * parts (head, tail) pairs: nb <= 256, parts <= 8

// Inclusive prefix sum over the 64 lanes of a wavefront (every lane active), with DPP: row shifts by
// 1, 2, 4 and 8 scan each row of 16 lanes (lanes shifted in from outside the row add 0), then
// row_bcast:15 adds the last lane of rows 0 and 2 to rows 1 and 3, and row_bcast:31 adds lane 31 to
// rows 2 and 3.  Each step compiles to one v_add_u32_dpp.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_or_zero(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWS, 0xF, true);
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  x += dpp_or_zero<0x111, 0xF>(x);  // row_shr:1
  x += dpp_or_zero<0x112, 0xF>(x);  // row_shr:2
  x += dpp_or_zero<0x114, 0xF>(x);  // row_shr:4
  x += dpp_or_zero<0x118, 0xF>(x);  // row_shr:8
  x += dpp_or_zero<0x142, 0xA>(x);  // row_bcast:15 into rows 1 and 3
  x += dpp_or_zero<0x143, 0xC>(x);  // row_bcast:31 into rows 2 and 3
  return x;
}

// Mean sampled weight of a graph (the statistic the reference's profiler_kernel computes), api.hip.
}  // namespace adds
struct adds_graph;
namespace adds {
float graph_mean_weight(const adds_graph* g);

// Host launchers (sssp_kernels.hip)
hipError_t launch_fill_u32(uint32_t* p, uint64_t count, uint32_t v, hipStream_t s);
hipError_t launch_fill_u64(uint64_t* p, uint64_t count, uint64_t v, hipStream_t s);
hipError_t launch_init(const KArgs& a, const uint32_t* srcs, uint32_t k, int replicated, uint32_t inf_bits,
                       hipStream_t s);
hipError_t launch_solve(const KArgs& a, bool is_float, bool replicated, uint32_t grid, hipStream_t s);
hipError_t launch_cleanup(const KArgs& a, hipStream_t s);
// The kernels of pull phase `slot` (1 .. pull_slots): count, short rows, long-row tasks.  Each returns at
// once unless launch slot - 1 of the persistent kernel requested this pull.
hipError_t launch_pull(const KArgs& a, bool is_float, bool replicated, uint32_t slot, uint32_t grid, hipStream_t s);
// Per-row partition of (col, w) into light (w <= delta) then heavy edges, the heavy ones in increasing
// weight band (heavy_band with band_lo, band_inv), written interleaved as (col, w) pairs, with per-row
// light counts.  aux: split_aux_words(m) words of scratch (long rows).
hipError_t launch_split(const uint32_t* row_ptr, const uint32_t* col, const uint32_t* w, uint32_t n, float delta,
                        float band_lo, float band_inv, bool is_float, uint2* cw2, uint32_t* lcnt, uint32_t* aux,
                        uint4* rinfo, uint4* phead, hipStream_t s);
// Rows longer than SPLIT_BIG edges are split by one 1024-thread workgroup each (k_split_big); the list
// of such rows needs at most m / SPLIT_BIG entries, after SPLIT_AUX_ROWS words: the list's length and
// the work counters of the two split kernels.
constexpr uint32_t SPLIT_BIG = 1024;
// Weight bands of a row's heavy part: band(x) = floor((x - band_lo) * band_inv) clamped to 0 .. SPLIT_BANDS-1,
// x the weight as a float (int weights converted).  Monotone in x, so a heavy edge of a higher band than
// a value's band weighs more than that value: a scan in band order can stop there (pull phases).
constexpr uint32_t SPLIT_BANDS = 32;
__host__ __device__ inline uint32_t heavy_band(float x, float band_lo, float band_inv) {
  return (uint32_t)fminf(fmaxf((x - band_lo) * band_inv, 0.0f), (float)(SPLIT_BANDS - 1));
}
constexpr uint32_t SPLIT_AUX_ROWS = 3;
constexpr uint64_t split_aux_words(uint64_t m) { return m / SPLIT_BIG + SPLIT_AUX_ROWS + 1; }

// Litmus (litmus.hip)
hipError_t launch_litmus(uint32_t* mem, uint32_t pairs, uint32_t rounds, uint32_t variant, unsigned long long* out,
                         uint32_t* xcc, hipStream_t s);
hipError_t launch_xcd_map(uint32_t* xcc, uint32_t workgroups, uint32_t threads, hipStream_t s);
hipError_t launch_gather_probe_init(uint2* cw, uint64_t m, uint32_t n, uint32_t* d, hipStream_t s);
hipError_t launch_gather_probe(const uint2* cw, const uint32_t* d, uint64_t m, uint32_t* out, uint32_t workgroups,
                               hipStream_t s);

// Backward (backward_kernels.hip)
struct BwdArgs {
  const uint32_t* rrow;  // transpose: in-edges of v at rrow[v]..rrow[v+1]
  const uint32_t* reid;  // transpose: forward edge id of each in-edge (ascending per row)
  const uint32_t* rsrc;  // transpose: tail vertex of each in-edge
  const float* rw;       // transpose: weight of each in-edge
  const float* dist;     // [k][n]
  const float* gdist;    // [k][n]
  const uint32_t* srcs;  // [k] (device)
  uint2* P;              // [k][n] (parent edge, parent vertex); edge 0xFFFFFFFF = unreached/unresolved,
                         // 0xFFFFFFFE = root
  uint2* C;              // [k][n] choices of an equal-distance round
  long long* Acc;        // [k][n] children's sums (dense form) or sums attached at a vertex (sparse
                         // form); shares C's memory
  uint32_t* cnt;         // [k][n] pending children (dense form)
  uint32_t* list;        // nonzero g entries (sparse form), list_cap of them
  uint32_t list_cap;
  uint32_t* owner;       // [k][n] sparse form: the entry whose segment holds the vertex (shares cnt's memory)
  uint32_t* seg_up;      // [list_cap] sparse form: the segment an entry's segment attaches to (or NONE)
  uint32_t* attach;      // [list_cap] sparse form: the vertex where it attaches
  unsigned long long* part;  // [4 * parent blocks] per-block counts of k_bwd_parent
  unsigned long long* GW;   // [m] exact fixed-point gradient
  float* gw;                // [m] output
  unsigned long long* ctr;  // [0] unresolved, [1] nonzero g, [2] largest |g| bits, [3] non-finite g,
                            // [4] listed entries, [5] resolved by a round
  uint32_t n;
  uint32_t k;
  uint64_t m;
  int S;  // fixed-point scale: G = rint(g * 2^S)
};
hipError_t launch_bwd_tweights(const uint32_t* reid, const float* w, float* rw, uint64_t m, hipStream_t s);
hipError_t launch_bwd_parent(const BwdArgs& b, hipStream_t s);
hipError_t launch_bwd_round(const BwdArgs& b, hipStream_t s);
hipError_t launch_bwd_sparse(const BwdArgs& b, uint32_t count, hipStream_t s);
constexpr uint32_t BWD_MAX_PARENT_BLOCKS = 16384;
uint32_t bwd_parent_blocks(uint32_t n);
hipError_t launch_bwd_climb(const BwdArgs& b, hipStream_t s);
hipError_t launch_bwd_convert(const BwdArgs& b, hipStream_t s);

}  // namespace adds
