/* adds.h -- C ABI of the MI355X-native ADDS single-source shortest-path solver.
 *
 * The reference exposes this path as a process (`sssp [-g gpu] [-o out] [-s src] graph.gr`) around one
 * host function, gg_main(CSRGraph& hg, CSRGraph& gg) (ads_int/kernel.cu:185, ads_float/kernel.cu:196),
 * whose CSRGraph is a struct of raw pointers (ads_int/csr_graph.h:90-97).  This header is that seam as a
 * C ABI: plain pointers and sizes, status codes instead of exit()/assert, stream-ordered calls.
 *
 * Threading: a graph handle is not internally synchronised; use one handle per host thread.  Every call
 * is ordered on the HIP stream it is given (NULL = the default stream).  The caller owns output buffers;
 * the handle owns the device-resident CSR and the solver workspace.
 *
 * Distances are 32-bit words: uint32 for ADDS_W_U32 graphs (unreached = ADDS_INF_U32) and IEEE binary32
 * for ADDS_W_F32 graphs (unreached = +infinity, bits ADDS_INF_F32_BITS).  Int distances follow the
 * reference's int32 arithmetic for every finite distance below 2^31 (its INF is INT32_MAX,
 * ads_int/csr_graph.h:20-23, so larger sums are undefined there).  Here every integer path sum is
 * computed exactly up to 2^32-3 and saturates at ADDS_SAT_U32 = 2^32-2 beyond: a sum that does not fit
 * never wraps to a small distance and never reads as unreached.  Every solver (adds_sssp*, near-far)
 * and the CPU oracle saturate the same way, so their results agree bit for bit on such inputs too.
 */
#ifndef ADDS_H
#define ADDS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ADDS_INF_U32 0xFFFFFFFFu
#define ADDS_SAT_U32 0xFFFFFFFEu /* saturated integer path sum (see above) */
#define ADDS_INF_F32_BITS 0x7F800000u

typedef enum {
  ADDS_OK = 0,
  ADDS_EINVAL = 1,    /* bad argument, malformed graph, source out of range */
  ADDS_ENOMEM = 2,    /* host or device allocation failed */
  ADDS_EOVERFLOW = 3, /* worklist pool or range-record store exhausted during a solve */
  ADDS_ETIMEOUT = 4,  /* device watchdog or host abort ended a solve */
  ADDS_EHIP = 5,      /* a HIP runtime call failed */
  ADDS_EIO = 6,       /* file could not be opened, read or written */
  ADDS_ENODEV = 7     /* no usable GPU */
} adds_status;

typedef enum { ADDS_W_U32 = 0, ADDS_W_F32 = 1 } adds_wtype;

typedef struct adds_graph adds_graph; /* opaque; owns the device-resident CSR */

/* Solver tuning.  Zero in any field selects the default (the reference's weight-profiler heuristic for
 * delta, ads_int/kernel.cu:238-257 + wl.h:395-449).  Results never depend on these values. */
typedef struct {
  float delta;           /* bucket width in distance units */
  uint32_t wg_per_cu;    /* persistent workgroups per compute unit */
  uint32_t chunk;        /* worklist items a wavefront claims at once (1..256; default 64, 128 or 256 by size) */
  uint32_t flags;        /* ADDS_F_* */
  uint32_t nbuckets;     /* bucket ring length (power of two, 4..256) */
  uint32_t parts;        /* sub-queues per bucket (1, 2, 4 or 8; 8 = one per XCD); (concurrency + 3) * parts <= 64 */
  uint32_t hub_degree;   /* edge ranges longer than this are split into chunks any wavefront can take */
  uint32_t task_edges;   /* edges per chunk of a split range (1..65536, default 64) */
  uint32_t concurrency;  /* buckets open for claiming at once (1..16, at most nbuckets - 2, and
                            (concurrency + 3) * parts <= 64) */
  uint32_t timeout_ms;   /* device watchdog (default 120000) */
  uint32_t tail_items;   /* when the head bucket moves on and at most this many items are left in the bucket
                            window, no more than a quarter of the most seen at a head move, the rest of the
                            solve runs as one bucket with its heavy edges relaxed at once (default 0: never);
                            saves the phase changes of a long sparse tail */
  uint32_t edge_windows; /* edges per wavefront per step of the edge pipeline, in 64-edge windows: 4
                            (default) or 8 (twice the edges in flight, loaded as 16-byte pairs; pays on
                            throughput-bound solves) */
  float light_delta;     /* edges with w <= light_delta are light: relaxed whenever their source is expanded;
                            the others are heavy: relaxed once, after the source's bucket has settled.  At
                            least delta (default delta); a value above every weight makes all edges light
                            (no heavy phases: pays where vertices are rarely expanded twice) */
  uint64_t pool_entries; /* worklist pool capacity in entries (default 4 per vertex-instance + 1 per 16
                            edge-instances + 8M, doubled
                            and remembered per graph after an overflow; a solve that overflows is rerun) */
  float pull_delta;      /* bucket width after the solve's first pull phase (0, the default: delta).  What
                            is left after such a pull is the sparse rest of the graph; a wider bucket there
                            saves phase changes (each a few claim round trips) at the price of a few
                            re-expansions of low-degree vertices */
  uint32_t order_degree; /* with ADDS_F_PULL_SPLIT: only vertices with at least this many light edges are
                            expanded in bucket order below the split point; the others are expanded in the
                            bucket that finds them, whatever their distance (0, the default: every vertex in
                            bucket order).  A vertex of many light edges expanded before its distance is
                            final rescans all of them; one of few costs little, and running those out of
                            order needs fewer, fuller buckets */
  uint32_t pull_early;   /* with ADDS_F_PULL_SPLIT, ADDS_F_PULL_ALL_LIGHT and pull_delta: the split point's pull
                            may come as soon as at most this many items are queued or in flight (after
                            four times as many were seen) instead of after the last narrow bucket; the items
                            still queued are carried over into the wide buckets, where every edge is light,
                            so the last hops of the narrow buckets overlap the rest of the solve (0, the
                            default: the pull waits for the split point) */
} adds_params;

/* flags */
#define ADDS_F_HOST_OUT 0x1u  /* dist output buffers are host memory (copied back, then synchronised) */
#define ADDS_F_NO_SYNC 0x2u   /* do not synchronise; status of the solve is read by adds_last_status */
#define ADDS_F_ATOMIC_PROBE 0x4u /* probe queue counters with atomic (cache-bypassing) loads only */
#define ADDS_F_NO_LOCAL 0x8u  /* every improvement goes through the queues (no local continuation); a
                                 tuning choice: local continuation wins on high-diameter graphs */
#define ADDS_F_NO_WAIT 0x10u  /* flushes issue their atomicMins without waiting for the results and queue
                                 every improvement (one that lost becomes a stale entry, dropped when
                                 claimed): a round trip less per flush; pays on throughput-bound solves */
#define ADDS_F_REPLICA_PARTS 0x20u /* many-source solves (adds_sssp_multi and its captures): the items of
                                 source i go to sub-queue part i % parts, so that with parts = 8 each
                                 source's distance row is worked on mostly by one XCD and its L2; pays on
                                 many sources of a power-law graph (256 sources on RMAT-20: -6%), loses on
                                 grids (hops leave the wavefront's part, so local continuation stops) */
#define ADDS_F_NO_PULL 0x40u  /* no pull phases (see adds_graph_set_symmetric): every heavy phase relaxes
                                 its edges from their sources */
#define ADDS_F_PULL_EAGER 0x80u /* request a pull phase at the first heavy phase with any heavy edge and
                                 always take it, whatever the cost estimate says (tests, tuning) */
#define ADDS_F_PULL_ALL_LIGHT 0x100u /* with pull_delta: after the re-bucketing pull every edge is light
                                 (expanding a vertex relaxes its whole row; no heavy phases), which suits
                                 the sparse rest of a power-law solve */
#define ADDS_F_PULL_SPLIT 0x200u /* symmetric graphs with light_delta > delta: the buckets below the
                                 lightest heavy edge's bucket run their light phases only (no heavy edge
                                 can reach them), then one pull phase relaxes all their heavy edges at
                                 once; narrow buckets order the first light phases (fewer re-expansions)
                                 without pushing the heavy edges of the dense start of the solve */
#define ADDS_F_LOCAL_SPARSE 0x400u /* local continuation only after claims of at most chunk / 4 items: the
                                 sparse stretches (chains of single hops at the end of a bucket) keep it,
                                 full claims hand their improvements to every wavefront (ignored with
                                 ADDS_F_NO_LOCAL) */
#define ADDS_F_FULL_FIRST 0x800u /* a claim tries the open queues that hold a full chunk before the others
                                 (head bucket first otherwise): when the head bucket is down to its last
                                 items, wavefronts take the plentiful buckets ahead instead of contending */

typedef struct {
  double ms;                  /* hipEvent time of the solve (init + persistent kernel + cleanup) */
  uint64_t vertices_expanded; /* worklist items whose edges were scanned (reference "work", int) */
  uint64_t edges_scanned;     /* edges relaxed (reference "work", float) */
  uint64_t relaxations;       /* atomic min operations issued on dist */
  uint64_t pushes;            /* worklist entries written (including re-bucketed ones) */
  uint64_t stale_skips;       /* popped entries superseded by a smaller distance */
  uint64_t hub_tasks;         /* chunks created by splitting long edge ranges */
  uint64_t rebuckets;         /* popped entries moved to a later bucket */
  uint64_t pool_chunks_used;  /* worklist pool chunks consumed */
  double kernel_ms;           /* hipEvent time of the persistent solver kernel alone */
  uint64_t probes;            /* queue probes issued by wavefronts looking for work */
  uint64_t cas_failures;      /* lost claim races */
  uint64_t idle_iterations;   /* idle-loop iterations of all wavefronts */
  uint64_t late_items;        /* items expanded more than one bucket after their own bucket */
  /* wavefront time by activity, in device wall-clock ticks (100 MHz) summed over all wavefronts */
  uint64_t ticks_probe;       /* looking for and claiming work */
  uint64_t ticks_items;       /* loading and filtering claimed items, filing, recording long ranges */
  uint64_t ticks_edges;       /* relaxing the claimed items' edges and pushing the improvements */
  uint64_t ticks_idle;        /* termination tests, head advances and back-off while no work is visible */
  uint64_t local_items;       /* improvements processed by the wavefront that found them (no queue trip) */
  uint64_t heavy_edges;       /* of edges_scanned, those relaxed in heavy phases */
  uint64_t ticks_flush;       /* of ticks_items + ticks_edges, the time in push-buffer flushes */
  uint64_t pulls;             /* pull phases performed (see adds_graph_set_symmetric) */
  uint64_t pull_edges;        /* of edges_scanned, in-edges scanned by pull phases */
  uint64_t pull_improved;     /* vertices whose distance a pull phase lowered */
} adds_stats;

/* ---- errors / device ---- */
const char* adds_strerror(adds_status s);
/* Selects the HIP device used by later calls made from this thread (reference -g, support.cu:48-56). */
adds_status adds_set_device(int device);
/* Number of compute units of the current device (0 if none). */
uint32_t adds_device_cu_count(void);
/* Detail text of the last error raised on this thread (HIP call and message, or the failed check). */
const char* adds_last_error_detail(void);

/* ---- graphs ----
 * Replaces CSRGraph::copy_to_gpu (csr_graph.cu:309-328): copies a host CSR (row_ptr[n+1], col[m],
 * w[m] 4-byte weights) into device memory owned by the handle.  Validates monotone row_ptr, row_ptr[n]
 * == m, col < n, and (float) weights that are finite and not negative. */
adds_status adds_graph_create(const uint32_t* row_ptr, const uint32_t* col, const void* w, uint32_t n,
                              uint64_t m, adds_wtype wt, adds_graph** out);
/* Same from device pointers (copied device-to-device; no validation of contents). */
adds_status adds_graph_create_device(const uint32_t* row_ptr_dev, const uint32_t* col_dev,
                                     const void* w_dev, uint32_t n, uint64_t m, adds_wtype wt,
                                     adds_graph** out);
/* Overwrites the edge weights from a device buffer of m 4-byte words (used by the autograd binding). */
adds_status adds_graph_set_weights_device(adds_graph* g, const void* w_dev, void* stream);
/* Declares (symmetric != 0) that every edge (u, v, w) of g has its reverse (v, u, w) with the same weight
 * bits, or withdraws the declaration.  No reference counterpart: it enables pull phases, which relax a
 * settled bucket's heavy edges from their targets, reading each unsettled vertex's in-edges from its own
 * row; they are taken when the bucket's heavy edges (its push-direction work) outnumber the unsettled
 * vertices' heavy in-edges (on RMAT-24 with a hub source, bucket 0: 931M against 29M).  Distances are
 * the same either way.  A declaration that is false gives wrong distances; adds_graph_set_weights_device
 * withdraws it (also for solves captured before).  A union of graphs (batched solves) is symmetric when
 * every member is.  Graphs start undeclared. */
adds_status adds_graph_set_symmetric(adds_graph* g, int symmetric);
/* Tests on the device whether g is symmetric in that sense with every row in strictly ascending order of
 * destination (an edge whose reverse is missing, with other weight bits, or out of order in its row makes
 * it read as not symmetric: a graph with unsorted rows or repeated edges is never declared), stores 1 or 0
 * in *symmetric and declares the graph accordingly.  No reference counterpart; the drop-in binaries call it
 * once per graph.  About 0.1 s per billion edges. */
adds_status adds_graph_check_symmetric(adds_graph* g, int* symmetric);
void adds_graph_destroy(adds_graph* g);
adds_status adds_graph_info(const adds_graph* g, uint32_t* n, uint64_t* m, adds_wtype* wt);
/* Device pointers of the resident CSR (row_ptr, col, w); valid until the handle is destroyed. */
adds_status adds_graph_device_arrays(const adds_graph* g, const uint32_t** row_ptr, const uint32_t** col,
                                     const void** w);

/* The solver's copy of the rows for light threshold light_delta (built if needed, then synchronised):
 * m (col, w) pairs, each row light edges first (w <= light_delta; int weights: w <= floor(light_delta)),
 * then its heavy edges in non-decreasing weight band, band(w) = min(31, floor((w - light_delta) /
 * light_delta)) in float arithmetic (int weights converted); inside a class the order is unspecified.
 * light_counts: n light-edge counts.  Device pointers owned by the graph, valid until the next split
 * with another threshold or adds_graph_destroy.  No reference counterpart: inspection (tests). */
adds_status adds_graph_split_rows(adds_graph* g, float light_delta, const void** cw, const uint32_t** light_counts);

/* ---- solves ----
 * Replaces gg_main's timed solve (ads_int/kernel.cu:223-262): dist[v] = shortest distance from src.
 * dist is a device buffer of n words, or host memory with ADDS_F_HOST_OUT. */
adds_status adds_sssp(adds_graph* g, uint32_t src, void* dist, const adds_params* p, adds_stats* stats,
                      void* stream);

/* adds_sssp that allocates nothing: no reference counterpart; it is for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid allocation).  The solver workspace lives in a
 * fixed 98 MiB device array of the library, which each GPU the library runs on holds until the process
 * ends; the graph handle already owns its row buffers.  Limits: the graph's worklist estimate
 * 4 n + m / 16 (or p->pool_entries) must be at most 8,388,608 entries (about 2M vertices) and
 * nbuckets * parts at most 512; larger inputs return ADDS_EINVAL with the reason in
 * adds_last_error_detail().  dist must be a device buffer of n words; the call is synchronous
 * (ADDS_F_HOST_OUT and ADDS_F_NO_SYNC return ADDS_EINVAL) and cannot grow its pool (a solve that needs
 * more returns ADDS_EOVERFLOW).  Calls share the device's array: calls that overlap on one device race
 * (calls from threads of one process are serialised).  stats->ms is host wall time, kernel_ms is 0. */
adds_status adds_sssp_noalloc(adds_graph* g, uint32_t src, void* dist, const adds_params* p, adds_stats* stats,
                              void* stream);

/* Bulk-synchronous near-far SSSP, the GPU baseline the reference compares ADDS with (its nf_int /
 * nf_float runs: nf_int.patch, nf_float.patch over Galois' GPU sssp).  Rounds relax a near worklist of
 * vertices below a threshold; vertices improved past it wait in a far pile until the threshold moves up
 * by multiples of delta.  delta <= 0 uses the reference's 32 * mean weight / mean degree
 * (nf_int.patch:334; truncated to an integer for uint32 weights).  dist and flags as for adds_sssp
 * (ADDS_F_HOST_OUT only).  stats: ms = the whole solve (rounds synchronise with the host),
 * vertices_expanded = worklist items processed (the reference's NF work count), edges_scanned,
 * relaxations = atomicMins issued, probes = rounds, rebuckets = threshold moves.  Same distances as
 * adds_sssp, bit for bit; integer path sums above 2^32-2 saturate. */
adds_status adds_sssp_nearfar(adds_graph* g, uint32_t src, void* dist, float delta, uint32_t flags,
                              adds_stats* stats, void* stream);

/* k sources on one graph in one persistent launch with a shared worklist; dist is [k][n] (device, or
 * host with ADDS_F_HOST_OUT).  No reference counterpart (batched form). */
adds_status adds_sssp_multi(adds_graph* g, const uint32_t* srcs, uint32_t k, void* dist,
                            const adds_params* p, adds_stats* stats, void* stream);
/* k instances on possibly different graphs in one launch: when every handle is the same graph this is
 * adds_sssp_multi; otherwise the graphs are solved as one block-diagonal union CSR (built on first use and
 * cached in graphs[0], keyed by the members' handles and weight versions: a weight update of any member
 * rebuilds it) from k sources.  dists[i] receives n_i words.  No reference counterpart. */
adds_status adds_sssp_batched(adds_graph* const* graphs, const uint32_t* srcs, uint32_t k,
                              void* const* dists, const adds_params* p, adds_stats* stats, void* stream);
/* Status of the last solve issued with ADDS_F_NO_SYNC on this handle (synchronises its stream). */
adds_status adds_last_status(adds_graph* g, adds_stats* stats);

/* Solves captured in a HIP graph (the reference's device-side launch of manager + workers,
 * ads_int/kernel.cu:130-149, becomes one host-side graph replay).  No reference counterpart as an API.
 *
 * adds_solve_graph_create captures one multi-source solve of g (dist fill, persistent kernel, worklist
 * cleanup) bound to the device buffers srcs_dev [k] and dist_dev ([k][n]).  The solve graph owns its own
 * worklist workspace, so eager solves, autotuning or other captures on the same handle never move memory
 * under it.  The handle g must outlive the solve graph; a later weight update of g is seen by replays.
 *
 * adds_solve_graph_create_batched captures k instances on different graphs in one persistent launch (the
 * block-diagonal union of adds_sssp_batched, built for and owned by the solve graph: a snapshot of the
 * members' CSR at capture, so later weight updates of the members are not seen).  srcs is host memory;
 * dists[i] is a device buffer of n_i words that every replay fills.
 *
 * adds_solve_graph_launch enqueues one replay on `stream`; adds_solve_graph_status waits for the last
 * replay and returns its status and counters (ms = replay time).  The pool of a capture is fixed: a replay
 * that exhausts it reports ADDS_EOVERFLOW (capture again with a larger pool_entries). */
typedef struct adds_solve_graph adds_solve_graph;
adds_status adds_solve_graph_create(adds_graph* g, const uint32_t* srcs_dev, uint32_t k, void* dist_dev,
                                    const adds_params* p, adds_solve_graph** out);
adds_status adds_solve_graph_create_batched(adds_graph* const* graphs, const uint32_t* srcs, uint32_t k,
                                            void* const* dists, const adds_params* p, adds_solve_graph** out);
adds_status adds_solve_graph_launch(adds_solve_graph* sg, void* stream);
adds_status adds_solve_graph_status(adds_solve_graph* sg, adds_stats* stats);
void adds_solve_graph_destroy(adds_solve_graph* sg);

/* Picks the fastest tuning point for this workload (k sources on g; k == 1 is a plain solve) by
 * coordinate descent over delta, light threshold, bucket ring length, sub-queues and open buckets,
 * workgroups per CU, claim size, reference chunk size and local continuation,
 * timing `trials` solves per point.  Every point's distances must be identical; a difference is
 * reported as an error.  No reference counterpart (the reference adapts delta online, wl.h:851-1310). */
adds_status adds_autotune(adds_graph* g, const uint32_t* srcs, uint32_t k, uint32_t trials, adds_params* best,
                          double* best_ms);
/* The same search for adds_sssp_batched over these graphs (tunes the block-diagonal union that
 * adds_sssp_batched builds and caches); the result is meant for adds_sssp_batched.  No reference
 * counterpart (batched forms are an addition). */
adds_status adds_autotune_batched(adds_graph* const* graphs, const uint32_t* srcs, uint32_t k, uint32_t trials,
                                  adds_params* best, double* best_ms);

/* Diagnostic with no reference counterpart: cross-XCD message-passing litmus test of the ordering recipe
 * the solver uses (relaxed agent-scope atomics ordered by waiting for completion).  `pairs` producer /
 * consumer workgroup pairs run `rounds` rounds; variant 0 publishes with a store, 1 with a fetch_add,
 * 2 checks atomicMin visibility, 3 is variant 0 without the wait (a control: it may show reordering).
 * *forbidden counts stale observations (must be 0 for variants 0-2), *lost counts rounds whose bounded
 * wait expired, *cross_xcd_pairs (may be NULL) the pairs whose two workgroups ran on different XCDs, read
 * from the XCC_ID hardware register of each workgroup. */
adds_status adds_litmus(uint32_t pairs, uint32_t rounds, uint32_t variant, uint64_t* forbidden, uint64_t* lost,
                        uint64_t* cross_xcd_pairs);
/* Diagnostic: the XCD (XCC_ID register) each workgroup of a launch of `workgroups` x `threads` ran on,
 * into xcc (host, `workgroups` words).  The solver groups its termination words by workgroup index mod 8,
 * which is the XCD under round-robin dispatch; this shows the placement. */
adds_status adds_xcd_map(uint32_t workgroups, uint32_t threads, uint32_t* xcc);
/* Diagnostic: the random-gather ceiling of the device for a distance array of n words.  A stream of m
 * (m <= 2^31) 8-byte (col, w) pairs with uniformly random col < n is read once and d[col] gathered for
 * every pair, with 8 gathers per lane in flight, at one to eight 256-thread workgroups per CU;
 * *gathers_per_s receives the best rate.  The solver's edge windows issue the same pattern (one d[v]
 * gather per edge), so this bounds their edge rate on graphs whose targets have no locality. */
adds_status adds_gather_ceiling(uint32_t n, uint64_t m, double* gathers_per_s);

/* ---- backward (float graphs) ----
 * No reference counterpart (the reference keeps no predecessor tree: ads_float/kernel.cu:82).  Given
 * forward distances dist [k][n] (device) from the k sources srcs (host array) and dL/d(dist) grad_dist
 * [k][n] (device), writes dL/dw into grad_w [m] (device).
 * Tree of each source, over its reached vertices: the source is the root; a vertex's parent edge is the
 * smallest edge id e = (u -> v) with d[u] < d[v] and fl(d[u] + w_e) == d[v]; a vertex with no such edge
 * (reached only through zero weights or adds that round away) takes, in rounds, the smallest tight edge
 * from an equal-distance vertex whose parent was chosen in an earlier round.
 * Gradient: with S = 62 - eM - eN (largest reached |g| = f 2^eM, 0.5 <= f < 1; 2^eN >= number of nonzero
 * reached g), dL/dw_e = ldexpf((float)GW_e, -S) where GW_e = sum of rint(g_s[t] 2^S) over the (s, t)
 * whose tree path uses e, summed exactly in 64-bit integers: the exact sum of the (scaled) inputs rounded
 * once, independent of any order -- bit-identical run to run and to the CPU reference.  ms: hipEvent
 * time of the pass.  ADDS_EINVAL when a reached grad_dist value is not finite or dist is not the
 * shortest-path solution from srcs. */
adds_status adds_sssp_backward_f32(adds_graph* g, const uint32_t* srcs, uint32_t k, const float* dist,
                                   const float* grad_dist, float* grad_w, double* ms, void* stream);

/* ---- host CSR, .gr files and distance text (reference csr_graph.cu / support.cu) ---- */
typedef struct {
  uint32_t n;
  uint64_t m;
  uint32_t* row_ptr; /* n+1 */
  uint32_t* col;     /* m */
  uint32_t* w;       /* m 4-byte weights (uint32 or float bits) */
  adds_wtype wtype;
} adds_host_csr;

void adds_host_csr_free(adds_host_csr* g);
/* Galois .gr v1 reader (csr_graph.cu:203-296).  sizeEdgeTy == 0 gives all-zero weights. */
adds_status adds_read_gr(const char* path, adds_wtype wt, adds_host_csr* out);
adds_status adds_write_gr(const char* path, const adds_host_csr* g);
/* Per-vertex text output (support.cu:78-96): "%d %u" (int), "%d %f" (float), "%d INF".  path "-" =
 * stdout. */
adds_status adds_write_dist_text(const char* path, const uint32_t* dist, uint32_t n, adds_wtype wt);

/* ---- text graph formats (no reference counterpart: the reference's graph lists name SuiteSparse and
 * DIMACS graphs that were converted to .gr outside it, graph_list_int / graph_list_float) ----
 * Shortest paths need non-negative weights, so a weight is |value|: the nearest float for ADDS_W_F32,
 * the nearest integer for ADDS_W_U32 (above 2^31-1, or NaN: ADDS_EOVERFLOW).  Edges keep file order
 * within a row; duplicates and self loops are kept.  Parse errors give ADDS_EIO with the line in
 * adds_last_error_detail().
 *
 * MatrixMarket coordinate file (SuiteSparse): entry (i, j, v) is the edge i-1 -> j-1; symmetric,
 * skew-symmetric and hermitian files also give j-1 -> i-1 for i != j; pattern files give weight 1;
 * complex entries use their real part; n = max(rows, cols). */
adds_status adds_read_mtx(const char* path, adds_wtype wt, adds_host_csr* out);
/* DIMACS shortest-path file (9th DIMACS Implementation Challenge): "c" comments, one "p sp <n> <m>",
 * then m arcs "a <u> <v> <w>", 1-based. */
adds_status adds_read_dimacs(const char* path, adds_wtype wt, adds_host_csr* out);

/* ---- synthetic graphs (seeded, deterministic; symmetric, no self loops, no duplicate edges) ---- */
typedef enum {
  ADDS_WK_U32_UNIFORM = 0, /* uint32 uniform in [lo, hi] */
  ADDS_WK_F32_UNIFORM = 1, /* float uniform in [lo, hi) */
  ADDS_WK_F32_LOGNORMAL = 2 /* float exp(N(0, lo^2)) */
} adds_wkind;

adds_status adds_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                          uint64_t seed, adds_wkind wk, double lo, double hi, adds_host_csr* out);
/* rows x cols 4-neighbour grid plus `shortcuts` random long-range undirected pairs. */
adds_status adds_gen_grid(uint32_t rows, uint32_t cols, uint64_t shortcuts, uint64_t seed, adds_wkind wk,
                          double lo, double hi, adds_host_csr* out);
/* Erdos-Renyi-style graph: n*avg_degree/2 uniform random undirected pairs. */
adds_status adds_gen_er(uint32_t n, double avg_degree, uint64_t seed, adds_wkind wk, double lo, double hi,
                        adds_host_csr* out);
/* Re-draws every weight of g with the generators' rule (so a generator call with this seed and weight
 * spec yields the same weights) and sets g->wtype accordingly. */
adds_status adds_gen_weights(adds_host_csr* g, uint64_t seed, adds_wkind wk, double lo, double hi);
/* Highest out-degree vertex (lowest id on ties). */
uint32_t adds_max_degree_vertex(const adds_host_csr* g);
/* Up to k distinct vertices sampled uniformly (seeded) from the component reached from `root`. */
uint32_t adds_sample_component(const adds_host_csr* g, uint32_t root, uint32_t k, uint64_t seed,
                               uint32_t* out);
/* n_r (vertices with finite distance) and m_r (sum of their out-degrees) of a distance array. */
void adds_reached_counts(const adds_host_csr* g, const uint32_t* dist, uint64_t* n_r, uint64_t* m_r);

#ifdef __cplusplus
}
#endif
#endif /* ADDS_H */
