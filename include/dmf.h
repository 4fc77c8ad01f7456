/*
 * dmf.h — C-ABI of the MI355X dynamic max-flow engine.
 *
 * Plain C types only (int32/int64 arrays, sizes, opaque handles). No C++ or
 * torch types cross this boundary and no exception escapes it: every entry
 * point returns an int status (DMF_OK = 0, negative on error).
 *
 * The reference (dyn_maxflow) has no library API; its interface is the pair
 * of on-disk formats, the bench drivers' argv/stdout, and two in-process host
 * functions.  Each declaration below names the reference code it replaces.
 */
#ifndef DMF_H
#define DMF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
#define DMF_OK 0
#define DMF_EINVAL (-1)   /* bad argument (vertex id out of range, s == t, ...) */
#define DMF_EIO (-2)      /* file cannot be opened / short read / short write   */
#define DMF_ENOARC (-3)   /* an update names an arc absent from the CSR         */
#define DMF_ENOMEM (-4)   /* host or device allocation failed                   */
#define DMF_EDEVICE (-5)  /* HIP runtime error                                  */
#define DMF_ENOCONV (-6)  /* round cap reached before convergence               */
#define DMF_EFORMAT (-7)  /* file or CSR inconsistent (negative sizes, ...)     */
#define DMF_EADMIT (-8)   /* serving: request exceeds the admission budget      */
#define DMF_EAGAIN (-9)   /* serving: result not ready within the timeout       */
#define DMF_ESIZE (-10)   /* input larger than a fixed workspace allows         */

/* Human-readable text of a status code (static storage). */
const char* dmf_strerror(int status);

/* ---------------------------------------------------------------- host CSR */
/*
 * Bidirectional CSR exactly as the reference stores it
 * (code/graph.hpp:21-30, written by input_create/create_bin.cpp:18-64):
 *   row[n+1]  offsets            (graph::indexofNodes)
 *   col[m]    arc head           (graph::edgeList)
 *   cap[m]    arc capacity       (graph::edgeLen)
 *   rev[m]    index of the reverse arc (graph::parallel_edge)
 * Memory returned by the dmf_* producers is owned by the library and released
 * with dmf_csr_free.
 */
typedef struct dmf_csr {
  int32_t n;
  int32_t m;
  int32_t* row;
  int32_t* col;
  int32_t* cap;
  int32_t* rev;
} dmf_csr;

void dmf_csr_free(dmf_csr* g);
/* Release any array returned by a dmf_* producer (update arrays, edge lists). */
void dmf_free(void* p);

/* Structural check of a CSR; the reference has none (its readers trust the
 * file).  Offsets start at 0, never decrease and end at m; heads are in
 * [0, n); capacities are >= 0; rev[e] is in [0, m), rev[rev[e]] == e, and the
 * head of rev[e] is the tail of e.  DMF_OK, DMF_EINVAL (null arrays, negative
 * sizes) or DMF_EFORMAT.  dmf_read_graph_bin and every entry point that builds
 * an engine from a caller's CSR run it first, so a malformed graph is an error
 * status and never an out-of-range index on the device.  One host pass, O(n + m). */
int dmf_csr_check(const dmf_csr* g);

/* Replaces graph::loadGraph (code/graph.hpp:202-223).  Layout, all int32 LE:
 * [n][m][row n+1][col m][cap m][rev m].  Unlike the reference, short reads and
 * inconsistent headers are reported (DMF_EIO / DMF_EFORMAT). */
int dmf_read_graph_bin(const char* path, dmf_csr* out);

/* Replaces writeCSR (input_create/create_bin.cpp:18-64); byte-identical output. */
int dmf_write_graph_bin(const char* path, const dmf_csr* g);

/* Replaces readEdgeList + writeCSR's construction (input_create/create_bin.cpp:67-124):
 * arcs in first-appearance order per tail, duplicate (u,v) capacities summed,
 * a reverse arc of capacity 0 added when absent, self-loops dropped,
 * n = max id + 1.  Hash-indexed, so linear in k (the reference's per-edge linear
 * search is quadratic on hubs).  k = number of input edges. */
int dmf_build_csr(int64_t k, const int32_t* u, const int32_t* v, const int32_t* w,
                  dmf_csr* out);

/* Replaces loadUpdateFile (code/update.hpp:52-89).  Layout, all int32 LE:
 * [k][src k][dst k][w k]; w is the new absolute capacity of arc src->dst.
 * Returns the k records as stored (self-loops included; the engine skips them,
 * as the reference reader does at code/update.hpp:82). */
int dmf_read_updates_bin(const char* path, int32_t* k, int32_t** src, int32_t** dst,
                         int32_t** w);

/* Replaces input_create/create_update_bin.cpp:11-52 (the .txt -> .bin step);
 * byte-identical output for the same records. */
int dmf_write_updates_bin(const char* path, int32_t k, const int32_t* src,
                          const int32_t* dst, const int32_t* w);

/* ------------------------------------------------------ seeded input tools */
/* (No reference counterpart with a seed: input_create/update_gen.cpp:31-32
 * seeds from std::random_device.  These are the reproducible replacements.) */

/* Directed R-MAT edge list: 2^scale * edge_factor samples, quadrant
 * probabilities a,b,c (d = 1-a-b-c), no noise; self-loops and duplicate (u,v)
 * dropped (first kept); capacities uniform in [cap_lo, cap_hi].  xoshiro256**
 * seeded through splitmix64. */
int dmf_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                 uint64_t seed, int32_t cap_lo, int32_t cap_hi, int64_t* k,
                 int32_t** u, int32_t** v, int32_t** w);

/* R-MAT over n vertices (any n >= 2): `samples` draws on the enclosing power
 * of two; draws with an endpoint >= n are dropped like self-loops and
 * duplicates.  For n = 2^scale and samples = n * edge_factor the output equals
 * dmf_gen_rmat's.  Used for the uneven independent instances of the batcher. */
int dmf_gen_rmat_n(int32_t n, int64_t samples, double a, double b, double c, uint64_t seed,
                   int32_t cap_lo, int32_t cap_hi, int64_t* k, int32_t** u, int32_t** v,
                   int32_t** w);

/* rows x cols 4-neighbour grid, id = r*cols + c, every neighbour pair in both
 * directions with independent capacities uniform in [cap_lo, cap_hi]. */
int dmf_gen_grid(int32_t rows, int32_t cols, uint64_t seed, int32_t cap_lo,
                 int32_t cap_hi, int64_t* k, int32_t** u, int32_t** v, int32_t** w);

/* s = argmax out-degree, t = argmax in-degree with t != s, ties -> lowest id,
 * degrees counted over the input edge list. */
int dmf_pick_source_sink(int32_t n, int64_t k, const int32_t* u, const int32_t* v,
                         int32_t* s, int32_t* t);

/* One update batch of `count` distinct input edges, sampled without replacement.
 * Each is an increase with probability inc_frac, else a decrease; the new
 * capacity is uniform in [cur+1, cap_hi] (increase) or [cap_lo, cur-1]
 * (decrease); an edge already at that bound is skipped and another drawn.
 * cur_w (length k) holds the current capacities and, if apply != 0, is updated
 * in place so successive batches build on each other.  Outputs have `*out_k`
 * records (< count only when the graph runs out of eligible edges). */
int dmf_gen_updates(int64_t k, const int32_t* u, const int32_t* v, int32_t* cur_w,
                    int64_t count, double inc_frac, int32_t cap_lo, int32_t cap_hi,
                    uint64_t seed, int apply, int32_t* out_k, int32_t** src,
                    int32_t** dst, int32_t** w);

/* ------------------------------------------------------------------ engine */
/* Variants: the three reference paths (code/README.md:36-42). */
#define DMF_VARIANT_DATA 0  /* worklist-driven       (code/dyn_data.h)  */
#define DMF_VARIANT_TOPO 1  /* all-vertex rounds     (code/dyn_topo.h)  */
#define DMF_VARIANT_PP 2    /* push-pull over a cut  (code/dyn_pp.h)    */
/* A push-pull batch continues from the state of a push-pull static solve
 * (data and topology batches in between are fine).  After a static solve of
 * another variant (or a reset) a batch that asks for push-pull runs
 * data-driven instead; dmf_counters.variant reports the variant that ran.
 * The learned scheduler picks push-pull only from such a state. */

#define DMF_SCHED_FIXED 0   /* global relabel every `relabel_interval` rounds */
#define DMF_SCHED_LEARNED 1 /* on-device MLP decides relabel + next variant   */

typedef struct dmf_params {
  int32_t variant;          /* DMF_VARIANT_*                                     */
  int32_t kc;               /* push/relabel cycles per vertex per round (>= 1)   */
  int32_t relabel_interval; /* rounds between global relabels (1 = reference)   */
  int32_t scheduler;        /* DMF_SCHED_*                                       */
  int32_t max_rounds;       /* safety cap on rounds per solve (0 = default)     */
  int32_t mode;             /* 0 = persistent one-launch solve, 1 = stepped,
                               2 = stepped with the BFS levels replayed from
                               captured hipGraph segments, 3 = round graphs:
                               each data-driven round of the fixed schedule
                               (relabel, worklist, push, repair) replayed
                               from one captured hipGraph with its loop
                               conditions on the device, one host check per
                               round; schedules it does not cover (learned,
                               cascade, topology-driven and push-pull
                               batches, lazy start) run as mode 2           */
  int32_t lazy_start;       /* data/topo batches (stepped drivers): skip the
                               global relabel before the first round when the
                               previous solve ended on fresh labels.  The
                               learned scheduler makes this choice itself.    */
  int32_t cascade;          /* data-driven rounds (stepped drivers): most push
                               sub-rounds chained on the device after each
                               round's push, each taking only the vertices
                               the pass before it gave excess, and relabels
                               that stop once every active vertex is labelled
                               (convergence is still checked on a full one).
                               0 = none, the reference's rounds; the learned
                               scheduler uses 1024 when this is 0.           */
} dmf_params;

/* Fills the reference's fixed schedule: data variant, kc = max(1, m/n),
 * global relabel every round (code/dyn_data_bench.cu:529-562, :65-79). */
void dmf_params_default(const dmf_csr* g, dmf_params* p);

/* Work counters of the last solve (static or batch), accumulated on device. */
typedef struct dmf_counters {
  uint64_t arc_visits;     /* adjacency entries examined                   */
  uint64_t vertex_visits;  /* vertices examined by any phase               */
  uint64_t pushes;         /* push operations                              */
  uint64_t relabels;       /* local relabels                               */
  uint64_t repairs;        /* label-repair saturations                     */
  uint64_t bfs_levels;     /* levels over all global relabels              */
  uint64_t global_relabels;
  uint64_t rounds;         /* push/repair rounds                           */
  uint64_t updates;        /* update records applied                       */
  uint64_t atomics;        /* global atomic RMW operations issued          */
  uint64_t launches;       /* kernel launches                              */
  double device_ms;        /* HIP-event span of the solve on its stream    */
  uint64_t bfs_arc_visits;     /* arc visits of the BFS level launches       */
  uint64_t bfs_vertex_visits;  /* vertex visits of the BFS level launches    */
  uint64_t bfs_level_launches; /* BFS level launches timed (profiling on)    */
  double bfs_level_ms;         /* their summed HIP-event durations           */
  uint64_t variant;            /* variant the solve ran (DMF_VARIANT_*)      */
} dmf_counters;

typedef struct dmf_engine dmf_engine;

/* Copies the CSR into a device-resident residual graph (cf. the per-call
 * allocation + H2D copies of code/dyn_data_bench.cu:182-217, kept resident
 * here across batches).  s,t are fixed for the engine's life. */
int dmf_engine_create(const dmf_csr* g, int32_t s, int32_t t, dmf_engine** out);
int dmf_engine_destroy(dmf_engine* e);

/* Run all later work on this HIP stream (hipStream_t passed as void*); NULL =
 * the engine's own stream. */
int dmf_engine_set_stream(dmf_engine* e, void* hip_stream);

/* Restore original capacities and zero flow (no solve). */
int dmf_engine_reset(dmf_engine* e);

/* Static solve from zero flow: replaces staticMaxFlow
 * (code/dyn_data_bench.cu:8-97, code/dyn_topo_bench.cu:16-75,
 * code/dyn_pp_bench.cu:155-239).  *flow = max-flow value (deficits folded). */
int dmf_engine_static_solve(dmf_engine* e, const dmf_params* p, int64_t* flow);

/* Apply one batch of absolute capacity updates and re-converge: replaces
 * OnAdd_kernel + Incremental (code/dyn_data.h:186-206,
 * code/dyn_data_bench.cu:98-170; topo/pp counterparts).  Host arrays; the copy
 * to HBM is outside the device-time span.  Duplicate arcs in one batch: the
 * last record wins.  A missing arc fails the whole batch with DMF_ENOARC, and
 * an out-of-range vertex id or a negative capacity with DMF_EINVAL; either
 * leaves the state untouched.  Self-loops are skipped. */
int dmf_engine_apply_batch(dmf_engine* e, int32_t k, const int32_t* src,
                           const int32_t* dst, const int32_t* w,
                           const dmf_params* p, int64_t* flow);

/* Same, with the three arrays already resident in device memory. */
int dmf_engine_apply_batch_device(dmf_engine* e, int32_t k, const int32_t* d_src,
                                  const int32_t* d_dst, const int32_t* d_w,
                                  const dmf_params* p, int64_t* flow);

/* Current max-flow value (excess[t] plus the deficits of v != s, as
 * recalculate_max_flow_kernel_39, code/dyn_data.h:208-215). */
int dmf_engine_flow(dmf_engine* e, int64_t* flow);

/* Source side of the minimum cut S_min (vertices reachable from s in the
 * residual graph of a maximum flow), side[v] = 1 for v in S_min.  On a
 * preflow the search also starts from every non-sink vertex with positive
 * excess (that excess flows back to s once the flow is completed, so the set
 * is the same); after dmf_engine_complete_flow it is plain reachability from
 * the sources.  No reference counterpart (the reference never reports a cut
 * set). */
int dmf_engine_min_cut(dmf_engine* e, uint8_t* side);

/* Turns the maximum preflow of the last solve into a maximum flow of the same
 * value, on the device: every non-terminal deficit is cancelled by pulling
 * flow back from the sinks, and every non-terminal excess stranded on the
 * source side is returned to the sources.  Afterwards excess[v] == 0 for every
 * v that is neither a source nor a sink, excess[t] is the flow value and the
 * per-arc flows cap - res satisfy capacity and conservation.  *flow (may be
 * NULL) receives the value, equal to dmf_engine_flow before the call.  Runs
 * outside any batch; later batches continue from the completed state.
 * DMF_ENOCONV if an imbalance cannot be routed (only possible when the last
 * solve did not converge).  Supersedes the deficit fold of
 * recalculate_max_flow_kernel_39 (code/dyn_data.h:208-215), which leaves the
 * state a preflow. */
int dmf_engine_complete_flow(dmf_engine* e, int64_t* flow);

/* Copy device state to host for checking: res/cap per arc (m), excess per
 * vertex (n), height per vertex (n).  Any pointer may be NULL. */
int dmf_engine_read_state(dmf_engine* e, int32_t* res, int32_t* cap, int64_t* excess,
                          int32_t* height);

int dmf_engine_counters(dmf_engine* e, dmf_counters* c);

/* Per-launch timing of the BFS level kernel: with on != 0, the stepped
 * driver brackets every level launch with HIP events on the engine's stream
 * (bfs_level_ms, bfs_level_launches in dmf_counters; graph-replayed levels
 * are timed per segment). */
int dmf_engine_set_profiling(dmf_engine* e, int on);

/* Bytes of device memory held by this engine. */
int64_t dmf_engine_device_bytes(const dmf_engine* e);

/* Upper bound of dmf_engine_create's device bytes for a graph of n vertices
 * and m arcs (update buffers excluded: 16 bytes per reserved update). */
int64_t dmf_engine_footprint(int32_t n, int32_t m);

/* Sizes the update buffers for batches of up to k records now, so that later
 * batches of that size allocate nothing. */
int dmf_engine_reserve_updates(dmf_engine* e, int32_t k);

/* Tries a few (lane_max, wide_t) degree-bucket thresholds -- the degree up to
 * which one lane scans a vertex, and above which a vertex is split into chunks
 * spread over many waves -- with a static solve each, keeps the fastest and
 * resets the engine.  No reference counterpart (its kernels use one thread per
 * vertex). */
int dmf_engine_autotune(dmf_engine* e, const dmf_params* p, int32_t* lane_max, int32_t* wide_t);

/* dmf_engine_autotune, then, for the fixed schedule (p->scheduler ==
 * DMF_SCHED_FIXED), the global-relabel interval: a static solve from reset is
 * timed under each interval in {1, 2, 3, 5, 8} with the chosen thresholds and
 * the fastest is written to p->relabel_interval (the learned scheduler times
 * its own relabels and keeps p unchanged).  Leaves the engine reset.  No
 * reference counterpart (the reference relabels every round). */
int dmf_engine_autotune_schedule(dmf_engine* e, dmf_params* p, int32_t* lane_max, int32_t* wide_t);

/* ---------------------------------------------------- batched instances */
/* Many independent (graph, s, t) instances kept resident as one disjoint union
 * and solved together: every phase of a solve covers all instances at once
 * (with the persistent driver, one launch per batch for all of them).
 * No reference counterpart (the reference solves one graph per process). */
int dmf_batch_create(int32_t ninst, const dmf_csr* graphs, const int32_t* s, const int32_t* t,
                     dmf_engine** out);
int32_t dmf_engine_instances(const dmf_engine* e);
/* flows[ninst] receives each instance's max-flow value. */
int dmf_batch_static_solve(dmf_engine* e, const dmf_params* p, int64_t* flows);
/* counts[i] update records for instance i, concatenated in instance order in
 * src/dst/w with instance-local vertex ids. */
int dmf_batch_apply(dmf_engine* e, const int32_t* counts, const int32_t* src, const int32_t* dst,
                    const int32_t* w, const dmf_params* p, int64_t* flows);
int dmf_batch_flows(dmf_engine* e, int64_t* flows);

/* ----------------------------------------------------- streaming server */
/* No reference counterpart (the reference drivers solve one update file per
 * process).  Residual graphs stay resident; update batches are solved as they
 * arrive.  Sessions (one resident graph each) are admitted only if their
 * device footprint fits hbm_budget; submitted batches are admitted, in
 * arrival order, into a device staging arena of staging_bytes when their
 * bytes fit, else they wait until solves free space.  Each session's batches
 * are applied in submission order by its own host thread and stream, so
 * solves of different sessions overlap on the device and staging copies
 * overlap with solves.  Thread-safe; results are collected once per ticket. */
typedef struct dmf_server dmf_server;
typedef struct dmf_server_config {
  int64_t hbm_budget;     /* device bytes the server may hold; 0 = 90% of free */
  int64_t staging_bytes;  /* staging arena for admitted batches; 0 = budget / 8 */
  int32_t max_sessions;   /* 0 = 64 */
  int32_t pad;
  dmf_params params;      /* solve parameters of every batch (kc <= 0: per-graph default) */
} dmf_server_config;
typedef struct dmf_result {
  uint64_t ticket;
  int32_t session;
  int32_t status;         /* DMF_OK or the batch's error */
  int64_t flow;           /* max-flow value after this batch */
  int32_t k;              /* update records in the batch */
  int32_t pad;
  double submit_ms, admit_ms, start_ms, done_ms; /* server clock */
  double device_ms;       /* device time of the solve */
  uint64_t arc_visits, rounds, global_relabels;
} dmf_result;
typedef struct dmf_server_stats {
  int32_t sessions, in_flight, queued, pad;
  uint64_t completed;
  int64_t hbm_budget, hbm_held, staged_bytes, peak_staged_bytes;
  double clock_ms;
} dmf_server_stats;
void dmf_server_config_default(dmf_server_config* c);  /* learned scheduler, stepped */
int dmf_server_create(const dmf_server_config* cfg, dmf_server** out);
/* Builds a resident session (engine + static solve); DMF_EADMIT if its
 * footprint plus max_batch update slots does not fit the remaining budget. */
int dmf_server_open(dmf_server* s, const dmf_csr* g, int32_t src, int32_t snk, int32_t max_batch,
                    int32_t* session, int64_t* static_flow);
/* Queues one batch (copied; returns at once).  DMF_EADMIT if k exceeds the
 * session's max_batch or the batch alone exceeds the staging arena. */
int dmf_server_submit(dmf_server* s, int32_t session, int32_t k, const int32_t* src,
                      const int32_t* dst, const int32_t* w, uint64_t* ticket);
/* Blocks until the ticket's result is ready (timeout_ms < 0: no limit;
 * DMF_EAGAIN on timeout) and returns it; each ticket is returned once. */
int dmf_server_wait(dmf_server* s, uint64_t ticket, double timeout_ms, dmf_result* out);
int dmf_server_get_stats(dmf_server* s, dmf_server_stats* out);
/* Finishes the session's queued batches, then frees it. */
int dmf_server_close(dmf_server* s, int32_t session);
int dmf_server_destroy(dmf_server* s);

/* ------------------------------------------ reference in-process entry points */
/* int recalculate_max_flow(graph&, std::vector<update>, int batchSize, int s,
 * int t, int kernel_cycles)  (code/dyn_data_bench.cu:171-298,
 * code/dyn_topo_bench.cu:120-239, code/dyn_pp_bench.cu:469-621):
 * fresh device state, static solve, the whole update list as one batch,
 * deficit fold, flow.  ms_static / ms_batch receive the two HIP-event spans
 * the reference prints ("static GPU Time", "DYN TIME"); either may be NULL. */
int dmf_recalculate_max_flow(const dmf_csr* g, int32_t k, const int32_t* src,
                             const int32_t* dst, const int32_t* w, int32_t s,
                             int32_t t, int32_t kc, int32_t variant, int64_t* flow,
                             double* ms_static, double* ms_batch);

/* int static_recalculate(...)  (code/dyn_data_bench.cu:325-492 and
 * counterparts): capacities set from the updates, then one static solve from
 * scratch.  ms receives "STATIC UPDATE GPU Time". */
int dmf_static_recalculate(const dmf_csr* g, int32_t k, const int32_t* src,
                           const int32_t* dst, const int32_t* w, int32_t s, int32_t t,
                           int32_t kc, int32_t variant, int64_t* flow, double* ms);

/* dmf_static_recalculate without any allocation: no reference counterpart.
 * For callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc): the residual state lives in a fixed
 * workspace of about 46 MiB compiled into the library, which every GPU the
 * library runs on holds until the process ends.  Graphs of at most 65,536
 * vertices and 1,048,576 arcs and at most 1,048,576 update records; larger
 * inputs fail with DMF_ESIZE before any device work.  The caller passes host
 * arrays as for dmf_static_recalculate and a HIP stream (NULL = the default
 * stream); the solve runs on the stepped driver with the fixed schedule and
 * returns the same flow.  Calls on one GPU share the workspace: calls that
 * overlap on one GPU race, so serialise them.  No device memory, stream, event,
 * pinned buffer or graph is created. */
int dmf_static_recalculate_noalloc(const dmf_csr* g, int32_t k, const int32_t* src,
                                   const int32_t* dst, const int32_t* w, int32_t s,
                                   int32_t t, int32_t kc, int32_t variant, void* hip_stream,
                                   int64_t* flow);

/* Number of HIP devices visible (0 on a machine without a GPU). */
int dmf_device_count(void);

#ifdef __cplusplus
}
#endif
#endif /* DMF_H */
