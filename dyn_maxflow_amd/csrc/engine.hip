// MI355X (gfx950) dynamic max-flow engine: host side and the C-ABI declared
// in include/dmf.h.  The device phases, the stepped kernels and the persistent
// cooperative kernel live in phases.hpp.
//
// Residual graph in HBM (structure of arrays, int32 unless noted):
//   row[n+1] col[m] rev[m]     CSR of forward + reverse arcs (code/graph.hpp:21-30)
//   cap[m] cap0[m]             current / original capacity  (d_weight)
//   res[m]                     residual of arc e             (d_residual_capacity)
//   rres[m]                    residual of rev[e], kept beside e so that scans
//                              over the reverse residual graph read one
//                              coalesced stream              (d_rev_residual_capacity)
//   excess[n] (int64)          net inflow                    (d_excess, int32 in the reference)
//   height[n]                  push-relabel label; doubles as the BFS level
//                              (n = unlabelled), replacing the reference's
//                              separate d_level/d_height pair
//   order[n]                   vertices in BFS visit order (frontiers are
//                              contiguous slices of it)
//   wl[n] wl2[n] part[n]       worklists (active / pull / stranded part)
//   side[n] pfinal[n] (u8)     push-pull partition (d_partB, d_partfinal)
//   htab (16-byte slots)       open-addressing (tail,head) -> arc index, so an
//                              update finds its arc in O(1) instead of the
//                              reference's scan of the tail's adjacency
//   wq_vert/wq_base            wide-vertex chunk queues; wide_vert/wide_base
//                              the static list of every vertex of degree > 1024
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/dmf.h"
#include "abi_guard.hpp"
#include "phases.hpp"
#include "sched.hpp"

using namespace dmfdev;

namespace {
#include "sched_weights.inc"
}

// ============================================================ host engine

struct dmf_engine {
  int32_t n = 0, m = 0, s = 0, t = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  Dev D{};
  Ctrl* h_ctrl = nullptr;  // pinned mirror
  int32_t* d_cap0 = nullptr;
  int32_t* d_usrc = nullptr;
  int32_t* d_udst = nullptr;
  int32_t* d_uw = nullptr;
  int32_t* d_uarc = nullptr;
  int32_t* d_urev = nullptr;
  int32_t* d_udelta = nullptr;
  int32_t ucap = 0;
  int coop_blocks = 0;
  int num_cu = 0;
  // Most blocks one stepped launch uses: a few per CU, each wave looping over
  // its share (every kernel is grid-stride).  A grid of thousands of
  // short-lived blocks is limited by how fast blocks are dispatched and
  // retired, not by the work in them.
  int grid_cap = 8192;
  // Push, repair and cascade launches (random-access discharge loops whose
  // waves wait on gathers and atomics) take a few more blocks per CU.
  int push_cap = 8192;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  dmf_counters last{};
  int64_t bytes = 0;
  uint64_t launches = 0;
  // learned scheduler
  SchedWeights* d_sched_w = nullptr;
  SchedState* d_sched = nullptr;
  SchedState* h_sched = nullptr;  // pinned mirror
  int last_levels = 0;
  int last_variant = -1;          // variant of the last solve (-1 = none yet)
  // The push-pull batch continues from the state of a push-pull static solve
  // (sink saturated, cut in side[]); after a data or topology static solve it
  // would end short of the maximum (the reference's pp always starts from its
  // own static solve).  Data and topology batches keep the state usable.
  bool pp_ready = false;
  // Labels the last data/topology solve converged on (a fresh relabel from
  // the sinks; order[0, visited_last) its labelled vertices).  A batch may
  // start on them instead of relabelling first: its updates can only make
  // them stale, and convergence is accepted on fresh labels only.
  bool labels_fresh = false;
  // Early-exit relabel (BFS from the sinks stopped once every active vertex is
  // labelled): 0 never, 1 with cascade sub-rounds (default), 2 under every
  // data-driven schedule.  DMF_EARLY_RELABEL overrides.
  int early_mode = 1;
  // Cascade push sub-rounds after each data-driven round's push and repair:
  // DMF_CASCADE overrides the solve's cap when set (>= 0); a sub-round gives
  // each listed vertex cascade_kc discharge cycles and lists only the
  // vertices it gave excess (cascade_relist 0; DMF_CASCADE_KC and
  // DMF_CASCADE_RELIST override both).  Stamps of the lists only grow.
  int cascade_env = -1;
  int cascade_kc = 1;
  int cascade_relist = 0;
  int32_t cstamp = 0;
  int32_t sched_tag = 0;  // tag of the last round whose decision a cascade launch takes
  // Two buffers for the BFS visit order: a restricted relabel tries as
  // bottom-up candidates only the labelled set of the solve's last complete
  // relabel from t (ord[full_buf], full_n vertices) and writes the other
  // buffer.  full_buf < 0: no such set yet in this solve.
  int32_t* ord[2] = {nullptr, nullptr};
  int full_buf = -1;
  int full_n = 0;
  int restrict_mode = 4;  // DMF_BFS_RESTRICT: 0 off, else Dev::bu_rest
  bool restricted = false;  // the last relabel ran a restricted bottom-up level
  int visited_last = 0;
  FILE* trace = nullptr;          // per-round scheduler trace (DMF_TRACE)
  int32_t ninst = 1;              // independent instances stored as one disjoint union
  std::vector<int64_t> voff;      // vertex offset of each instance (batched engines)
  std::vector<int32_t> h_row;     // host copy of the CSR offsets (wide-list rebuilds)
  // Vertex scattering: h_vmap[v] is the engine's id of caller vertex v and
  // h_row_caller the caller's offsets (arcs keep their order within a vertex).
  // Both empty, and d_vmap null, when the engine keeps the caller's numbering.
  std::vector<int32_t> h_vmap, h_row_caller;
  int32_t* d_vmap = nullptr;
  int64_t wide_bytes = 0;
  uint64_t solve_id = 0;
  std::chrono::steady_clock::time_point t_solve;
  // graph-replayed BFS (mode 2): segments of 2^i level pairs per BFS mode
  static constexpr int kSegSizes = 7;
  Dev* d_dev = nullptr;
  hipGraph_t seg_graph[kBfsModes][kSegSizes] = {};
  hipGraphExec_t seg_exec[kBfsModes][kSegSizes] = {};
  bool use_graph = false;
  bool graph_failed = false;
  int seg_levels[kBfsModes] = {};
  // the same for early-exit relabels from t, which stop shallower than full ones
  int seg_levels_early = 0;
  // round graphs (mode 3): a whole data-driven round with 2*i level launches
  // in slot i, captured for push/relabel cycles kc (a kernel argument fixed
  // at capture); round_depth is the deepest relabel of the last such solve
  static constexpr int kRoundSlots = 33;  // up to 64 level launches
  hipGraph_t round_graph[kRoundSlots] = {};
  hipGraphExec_t round_exec[kRoundSlots] = {};
  int round_kc[kRoundSlots] = {};
  int round_depth = 0;
  bool use_rounds = false;
  // A round queued behind a relabel's first level segment (data-driven loops
  // with cascade sub-rounds): round_behind launches it; relabel_graph records
  // from its first host check whether it ran (spec_ran: the BFS was over and
  // the list not empty), the list sizes it used and how much of the cascade
  // chain it queued.
  const std::function<int()>* round_behind = nullptr;
  bool spec_launched = false;
  bool spec_ran = false;
  int32_t spec_cnt = 0;
  unsigned long long spec_wpk = 0;
  int spec_sub = 0;
  SchedRound spec_sr{};
  // per-launch BFS level timing (dmf_engine_set_profiling)
  bool profile = false;
  hipEvent_t pev0 = nullptr, pev1 = nullptr;
  // push-pull rounds: the A side's pull and repair run on a second stream
  // beside the B side's push and repair (created on first use)
  hipStream_t pull_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

namespace {

#define HIP_OK(expr)                                                          \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      std::fprintf(stderr, "dmf: HIP error %s at %s:%d: %s\n", hipGetErrorName(_e), __FILE__, \
                   __LINE__, #expr);                                          \
      return DMF_EDEVICE;                                                     \
    }                                                                         \
  } while (0)

template <class T>
int dalloc(dmf_engine* E, T** p, size_t count) {
  if (count == 0) count = 1;
  if (hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * count) != hipSuccess) {
    *p = nullptr;
    return DMF_ENOMEM;
  }
  E->bytes += static_cast<int64_t>(sizeof(T) * count);
  return DMF_OK;
}

constexpr int kGridBlocksPerCu = 4;
// Push/repair/cascade grids: 6 blocks per CU measured 2-3% faster per learned
// RMAT-20 batch than 4 (8 the same as 6); the BFS levels stay at 4.
constexpr int kPushBlocksPerCu = 6;
// Cascade sub-rounds launched between host checks, and the learned
// scheduler's cap on them per round when the parameters give none.
constexpr int kCascadeChunk = 8;
constexpr int kLearnedCascade = 1024;

int blocks_for(int64_t work) {
  int64_t b = (work + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return static_cast<int>(b);
}

int sync_ctrl(dmf_engine* E) {
  HIP_OK(hipMemcpyAsync(E->h_ctrl, E->D.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, E->stream));
  HIP_OK(hipStreamSynchronize(E->stream));
  return DMF_OK;
}

#define LAUNCH_ON(st, kernel, blocks, ...)                                            \
  do {                                                                                \
    hipLaunchKernelGGL(kernel, dim3(std::min<int>((blocks), E->grid_cap)), dim3(kBlock), 0, (st), \
                       __VA_ARGS__);                                                  \
    ++E->launches;                                                                    \
    HIP_OK(hipGetLastError());                                                        \
  } while (0)
#define LAUNCH(kernel, blocks, ...) LAUNCH_ON(E->stream, kernel, blocks, __VA_ARGS__)
// push, repair and cascade launches: capped at push_cap instead of grid_cap
#define LAUNCH_PUSH_ON(st, kernel, blocks, ...)                                       \
  do {                                                                                \
    hipLaunchKernelGGL(kernel, dim3(std::min<int>((blocks), E->push_cap)), dim3(kBlock), 0, (st), \
                       __VA_ARGS__);                                                  \
    ++E->launches;                                                                    \
    HIP_OK(hipGetLastError());                                                        \
  } while (0)
#define LAUNCH_PUSH(kernel, blocks, ...) LAUNCH_PUSH_ON(E->stream, kernel, blocks, __VA_ARGS__)

int zero_dev(dmf_engine* E, void* p, size_t bytes) {
  HIP_OK(hipMemsetAsync(p, 0, bytes, E->stream));
  return DMF_OK;
}

// Grid for a phase over `cnt` list entries (64 per wave) plus `chunks` wide chunks.
int phase_blocks(int64_t cnt, int64_t chunks) {
  const int64_t waves = (cnt + kWave - 1) / kWave + chunks;
  return blocks_for(waves * kWave);
}

int64_t chunks_of(unsigned long long packed) { return static_cast<int64_t>(packed & 0xffffffffu); }

// BFS levels: a wave expands kGroup frontier vertices at a time, so one wave
// per kGroup vertices (plus one per wide chunk) puts every group in flight.
int bfs_blocks(int64_t cnt, int64_t chunks) {
  const int64_t waves = (cnt + kGroup - 1) / kGroup + chunks;
  return blocks_for(waves * kWave);
}

// Stepped global relabel: host loop over levels.  Returns visited count.
int relabel_stepped(dmf_engine* E, int mode, int part_cnt, int* visited, bool early) {
  Dev& D = E->D;
  int rc;
  hipLaunchKernelGGL(k_relabel_prep, dim3(1), dim3(kWave), 0, E->stream, D, nullptr, early ? 1 : 0);
  ++E->launches;
  HIP_OK(hipGetLastError());
  if (mode == BFS_FINAL) LAUNCH(k_bfs_init_final, blocks_for(part_cnt), D, part_cnt);
  else LAUNCH(k_bfs_init, blocks_for(E->n), D, mode, early ? 1 : 0);
  int lo = 0, lvl = 0;
  for (;;) {
    if ((rc = sync_ctrl(E))) return rc;
    if (E->profile && lvl > 0) {  // the previous level's launch has completed
      float ms = 0.f;
      HIP_OK(hipEventElapsedTime(&ms, E->pev0, E->pev1));
      E->last.bfs_level_ms += ms;
      E->last.bfs_level_launches += 1;
    }
    const int32_t cnt = E->h_ctrl->qcnt[lvl % 3];
    const unsigned long long wpk = E->h_ctrl->wq[lvl % 3];
    if (cnt == 0) break;
    const int bu = use_bottom_up(D, mode, lo, cnt, chunks_of(wpk)) ? 1 : 0;
    const int grid = std::min(bu ? blocks_for(E->n) : bfs_blocks(cnt, chunks_of(wpk)), E->grid_cap);
    if (E->profile) {
      // events stamped by the launch itself: the span is the kernel's execution
      hipExtLaunchKernelGGL(k_bfs_level, dim3(grid), dim3(kBlock), 0, E->stream, E->pev0, E->pev1, 0, D, mode, lo,
                            cnt, lvl, wpk, bu);
      ++E->launches;
      HIP_OK(hipGetLastError());
    } else {
      LAUNCH(k_bfs_level, grid, D, mode, lo, cnt, lvl, wpk, bu);
    }
    lo += cnt;
    ++lvl;
    if (lvl > E->n + 1) return DMF_EDEVICE;
  }
  E->last.bfs_levels += lvl;
  E->last.global_relabels += 1;
  E->last_levels = lvl;
  *visited = lo;
  return DMF_OK;
}

constexpr int kNoGraph = 1;  // internal: capture refused, use per-level launches

// Captures a segment of 2^idx level launches (idx >= 1, so every segment
// starts and ends on cursor slot 0) for BFS mode `mode`.
int capture_segment(dmf_engine* E, int mode, int idx) {
  if (E->seg_exec[mode][idx]) return DMF_OK;
  // fixed at capture: at most 8 blocks per CU (a full grid makes every empty
  // trailing level of a segment expensive)
  const int grid = std::min(phase_blocks(E->n, chunks_of(E->D.wide_packed)), E->grid_cap);
  hipStream_t st = E->stream;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess) return kNoGraph;
  bool ok = true;
  for (int i = 0; i < (1 << idx) && ok; ++i) {
    hipLaunchKernelGGL(k_bfs_level_dev, dim3(grid), dim3(kBlock), 0, st, E->d_dev, E->D.ctrl, mode, i & 1);
    ok = hipGetLastError() == hipSuccess;
  }
  hipGraph_t g = nullptr;
  const bool ended = hipStreamEndCapture(st, &g) == hipSuccess;
  if (!ok || !ended || !g) {
    if (g) (void)hipGraphDestroy(g);
    return kNoGraph;
  }
  hipGraphExec_t x = nullptr;
  if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    return kNoGraph;
  }
  E->seg_graph[mode][idx] = g;
  E->seg_exec[mode][idx] = x;
  return DMF_OK;
}

void destroy_segments(dmf_engine* E) {
  for (int a = 0; a < kBfsModes; ++a)
    for (int b = 0; b < dmf_engine::kSegSizes; ++b) {
      if (E->seg_exec[a][b]) (void)hipGraphExecDestroy(E->seg_exec[a][b]);
      if (E->seg_graph[a][b]) (void)hipGraphDestroy(E->seg_graph[a][b]);
      E->seg_exec[a][b] = nullptr;
      E->seg_graph[a][b] = nullptr;
    }
  for (int b = 0; b < dmf_engine::kRoundSlots; ++b) {
    if (E->round_exec[b]) (void)hipGraphExecDestroy(E->round_exec[b]);
    if (E->round_graph[b]) (void)hipGraphDestroy(E->round_graph[b]);
    E->round_exec[b] = nullptr;
    E->round_graph[b] = nullptr;
  }
}

// Kernels in one captured round of slot idx: begin, roots, 2*idx levels, mid,
// worklist, list, push, repair.
int round_launches(int idx) { return 7 + 2 * idx; }
// Slot whose round holds at least depth + 1 level launches (a relabel of
// depth d takes d + 1: the last finds its frontier empty), capped at 64.
int round_slot(int depth) { return std::min(dmf_engine::kRoundSlots - 1, std::max(1, (depth + 2) / 2)); }

// Captures one data-driven round of the fixed schedule with 2*idx BFS level
// launches (an even count, so the levels end on cursor slot 0, which
// k_round_mid reads).  Grids are fixed at capture: every kernel strides over
// whatever its list holds at run time.  The captured kernels take the engine
// state by value, so the graphs are rebuilt whenever that state is
// (setup_wide).
int capture_round(dmf_engine* E, int idx, int kc) {
  if (E->round_exec[idx] && E->round_kc[idx] == kc) return DMF_OK;
  if (E->round_exec[idx]) {
    (void)hipGraphExecDestroy(E->round_exec[idx]);
    (void)hipGraphDestroy(E->round_graph[idx]);
    E->round_exec[idx] = nullptr;
    E->round_graph[idx] = nullptr;
  }
  const Dev& D = E->D;
  const int gn = std::min(blocks_for(E->n), E->grid_cap);
  const int glev = std::min(phase_blocks(E->n, chunks_of(D.wide_packed)), E->grid_cap);
  const int glist = std::min(phase_blocks(E->n, chunks_of(D.wave_packed)), E->push_cap);  // push and repair
  hipStream_t st = E->stream;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess) return kNoGraph;
  hipLaunchKernelGGL(k_round_begin, dim3(1), dim3(kWave), 0, st, D);
  hipLaunchKernelGGL(k_round_bfs_init, dim3(gn), dim3(kBlock), 0, st, D);
  for (int i = 0; i < 2 * idx; ++i)
    hipLaunchKernelGGL(k_bfs_level_dev, dim3(glev), dim3(kBlock), 0, st, E->d_dev, D.ctrl, static_cast<int>(BFS_SINK),
                       i & 1);
  hipLaunchKernelGGL(k_round_mid, dim3(1), dim3(kWave), 0, st, D);
  hipLaunchKernelGGL(k_round_wl, dim3(gn), dim3(kBlock), 0, st, D);
  hipLaunchKernelGGL(k_round_list, dim3(1), dim3(kWave), 0, st, D);
  hipLaunchKernelGGL(k_round_push, dim3(glist), dim3(kBlock), 0, st, D, kc);
  hipLaunchKernelGGL(k_round_repair, dim3(glist), dim3(kBlock), 0, st, D);
  const bool ok = hipGetLastError() == hipSuccess;
  hipGraph_t g = nullptr;
  const bool ended = hipStreamEndCapture(st, &g) == hipSuccess;
  if (!ok || !ended || !g) {
    if (g) (void)hipGraphDestroy(g);
    return kNoGraph;
  }
  hipGraphExec_t x = nullptr;
  if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    return kNoGraph;
  }
  E->round_graph[idx] = g;
  E->round_exec[idx] = x;
  E->round_kc[idx] = kc;
  return DMF_OK;
}

// Global relabel with the level loop replayed from captured graphs: one host
// check per segment instead of one per level.  The first segment is as long as
// the previous BFS of the same kind was deep; later ones grow.  Same result
// as relabel_stepped (the same level kernel body runs on the same frontiers).
// list (BFS_SINK): the worklist of the round that follows is built behind the
// first segment (k_cap_after for an early-exit relabel, k_build_wl_after) and
// taken from the same host check; a deeper BFS runs on and the two are
// launched again once it is over.
int relabel_graph(dmf_engine* E, int mode, int part_cnt, int* visited, bool early, bool list) {
  Dev& D = E->D;
  int rc;
  const int gn = std::min(blocks_for(E->n), E->grid_cap);
  auto list_behind = [&]() -> int {
    if (early) hipLaunchKernelGGL(k_cap_after, dim3(gn), dim3(kBlock), 0, E->stream, D);
    hipLaunchKernelGGL(k_build_wl_after, dim3(gn), dim3(kBlock), 0, E->stream, D);
    E->launches += early ? 2 : 1;
    HIP_OK(hipGetLastError());
    return DMF_OK;
  };
  bool first = true;
  E->spec_launched = E->spec_ran = false;
  hipLaunchKernelGGL(k_relabel_prep, dim3(1), dim3(kWave), 0, E->stream, D, E->d_dev, early ? 1 : 0);
  ++E->launches;
  HIP_OK(hipGetLastError());
  if (mode == BFS_FINAL) LAUNCH(k_bfs_init_final, blocks_for(part_cnt), D, part_cnt);
  else LAUNCH(k_bfs_init, blocks_for(E->n), D, mode, early ? 1 : 0);
  int idx = 1;
  // the first segment as long as the last relabel of the same kind was deep
  const int predicted = early ? E->seg_levels_early : E->seg_levels[mode];
  while (idx + 1 < dmf_engine::kSegSizes && (1 << idx) < predicted + 1) ++idx;
  int lvl = 0;
  for (;;) {
    if ((rc = capture_segment(E, mode, idx))) return rc;
    if (E->profile) HIP_OK(hipEventRecord(E->pev0, E->stream));
    HIP_OK(hipGraphLaunch(E->seg_exec[mode][idx], E->stream));
    if (E->profile) HIP_OK(hipEventRecord(E->pev1, E->stream));
    ++E->launches;
    if (list && first && (rc = list_behind())) return rc;
    const bool spec = list && first && E->round_behind != nullptr;
    if (spec && (rc = (*E->round_behind)())) return rc;
    const bool first_check = first;
    first = false;
    if ((rc = sync_ctrl(E))) return rc;
    if (first_check && spec) {
      const Ctrl& h = *E->h_ctrl;
      E->spec_launched = true;
      E->spec_ran = !h.bfs_inc && !h.r_skip;
      E->spec_cnt = h.r_cnt;
      E->spec_wpk = h.r_wpk;
    }
    if (E->profile) {
      float ms = 0.f;
      HIP_OK(hipEventElapsedTime(&ms, E->pev0, E->pev1));
      E->last.bfs_level_ms += ms;
      E->last.bfs_level_launches += static_cast<uint64_t>(1) << idx;
    }
    lvl = E->h_ctrl->bcur[0][1];
    // behind the first segment, k_build_wl_after recorded whether the BFS
    // was over (a round queued there reuses the ring counters once it was)
    if (first_check && list ? !E->h_ctrl->bfs_inc : E->h_ctrl->qcnt[lvl % 3] == 0) break;
    if (lvl > E->n + 1) return DMF_EDEVICE;
    if (idx + 1 < dmf_engine::kSegSizes) ++idx;
  }
  if (list && E->h_ctrl->bfs_inc) {
    // the first segment did not finish the BFS: list now
    if ((rc = list_behind())) return rc;
    if ((rc = sync_ctrl(E))) return rc;
  }
  (early ? E->seg_levels_early : E->seg_levels[mode]) = lvl;
  E->last.bfs_levels += lvl;
  E->last.global_relabels += 1;
  E->last_levels = lvl;
  *visited = E->h_ctrl->bcur[0][0];
  return DMF_OK;
}

// early (BFS_SINK only): stop once every vertex active now is labelled; the
// control block's bfs_cut then tells that the labels are not complete.
// *listed (may be null): the round's worklist was built behind the relabel
// (graph-replayed levels from t only); otherwise the caller builds it.
int relabel_run(dmf_engine* E, int mode, int part_cnt, int* visited, bool early, bool* listed) {
  const bool list = listed != nullptr && mode == BFS_SINK;
  if (listed) *listed = false;
  if (E->use_graph && !E->graph_failed) {
    const int rc = relabel_graph(E, mode, part_cnt, visited, early, list);
    if (rc == DMF_OK && list) *listed = true;
    if (rc != kNoGraph) return rc;
    // capture or instantiation refused: keep the host-driven level loop for
    // this engine; relabel_stepped starts the BFS over from its roots
    E->graph_failed = true;
    std::fprintf(stderr, "dmf: BFS graph capture unavailable, using per-level launches\n");
  }
  return relabel_stepped(E, mode, part_cnt, visited, early);
}

// Global relabel.  restrict_prev (from t only): bottom-up levels try only the
// labelled set of the solve's last complete relabel from t.  A vertex that
// became able to reach t since then is missed by such a level and keeps
// height n, so a relabel that ran one is not taken as fresh (E->restricted;
// its caller confirms convergence with an unrestricted relabel).  Relabels
// from t write the order buffer that does not hold that set, and a complete
// one (not cut short by an early exit) becomes the set for the next.
int relabel(dmf_engine* E, int mode, int part_cnt, int* visited, bool early = false, bool restrict_prev = false,
            bool* listed = nullptr) {
  early = early && mode == BFS_SINK;
  Dev& D = E->D;
  const bool two = mode == BFS_SINK && E->ord[1] != nullptr;
  const bool rs = two && restrict_prev && E->full_buf >= 0 && E->full_n > 0 && E->restrict_mode > 0;
  if (two && E->full_buf >= 0) D.order = E->ord[E->full_buf ^ 1];
  if (rs) {
    D.prev = E->ord[E->full_buf];
    D.nprev = E->full_n;
    D.bu_rest = E->restrict_mode;
  }
  const int rc = relabel_run(E, mode, part_cnt, visited, early, listed);
  D.prev = nullptr;
  D.nprev = 0;
  // labels of a restricted relabel whose levels all ran top-down are exact
  E->restricted = rs && E->h_ctrl->bfs_rbu != 0;
  if (rc) return rc;
  if (two && !(early && E->h_ctrl->bfs_cut)) {
    E->full_buf = D.order == E->ord[1] ? 1 : 0;
    E->full_n = *visited;
  }
  return DMF_OK;
}

int reset_lists(dmf_engine* E) {
  Dev& D = E->D;
  int rc;
  if ((rc = zero_dev(E, &D.ctrl->wl_cnt, 3 * sizeof(int32_t)))) return rc;  // wl, wl2, part
  if ((rc = zero_dev(E, &D.ctrl->wq[kQActive], 2 * sizeof(unsigned long long)))) return rc;
  return DMF_OK;
}

double since_solve_ms(const dmf_engine* E) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - E->t_solve).count();
}

// One line of the scheduler trace (DMF_TRACE) for the decision in h_sched.
void sched_trace(dmf_engine* E, const char* what, int round, int variant, int relabelled) {
  if (!E->trace) return;
  const SchedState& S = *E->h_sched;
  std::fprintf(E->trace,
               "{\"solve\":%llu,\"what\":\"%s\",\"round\":%d,\"variant\":%d,\"relabelled\":%d,"
               "\"k\":%d,\"t_ms\":%.4f,\"feat\":[%.6g,%.6g,%.6g,%.6g,%.6g,%.6g,%.6g,%.6g],"
               "\"out\":[%.6g,%.6g,%.6g,%.6g]}\n",
               (unsigned long long)E->solve_id, what, round, variant, relabelled, E->D.k, since_solve_ms(E),
               S.feat[0], S.feat[1], S.feat[2], S.feat[3], S.feat[4], S.feat[5], S.feat[6], S.feat[7],
               S.out[0], S.out[1], S.out[2], S.out[3]);
}

// Fraction of the batch's updates that raise a capacity (scheduler feature).
float inc_fraction(const dmf_engine* E) {
  const int k = E->D.k;
  return k > 0 ? static_cast<float>(E->h_ctrl->inc) / static_cast<float>(k) : 0.5f;
}

// Queues the scheduler MLP on the device and the copy of its decision to the
// host (inc < 0: the kernel reads the batch's increase count itself).
int launch_sched(dmf_engine* E, int phase, int rounds_since, double active, float inc) {
  Dev& D = E->D;
  hipLaunchKernelGGL(k_sched, dim3(1), dim3(kWave), 0, E->stream, D, E->d_sched_w, E->d_sched, phase, D.k, inc,
                     rounds_since, E->last_levels, active);
  ++E->launches;
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(E->h_sched, E->d_sched, sizeof(SchedState), hipMemcpyDeviceToHost, E->stream));
  return DMF_OK;
}

// Evaluates the scheduler MLP on the device and reads its decision back.
int run_sched(dmf_engine* E, int phase, int rounds_since, const char* what, int round, int variant,
              int relabelled, double active) {
  int rc;
  if ((rc = launch_sched(E, phase, rounds_since, active, inc_fraction(E)))) return rc;
  HIP_OK(hipStreamSynchronize(E->stream));
  sched_trace(E, what, round, variant, relabelled);
  return DMF_OK;
}

// The round-graph driver's data-driven loop (fixed schedule, no cascade
// sub-rounds, early-exit relabels or lazy start): one replay per round and
// one host check after it.  A relabel deeper than the captured level launches
// stops the replay (r_stop); the host then runs further level segments from
// the device cursor, finishes the round with the same kernels launched one by
// one, and captures a longer round for the rounds after it.  The order buffer
// is pinned to the first of the two (this driver never restricts bottom-up
// levels to a previous labelled set).  Same phases, lists and conditions as
// the host-checked loop of solve_stepped, so the same rounds and results.
int rounds_loop(dmf_engine* E, const Cfg& C, int* round, int* visited, bool* conv) {
  Dev& D = E->D;
  int rc;
  if (E->ord[0]) D.order = E->ord[0];
  D.prev = nullptr;
  D.nprev = 0;
  E->full_buf = -1;
  hipLaunchKernelGGL(k_round_init, dim3(1), dim3(kWave), 0, E->stream, D, E->d_dev, C.interval, C.max_rounds);
  ++E->launches;
  HIP_OK(hipGetLastError());
  // rounds as long as the deepest relabel of the last solve of this driver
  int idx = round_slot(E->round_depth);
  const int gn = std::min(blocks_for(E->n), E->grid_cap);
  const int glist = std::min(phase_blocks(E->n, chunks_of(D.wave_packed)), E->push_cap);  // push and repair
  int deepest = 0;
  for (;;) {
    if ((rc = capture_round(E, idx, C.kc))) return rc;
    HIP_OK(hipGraphLaunch(E->round_exec[idx], E->stream));
    E->launches += static_cast<uint64_t>(round_launches(idx));
    if ((rc = sync_ctrl(E))) return rc;
    if (E->h_ctrl->r_stop) {
      // the relabel goes on past the captured levels: segments of level
      // launches from the cursor until the frontier is empty
      int seg = 1;
      while (seg + 1 < dmf_engine::kSegSizes && (1 << seg) < 2 * idx) ++seg;
      for (;;) {
        if ((rc = capture_segment(E, BFS_SINK, seg))) return rc;
        HIP_OK(hipGraphLaunch(E->seg_exec[BFS_SINK][seg], E->stream));
        E->launches += static_cast<uint64_t>(1) << seg;
        if ((rc = sync_ctrl(E))) return rc;
        const int lvl = E->h_ctrl->bcur[0][1];
        if (E->h_ctrl->qcnt[lvl % 3] == 0) {
          deepest = std::max(deepest, lvl);
          break;
        }
        if (lvl > E->n + 1) return DMF_EDEVICE;
        if (seg + 1 < dmf_engine::kSegSizes) ++seg;
      }
      hipLaunchKernelGGL(k_round_resume, dim3(1), dim3(kWave), 0, E->stream, D);
      hipLaunchKernelGGL(k_round_mid, dim3(1), dim3(kWave), 0, E->stream, D);
      hipLaunchKernelGGL(k_round_wl, dim3(gn), dim3(kBlock), 0, E->stream, D);
      hipLaunchKernelGGL(k_round_list, dim3(1), dim3(kWave), 0, E->stream, D);
      hipLaunchKernelGGL(k_round_push, dim3(glist), dim3(kBlock), 0, E->stream, D, C.kc);
      hipLaunchKernelGGL(k_round_repair, dim3(glist), dim3(kBlock), 0, E->stream, D);
      E->launches += 6;
      HIP_OK(hipGetLastError());
      idx = std::max(idx, round_slot(deepest));
      continue;
    }
    if (E->h_ctrl->r_done) break;
  }
  const Ctrl& c = *E->h_ctrl;
  *round = c.r_rounds;
  *visited = c.visited;
  *conv = c.r_noconv == 0;
  E->last.bfs_levels += static_cast<uint64_t>(c.r_levels);
  E->last.global_relabels += static_cast<uint64_t>(c.r_relabels);
  // depth of the last relabel: the converging one
  E->last_levels = c.bcur[0][1];
  E->seg_levels[BFS_SINK] = std::max(c.bcur[0][1], 1);
  E->round_depth = std::max(c.r_maxlvl, deepest);
  return DMF_OK;
}

// The second stream of push-pull rounds and its fork/join events, created on
// first use (default priority, like the engine's own stream).
int ensure_pull_stream(dmf_engine* E) {
  if (E->pull_stream) return DMF_OK;
  HIP_OK(hipStreamCreateWithFlags(&E->pull_stream, hipStreamNonBlocking));
  HIP_OK(hipEventCreateWithFlags(&E->ev_fork, hipEventDisableTiming));
  HIP_OK(hipEventCreateWithFlags(&E->ev_join, hipEventDisableTiming));
  return DMF_OK;
}

// Stepped solve (one kernel per phase, host-checked conditions).
int solve_stepped(dmf_engine* E, const Cfg& C0) {
  Dev& D = E->D;
  int rc;
  const bool labels_in = E->labels_fresh;
  E->labels_fresh = false;  // set again only by a converged data-driven loop
  E->full_buf = -1;         // the update may have let more vertices reach t
  Cfg C = C0;
  const bool learned = C.scheduler == DMF_SCHED_LEARNED;
  const bool consult = learned || E->trace != nullptr;
  const int n = E->n;
  const int64_t static_chunks = chunks_of(D.wave_packed);  // topology-driven push and repair
  if (C.entry == 0) {
    LAUNCH(k_init_residual, blocks_for(std::max(E->m, E->n)), D);
    LAUNCH(k_saturate_source, blocks_for(std::min<int64_t>(E->m, 1 << 20)), D);
    if (C.variant == DMF_VARIANT_PP) LAUNCH(k_saturate_sink_pp, blocks_for(std::min<int64_t>(E->m, 1 << 20)), D);
  } else if (C.entry == 1) {
    LAUNCH(k_upd_lookup, blocks_for(D.k), D);
    // batch start: the scheduler picks the variant (push-pull only from a
    // push-pull static state); its decision comes back with the lookup's
    // counters in one host check
    if (consult && (rc = launch_sched(E, 0, 0, 0.0, -1.f))) return rc;
    if ((rc = sync_ctrl(E))) return rc;
    if (E->h_ctrl->missing || E->h_ctrl->invalid) {  // caller reports; owners cleared
      LAUNCH(k_upd_release, blocks_for(D.k), D);
      return DMF_OK;
    }
    if (consult) {
      sched_trace(E, "batch", 0, C.variant, 0);
      if (learned && E->h_sched->valid && (E->h_sched->variant != DMF_VARIANT_PP || E->pp_ready))
        C.variant = E->h_sched->variant;
    }
    if (C.variant == DMF_VARIANT_PP && !E->pp_ready) C.variant = DMF_VARIANT_DATA;
  }
  const bool pp = C.variant == DMF_VARIANT_PP;
  const bool topo = C.variant == DMF_VARIANT_TOPO;
  if (C.entry == 1) {
    // the apply releases the arcs it updates (no release pass)
    LAUNCH(k_upd_apply, blocks_for(std::max<int64_t>(D.k, D.wide_packed >> 32)), D, 0, pp ? 1 : 0);
    if (pp) {
      LAUNCH(k_pp_cut, phase_blocks(n, chunks_of(D.wide_packed)), D);
    } else {
      LAUNCH(k_upd_clamp, blocks_for(D.k), D);
      LAUNCH(k_saturate_source, blocks_for(std::min<int64_t>(E->m, 1 << 20)), D);
    }
  }
  int round = 0;
  bool conv = true;
  double active = 0;  // vertices listed for the last round (scheduler feature)
  if (!pp || C.entry == 0) {
    int since = C.interval, visited = 0;
    bool want_relabel = true;  // scheduler decision for the coming round
    if (C.entry == 1 && labels_in && !topo) {
      // lazy start (data-driven only: its loop accepts convergence on fresh
      // labels alone): keep the previous solve's labels for the first round
      const bool skip = learned ? (E->h_sched->valid && E->h_sched->relabel == 0) : C.lazy != 0;
      if (skip) {
        since = 0;
        visited = E->visited_last;
        want_relabel = false;
      }
    }
    // Early-exit relabels label only as deep as the active vertices lie, so
    // they never count as fresh: the first relabel of a solve and every
    // relabel that confirms convergence are full.
    // Cascade sub-rounds: the fixed schedules only when asked for, the learned
    // scheduler by default; the chain is launched in chunks and stops when a
    // chunk leaves no vertex listed.
    int sub = topo ? 0 : (C.cascade > 0 ? C.cascade : (learned ? kLearnedCascade : 0));
    if (E->cascade_env >= 0) sub = topo ? 0 : E->cascade_env;
    if (!E->D.cmark) sub = 0;
    const bool try_early = !topo && E->D.tgt && (E->early_mode >= 2 || (E->early_mode == 1 && sub > 0));
    bool need_full = true;
    // round graphs: the fixed schedule's data-driven loop with its conditions
    // on the device (other schedules keep the host-checked loop below)
    bool rounds = E->use_rounds && !E->graph_failed && !topo && !consult && sub == 0 && !try_early && want_relabel;
    if (rounds) {
      rc = rounds_loop(E, C, &round, &visited, &conv);
      if (rc == kNoGraph) {
        // capture refused: the host-checked loop carries on from the current
        // state (it starts with a relabel) on cleared lists
        E->graph_failed = true;
        std::fprintf(stderr, "dmf: round graph capture unavailable, using per-phase launches\n");
        if ((rc = reset_lists(E))) return rc;
        rounds = false;
      } else if (rc) {
        return rc;
      }
    }
    const int ckc = E->cascade_kc > 0 ? E->cascade_kc : C.kc;
    // the push grid: the first sub-rounds carry lists as long as the round's
    // own; 2 blocks per CU measured 2-3% slower per learned batch on RMAT-20
    // than 4
    const int cgrid = std::max(1, E->push_cap);
    // the scheduler's round-end decision rides on the first cascade launch
    // that finds the chain's list empty
    const SchedWeights* sw = consult ? E->d_sched_w : nullptr;
    // With cascade sub-rounds, a round that follows a graph-replayed relabel is
    // queued behind the relabel's first level segment and worklist: push and
    // repair size themselves from the list k_list_snap records on the device,
    // and the chain's first kCascadeChunk sub-rounds (with the scheduler's
    // decision) follow, so one host check serves the relabel, the list and
    // the round.  When the BFS outlasts that segment, or the list is empty,
    // the queued kernels do nothing and the round proceeds as below.
    const std::function<int()> round_behind = [&]() -> int {
      int rc2;
      if (E->cstamp > INT_MAX - (sub + 2)) {
        if ((rc2 = zero_dev(E, D.cmark, sizeof(int32_t) * static_cast<size_t>(n)))) return rc2;
        E->cstamp = 0;
      }
      hipLaunchKernelGGL(k_list_snap, dim3(1), dim3(kWave), 0, E->stream, D);
      ++E->launches;
      HIP_OK(hipGetLastError());
      LAUNCH_PUSH(k_push, E->push_cap, D, 1, -1, C.kc, PM_ALL, 0ull, 0, 0, ++E->cstamp, E->cascade_relist);
      LAUNCH_PUSH(k_repair, E->push_cap, D, 1, -1, RM_ALL, 0ull);
      // rounds since the relabel: 1; active and the relabel's depth: on the device
      E->spec_sr = SchedRound{D.k, inc_fraction(E), 1, -1, -1.0, ++E->sched_tag, 1};
      const int chunk = std::min(kCascadeChunk, sub);
      int slot = 0;
      for (int j = 0; j < chunk; ++j) {
        LAUNCH_PUSH(k_cascade, cgrid, D, slot, ckc, ++E->cstamp, E->cascade_relist, sw, E->d_sched, E->spec_sr);
        slot = (slot + 1) % 3;
      }
      E->spec_sub = chunk;
      if (consult)
        HIP_OK(hipMemcpyAsync(E->h_sched, E->d_sched, sizeof(SchedState), hipMemcpyDeviceToHost, E->stream));
      return DMF_OK;
    };
    for (; !rounds;) {
      bool fresh = false;
      bool listed = false;  // the worklist was built behind this round's relabel
      bool spec_ran = false;  // ... and the round itself, with its first cascade chunk
      const bool due = learned ? (want_relabel || since >= 16) : since >= C.interval;
      if (due) {
        // the relabels that may stop early may also restrict their bottom-up
        // levels to the last complete labelled set
        const bool early = try_early && !need_full;
        // data-driven rounds build their worklist behind the relabel (one
        // host check for both) where the levels are graph-replayed
        E->round_behind = (!topo && sub > 0 && round < C.max_rounds) ? &round_behind : nullptr;
        rc = relabel(E, BFS_SINK, 0, &visited, early, early, topo ? nullptr : &listed);
        E->round_behind = nullptr;
        if (rc) return rc;
        spec_ran = listed && E->spec_launched && E->spec_ran;
        since = 0;
        const bool cut = early && E->h_ctrl->bfs_cut;
        fresh = !cut && !E->restricted;
        if (cut && !listed) LAUNCH(k_cap_unlabelled, blocks_for(n), D, std::max(0, E->last_levels - 1));
        need_full = false;
      }
      int32_t cnt = 0;
      unsigned long long wpk = 0;
      bool sched_done = false;  // the cascade chain took the round-end decision
      if (!topo) {
        // the list counters were cleared by the last push launch (or k_begin)
        // a labelled set that is a large part of the graph is listed by one
        // coalesced pass over every vertex instead of a gather through order[]
        if (!listed) {
          const bool all = visited > n / 4;
          LAUNCH(k_build_wl, blocks_for(all ? n : visited), D, all ? -1 : visited);
          if ((rc = sync_ctrl(E))) return rc;
        }
        // a round queued behind the relabel took the list sizes on the
        // device and cleared the counters
        cnt = spec_ran ? E->spec_cnt : E->h_ctrl->wl_cnt;
        wpk = spec_ran ? E->spec_wpk : E->h_ctrl->wq[kQActive];
        if (cnt == 0 && wpk == 0) {
          if (fresh) break;
          since = C.interval;  // confirm convergence on fresh labels
          want_relabel = true;
          need_full = true;
          continue;
        }
      }
      if (round >= C.max_rounds) {
        conv = false;
        break;
      }
      if (topo) {
        const int fs = round & 1;
        // every vertex up to wave_lane one per lane, the static wave list for
        // the rest
        Dev Dt = D;
        Dt.lane_max = D.wave_lane;
        LAUNCH_PUSH(k_push, phase_blocks(n, static_chunks), Dt, 0, n, C.kc, PM_ALL, 0ull, fs);
        LAUNCH_PUSH(k_repair, phase_blocks(n, static_chunks), Dt, 0, n, RM_ALL, 0ull);
        if ((rc = sync_ctrl(E))) return rc;
        if (!E->h_ctrl->flag[fs]) break;
        active = static_cast<double>(n);
      } else {
        active = static_cast<double>(cnt) + static_cast<double>(wpk >> 32);
        int slot = 0, done_sub = 0;
        SchedRound sr{};
        bool chain_over = sub == 0;
        if (spec_ran) {
          // push, repair and the chain's first sub-rounds ran behind the
          // relabel; its host check also read the chain's state
          sr = E->spec_sr;
          done_sub = E->spec_sub;
          slot = done_sub % 3;
          if (done_sub >= sub) {
            chain_over = true;
          } else if (E->h_ctrl->qcnt[slot] == 0 && E->h_ctrl->wq[slot] == 0) {
            sched_done = consult && E->h_sched->tag == sr.tag;
            chain_over = true;
          }
        } else {
          if (sub > 0 && E->cstamp > INT_MAX - (sub + 2)) {
            // list stamps only grow: start them again from a cleared mark array
            if ((rc = zero_dev(E, D.cmark, sizeof(int32_t) * static_cast<size_t>(n)))) return rc;
            E->cstamp = 0;
          }
          LAUNCH_PUSH(k_push, phase_blocks(cnt, chunks_of(wpk)), D, 1, cnt, C.kc, PM_ALL, wpk, 0, sub ? 0 : -1,
                      sub ? ++E->cstamp : 0, E->cascade_relist);
          LAUNCH_PUSH(k_repair, phase_blocks(cnt, chunks_of(wpk)), D, 1, cnt, RM_ALL, wpk);
          sr = SchedRound{D.k, inc_fraction(E), since + 1, E->last_levels, active, ++E->sched_tag, 0};
        }
        // cascade sub-rounds: each pushes from the list the pass before it
        // wrote, launched kCascadeChunk at a time with one host check between
        // chunks (an empty list ends the chain)
        if (!chain_over) {
          for (; done_sub < sub;) {
            const int chunk = std::min(kCascadeChunk, sub - done_sub);
            for (int j = 0; j < chunk; ++j) {
              LAUNCH_PUSH(k_cascade, cgrid, D, slot, ckc, ++E->cstamp, E->cascade_relist, sw, E->d_sched, sr);
              slot = (slot + 1) % 3;
            }
            done_sub += chunk;
            if (done_sub >= sub) break;
            if (consult)
              HIP_OK(hipMemcpyAsync(E->h_sched, E->d_sched, sizeof(SchedState), hipMemcpyDeviceToHost, E->stream));
            if ((rc = sync_ctrl(E))) return rc;
            if (E->h_ctrl->qcnt[slot] == 0 && E->h_ctrl->wq[slot] == 0) {
              sched_done = consult && E->h_sched->tag == sr.tag;
              break;
            }
          }
        }
      }
      ++round;
      ++since;
      if (consult) {
        // the trace records whether a relabel (full or early-exit) preceded the round
        if (sched_done) sched_trace(E, "round", round, C.variant, due ? 1 : 0);
        else if ((rc = run_sched(E, 1, since, "round", round, C.variant, due ? 1 : 0, active))) return rc;
        want_relabel = learned ? (!E->h_sched->valid || E->h_sched->relabel != 0) : true;
      }
    }
    // the push-pull static solve leaves its cut in side[] (find_cut,
    // code/dyn_pp_bench.cu:231); other variants leave side[] alone
    if (conv && pp) LAUNCH(k_find_cut, blocks_for(n), D);
    // the data-driven loop ends right after a fresh relabel: its labels and
    // labelled set order[0, visited) are what a lazy-start batch begins from
    if (conv && !pp && !topo) {
      E->labels_fresh = true;
      E->visited_last = visited;
    }
  } else {
    // push-pull batch; stale two-sided labels stay valid between relabels and
    // convergence is only accepted on fresh ones (part_final reads them)
    int since = C.interval, visited = 0;
    bool want_relabel = true;
    for (;;) {
      bool fresh = false;
      const bool due = learned ? (want_relabel || since >= 16) : since >= C.interval;
      if (due) {
        if ((rc = relabel(E, BFS_PP, 0, &visited))) return rc;
        since = 0;
        fresh = true;
      }
      LAUNCH(k_build_pp_wl, blocks_for(visited), D, visited);
      if ((rc = sync_ctrl(E))) return rc;
      const int np = E->h_ctrl->wl_cnt, nq = E->h_ctrl->wl2_cnt;
      const unsigned long long wp = E->h_ctrl->wq[kQActive], wq = E->h_ctrl->wq[kQPull];
      if (np + nq == 0 && wp == 0 && wq == 0) {
        if (fresh) break;
        since = C.interval;
        want_relabel = true;
        continue;
      }
      if (round >= C.max_rounds) {
        conv = false;
        break;
      }
      const bool side_b = np || wp, side_a = nq || wq;
      if (side_a && side_b) {
        if ((rc = ensure_pull_stream(E))) return rc;
        // The reference pulls on one stream and pushes on another
        // (code/dyn_pp_bench.cu:42-60).  The B side's push and repair touch
        // only arcs and excesses between B vertices, the A side's pull and
        // repair only those between A vertices (the cut between them is
        // saturated and side[] does not change within a round), so the two
        // chains run at once: A on the second stream, forked from and joined
        // back into the engine's stream.
        HIP_OK(hipEventRecord(E->ev_fork, E->stream));
        HIP_OK(hipStreamWaitEvent(E->pull_stream, E->ev_fork, 0));
        LAUNCH_PUSH_ON(E->pull_stream, k_push, phase_blocks(nq, chunks_of(wq)), D, 2, nq, C.kc, PM_ONLY_A, wq, 0);
        LAUNCH_PUSH_ON(E->pull_stream, k_repair, phase_blocks(nq, chunks_of(wq)), D, 2, nq, RM_PULL_A, wq);
        HIP_OK(hipEventRecord(E->ev_join, E->pull_stream));
        LAUNCH_PUSH(k_push, phase_blocks(np, chunks_of(wp)), D, 1, np, C.kc, PM_ONLY_B, wp, 0);
        LAUNCH_PUSH(k_repair, phase_blocks(np, chunks_of(wp)), D, 1, np, RM_B, wp);
        HIP_OK(hipStreamWaitEvent(E->stream, E->ev_join, 0));
      } else {
        if (side_b) LAUNCH_PUSH(k_push, phase_blocks(np, chunks_of(wp)), D, 1, np, C.kc, PM_ONLY_B, wp, 0);
        if (side_a) LAUNCH_PUSH(k_push, phase_blocks(nq, chunks_of(wq)), D, 2, nq, C.kc, PM_ONLY_A, wq, 0);
        if (side_b) LAUNCH_PUSH(k_repair, phase_blocks(np, chunks_of(wp)), D, 1, np, RM_B, wp);
        if (side_a) LAUNCH_PUSH(k_repair, phase_blocks(nq, chunks_of(wq)), D, 2, nq, RM_PULL_A, wq);
      }
      active = static_cast<double>(np) + nq + static_cast<double>(wp >> 32) + static_cast<double>(wq >> 32);
      ++round;
      ++since;
      if (consult) {
        if ((rc = run_sched(E, 1, since, "round", round, C.variant, fresh ? 1 : 0, active))) return rc;
        want_relabel = learned ? (!E->h_sched->valid || E->h_sched->relabel != 0) : true;
      }
    }
    if (conv) {
      if ((rc = reset_lists(E))) return rc;
      LAUNCH(k_find_part, blocks_for(n), D);
      if ((rc = sync_ctrl(E))) return rc;
      const int32_t parts = E->h_ctrl->part_cnt;
      if (parts > 0) {
        for (;;) {
          int visited = 0;
          if ((rc = relabel(E, BFS_FINAL, parts, &visited))) return rc;
          LAUNCH(k_build_final_wl, blocks_for(parts), D, parts);
          if ((rc = sync_ctrl(E))) return rc;
          const int32_t cnt = E->h_ctrl->wl_cnt;
          const unsigned long long wpk = E->h_ctrl->wq[kQActive];
          if (cnt == 0 && wpk == 0) break;
          if (round >= C.max_rounds) {
            conv = false;
            break;
          }
          LAUNCH_PUSH(k_push, phase_blocks(cnt, chunks_of(wpk)), D, 1, cnt, C.kc, PM_ALL, wpk, 0);
          LAUNCH_PUSH(k_repair, phase_blocks(cnt, chunks_of(wpk)), D, 1, cnt, RM_PFINAL, wpk);
          ++round;
        }
        if (conv) LAUNCH(k_find_final_cut, blocks_for(n), D);
      }
    }
  }
  hipLaunchKernelGGL(k_flow_prep, dim3(1), dim3(kBlock), 0, E->stream, D);
  ++E->launches;
  HIP_OK(hipGetLastError());
  LAUNCH(k_flow, blocks_for(n), D);
  E->last.rounds = round;
  E->last_variant = C.variant;
  if (!conv) return DMF_ENOCONV;
  return DMF_OK;
}

int solve_persistent(dmf_engine* E, const Cfg& C) {
  Dev D = E->D;
  // topology-driven: every vertex up to wave_lane one per lane, the static
  // wave list for the rest (the persistent kernel's topology rounds use no
  // queues)
  if (C.variant == DMF_VARIANT_TOPO) D.lane_max = D.wave_lane;
  Cfg c = C;
  E->last_variant = C.variant;
  void* args[] = {&D, &c};
  hipError_t err = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_solve_persistent),
                                              dim3(E->coop_blocks), dim3(kBlock), args, 0,
                                              E->stream);
  ++E->launches;
  if (err != hipSuccess) {
    std::fprintf(stderr, "dmf: cooperative launch failed: %s\n", hipGetErrorString(err));
    return DMF_EDEVICE;
  }
  return DMF_OK;
}

Cfg make_cfg(const dmf_engine* E, const dmf_params* p, int entry) {
  Cfg c{};
  dmf_params d;
  if (!p) {
    d.variant = DMF_VARIANT_DATA;
    d.kc = std::max(1, E->m / std::max(1, E->n));
    d.relabel_interval = 1;
    d.scheduler = DMF_SCHED_FIXED;
    d.max_rounds = 0;
    d.mode = 0;
    d.lazy_start = 0;
    d.cascade = 0;
    p = &d;
  }
  c.variant = p->variant;
  c.kc = p->kc > 0 ? p->kc : 1;
  c.interval = p->relabel_interval > 0 ? p->relabel_interval : 1;
  c.max_rounds = p->max_rounds > 0
                     ? p->max_rounds
                     : static_cast<int>(std::min<int64_t>(4LL * E->n + 1000, INT32_MAX));
  c.entry = entry;
  c.scheduler = p->scheduler;
  c.lazy = p->lazy_start;
  c.cascade = p->cascade > 0 ? p->cascade : 0;
  return c;
}

// Zero the control block's counters before a solve.
int begin_solve(dmf_engine* E) {
  E->last = dmf_counters{};
  E->launches = 0;
  hipLaunchKernelGGL(k_begin, dim3(1), dim3(kBlock), 0, E->stream, E->D);
  ++E->launches;
  HIP_OK(hipGetLastError());
  return DMF_OK;
}

int finish_solve(dmf_engine* E, int rc_driver, int64_t* flow) {
  int rc = sync_ctrl(E);
  if (rc) return rc;
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, E->ev0, E->ev1));
  const Ctrl& c = *E->h_ctrl;
  unsigned long long ctr[C_N] = {};
  {
    std::vector<unsigned long long> slots(static_cast<size_t>(C_N) * kSlots);
    HIP_OK(hipMemcpyAsync(slots.data(), E->D.ctrs, sizeof(unsigned long long) * slots.size(),
                          hipMemcpyDeviceToHost, E->stream));
    HIP_OK(hipStreamSynchronize(E->stream));
    for (int i = 0; i < C_N; ++i)
      for (int j = 0; j < kSlots; ++j) ctr[i] += slots[static_cast<size_t>(i) * kSlots + j];
  }
  E->last.arc_visits = ctr[C_ARC];
  E->last.vertex_visits = ctr[C_VERT];
  E->last.pushes = ctr[C_PUSH];
  E->last.relabels = ctr[C_RELABEL];
  E->last.repairs = ctr[C_REPAIR];
  if (ctr[C_LEVELS] || ctr[C_GRELABEL]) {  // persistent driver counts on device
    E->last.bfs_levels = ctr[C_LEVELS];
    E->last.global_relabels = ctr[C_GRELABEL];
  }
  if (ctr[C_ROUNDS]) E->last.rounds = ctr[C_ROUNDS];
  E->last.updates = ctr[C_UPD] ? ctr[C_UPD] : E->last.updates;
  E->last.atomics = ctr[C_ATOMIC];
  E->last.bfs_arc_visits = ctr[C_BFSARC];
  E->last.bfs_vertex_visits = ctr[C_BFSVERT];
  E->last.launches = E->launches;
  E->last.device_ms = ms;
  E->last.variant = static_cast<uint64_t>(E->last_variant);
  if (c.missing) return DMF_ENOARC;
  if (c.invalid) return DMF_EINVAL;
  if (rc_driver) return rc_driver;
  if (c.status) return c.status;
  if (flow) *flow = c.flow;
  return DMF_OK;
}

int run_solve(dmf_engine* E, const dmf_params* p, int entry, int64_t* flow) {
  Cfg C = make_cfg(E, p, entry);
  if (C.variant < 0 || C.variant > 2 || C.scheduler < 0 || C.scheduler > 1) return DMF_EINVAL;
  if (p && (p->mode < 0 || p->mode > 3)) return DMF_EINVAL;
  if (entry == 1 && C.variant == DMF_VARIANT_PP && !E->pp_ready) C.variant = DMF_VARIANT_DATA;
  int rc = begin_solve(E);
  if (rc) return rc;
  HIP_OK(hipEventRecord(E->ev0, E->stream));
  ++E->solve_id;
  E->t_solve = std::chrono::steady_clock::now();
  // the scheduler runs between rounds on the host-driven path
  const bool stepped = (p && p->mode >= 1) || E->coop_blocks <= 0 ||
                       C.scheduler == DMF_SCHED_LEARNED || E->trace != nullptr;
  E->use_graph = p && (p->mode == 2 || p->mode == 3);
  E->use_rounds = p && p->mode == 3;
  if (!stepped) E->labels_fresh = false;
  int rc_driver = stepped ? solve_stepped(E, C) : solve_persistent(E, C);
  if (rc_driver == DMF_EDEVICE) return rc_driver;
  HIP_OK(hipEventRecord(E->ev1, E->stream));
  if (entry == 1) E->last.updates = static_cast<uint64_t>(E->D.k);
  rc = finish_solve(E, rc_driver, flow);
  if (entry == 0) E->pp_ready = rc == DMF_OK && E->last_variant == DMF_VARIANT_PP;
  if (E->trace) {
    std::fprintf(E->trace,
                 "{\"solve\":%llu,\"what\":\"summary\",\"entry\":%d,\"variant\":%d,\"scheduler\":%d,"
                 "\"interval\":%d,\"lazy\":%d,\"cascade\":%d,\"k\":%d,\"device_ms\":%.4f,\"wall_ms\":%.4f,"
                 "\"rounds\":%llu,\"relabels\":%llu,\"rc\":%d}\n",
                 (unsigned long long)E->solve_id, entry, E->last_variant, C.scheduler, C.interval, C.lazy, C.cascade,
                 E->D.k, E->last.device_ms, since_solve_ms(E), (unsigned long long)E->last.rounds,
                 (unsigned long long)E->last.global_relabels, rc);
    std::fflush(E->trace);
  }
  return rc;
}

// One phase of the preflow-to-flow completion.  deficits: every non-terminal
// deficit pulls flow back along the forward residual graph from the sinks
// (BFS_CT labels, never from a source); otherwise every non-terminal excess is
// pushed back along the residual graph to the sources (BFS_CS labels, never
// into a sink).  Labels are recomputed before every round, so the rounds need
// no label repair; the phase ends when no labelled vertex is imbalanced.
int complete_phase(dmf_engine* E, bool deficits, int kc, int max_rounds) {
  Dev& D = E->D;
  const int mode = deficits ? BFS_CT : BFS_CS;
  for (int round = 0;; ++round) {
    int visited = 0, rc;
    if ((rc = relabel(E, mode, 0, &visited))) return rc;
    if ((rc = reset_lists(E))) return rc;
    LAUNCH(k_build_imbalance_wl, blocks_for(visited), D, visited, deficits ? 1 : 0);
    if ((rc = sync_ctrl(E))) return rc;
    const int32_t cnt = deficits ? E->h_ctrl->wl2_cnt : E->h_ctrl->wl_cnt;
    const unsigned long long wpk = E->h_ctrl->wq[deficits ? kQPull : kQActive];
    if (cnt == 0 && wpk == 0) return DMF_OK;
    if (round >= max_rounds) return DMF_ENOCONV;
    if (deficits) LAUNCH_PUSH(k_push, phase_blocks(cnt, chunks_of(wpk)), D, 2, cnt, kc, PM_NO_SRC, wpk, 0);
    else LAUNCH_PUSH(k_push, phase_blocks(cnt, chunks_of(wpk)), D, 1, cnt, kc, PM_NO_SNK, wpk, 0);
  }
}

// Device bytes per update record an engine holds (source, destination, new
// capacity, arc, reverse arc, capacity change).
constexpr int64_t kUpdateSlotBytes = 24;

int ensure_update_capacity(dmf_engine* E, int32_t k) {
  if (k <= E->ucap) return DMF_OK;
  int32_t cap = std::max<int32_t>(k, 1024);
  hipFree(E->d_usrc);
  hipFree(E->d_udst);
  hipFree(E->d_uw);
  hipFree(E->d_uarc);
  hipFree(E->d_urev);
  hipFree(E->d_udelta);
  E->bytes -= static_cast<int64_t>(E->ucap) * kUpdateSlotBytes;
  E->d_usrc = E->d_udst = E->d_uw = E->d_uarc = E->d_urev = E->d_udelta = nullptr;
  E->ucap = 0;
  int rc;
  if ((rc = dalloc(E, &E->d_usrc, cap)) || (rc = dalloc(E, &E->d_udst, cap)) ||
      (rc = dalloc(E, &E->d_uw, cap)) || (rc = dalloc(E, &E->d_uarc, cap)) ||
      (rc = dalloc(E, &E->d_urev, cap)) || (rc = dalloc(E, &E->d_udelta, cap)))
    return rc;
  E->ucap = cap;
  return DMF_OK;
}

}  // namespace

namespace {

// (Re)builds the wide-vertex structures for thresholds lane_max and wide_t: the
// static list of every vertex of degree > wide_t with chunk bases, its bitmap,
// the wave list and the chunk queues sized for it.  The new lists are built in
// full before the old ones are released, so on failure the engine keeps its
// previous thresholds and lists intact.
int setup_wide(dmf_engine* E, int32_t lane_max, int32_t wide_t) {
  Dev& D = E->D;
  std::vector<int32_t> wide_v, wide_b, wave_v, wave_b;
  int64_t wide_chunks = 0, wave_chunks = 0;
  // Vertices above lane_max: push and repair give each its own wave (one per
  // chunk) from the active queues, which have a slot for every such vertex.
  // Topology-driven rounds take theirs from the static wave list of the
  // vertices above wave_lane (the rest one per lane).
  const int32_t queue_t = std::min(lane_max, wide_t);
  const int32_t wave_lane = static_wave_t(lane_max, wide_t);
  size_t nqueue_v = 0;
  for (int32_t v = 0; v < E->n; ++v) {
    const int32_t d = E->h_row[v + 1] - E->h_row[v];
    if (d > queue_t) ++nqueue_v;
    if (d > wave_lane) {
      wave_v.push_back(v);
      wave_b.push_back(static_cast<int32_t>(wave_chunks));
      wave_chunks += (d + kChunk - 1) / kChunk;
    }
    if (d > wide_t) {
      wide_v.push_back(v);
      wide_b.push_back(static_cast<int32_t>(wide_chunks));
      wide_chunks += (d + kChunk - 1) / kChunk;
    }
  }
  const size_t nwide = std::max<size_t>(1, wide_v.size());
  const size_t nqueue = std::max<size_t>(1, nqueue_v);
  const size_t nwave = std::max<size_t>(1, wave_v.size());
  std::vector<uint32_t> wbits((static_cast<size_t>(E->n) + 31) / 32 + 1, 0u);
  for (int32_t v : wide_v) wbits[static_cast<uint32_t>(v) >> 5] |= 1u << (v & 31);
  HIP_OK(hipStreamSynchronize(E->stream));  // no kernel uses the old lists from here on
  int32_t *wq_vert = nullptr, *wq_base = nullptr, *wide_vert = nullptr, *wide_base = nullptr;
  int32_t *wave_vert = nullptr, *wave_base = nullptr;
  uint32_t* widebits = nullptr;
  int4* bfs_ct = nullptr;
  const size_t ct_cap = static_cast<size_t>(std::max<int64_t>(1, wide_chunks));
  const int64_t before = E->bytes;
  auto put = [](void* dst, const void* src, size_t bytes) {
    return bytes == 0 || hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  int rc;
  if (!(rc = dalloc(E, &wq_vert, nqueue * kNumQ)) && !(rc = dalloc(E, &wq_base, nqueue * kNumQ)) &&
      !(rc = dalloc(E, &wide_vert, nwide)) && !(rc = dalloc(E, &wide_base, nwide)) &&
      !(rc = dalloc(E, &wave_vert, nwave)) && !(rc = dalloc(E, &wave_base, nwave)) &&
      !(rc = dalloc(E, &widebits, wbits.size())) && !(rc = dalloc(E, &bfs_ct, ct_cap * kNumRing))) {
    const bool copied = put(wide_vert, wide_v.data(), sizeof(int32_t) * wide_v.size()) &&
                        put(wide_base, wide_b.data(), sizeof(int32_t) * wide_b.size()) &&
                        put(wave_vert, wave_v.data(), sizeof(int32_t) * wave_v.size()) &&
                        put(wave_base, wave_b.data(), sizeof(int32_t) * wave_b.size()) &&
                        put(widebits, wbits.data(), sizeof(uint32_t) * wbits.size());
    if (!copied) rc = DMF_EDEVICE;
  }
  if (rc) {
    for (void* p : {static_cast<void*>(wq_vert), static_cast<void*>(wq_base), static_cast<void*>(wide_vert),
                    static_cast<void*>(wide_base), static_cast<void*>(wave_vert),
                    static_cast<void*>(wave_base), static_cast<void*>(widebits), static_cast<void*>(bfs_ct)})
      if (p) (void)hipFree(p);
    E->bytes = before;
    return rc;
  }
  for (void* p : {static_cast<void*>(D.wq_vert), static_cast<void*>(D.wq_base),
                  const_cast<void*>(static_cast<const void*>(D.wide_vert)),
                  const_cast<void*>(static_cast<const void*>(D.wide_base)),
                  const_cast<void*>(static_cast<const void*>(D.wave_vert)),
                  const_cast<void*>(static_cast<const void*>(D.wave_base)),
                  const_cast<void*>(static_cast<const void*>(D.widebits)), static_cast<void*>(D.bfs_ct)})
    if (p) (void)hipFree(p);
  const int64_t new_bytes = E->bytes - before;
  E->bytes -= E->wide_bytes;
  E->wide_bytes = new_bytes;
  // BFS segments captured for the old lists launch the old chunk count's grid
  destroy_segments(E);
  D.lane_max = lane_max;
  D.wave_lane = wave_lane;
  D.wq_vert = wq_vert;
  D.wq_base = wq_base;
  D.wq_cap = static_cast<int32_t>(nqueue);
  D.wide_vert = wide_vert;
  D.wide_base = wide_base;
  D.widebits = widebits;
  D.bfs_ct = bfs_ct;
  D.bfs_ct_cap = static_cast<int32_t>(ct_cap);
  D.wide_t = wide_t;
  D.wide_packed = (static_cast<unsigned long long>(wide_v.size()) << 32) |
                  static_cast<unsigned long long>(wide_chunks);
  D.wave_vert = wave_vert;
  D.wave_base = wave_base;
  D.wave_packed = (static_cast<unsigned long long>(wave_v.size()) << 32) |
                  static_cast<unsigned long long>(wave_chunks);
  return DMF_OK;
}

// ---------------------------------------------------------------- vertex scattering
// On graphs with hub vertices whose labels fit in one XCD's L2, the engine numbers
// the vertices in a seeded random order.  On RMAT-20 (1 M vertices, edge factor
// 16, 10 batches of 1-10% updates) this cut the mean global-relabel level from
// 54-55 to 48 us at the same arc visits, and the plain batch by about 7%, with
// equal flows; a degree-sorted numbering gave 53 us.  On RMAT-23 (8.4 M
// vertices, 16 MB of 16-bit labels) streaming batches were 20% slower scattered:
// there the generator's id locality saves L2 misses that a random order adds.
// Grids and meshes keep their numbering, whose neighbour locality a random
// order would destroy.  DMF_VERTEX_SCATTER=0/1 overrides the choice.
constexpr int64_t kScatterSkew = 32;          // max degree >= 32 x mean degree
constexpr int32_t kScatterMaxN = 1 << 21;     // 2 B labels of 2 M vertices fill 4 MB

bool want_scatter(const dmf_csr* g) {
  if (const char* e = std::getenv("DMF_VERTEX_SCATTER")) return std::atoi(e) != 0;
  if (g->n > kScatterMaxN) return false;
  int64_t maxd = 0;
  for (int32_t v = 0; v < g->n; ++v) maxd = std::max<int64_t>(maxd, g->row[v + 1] - g->row[v]);
  const int64_t mean = std::max<int64_t>(1, static_cast<int64_t>(g->m) / std::max(1, g->n));
  return maxd >= kScatterSkew * mean;
}

// Runs fn(v0, v1) over arc-balanced vertex ranges of the offsets row, on up to
// 16 threads for large graphs (inline where a thread cannot be had).
template <class F>
void for_vertex_ranges(const int32_t* row, int32_t n, F&& fn) {
  const int64_t m = row[n];
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int32_t parts = m < (1 << 22) ? 1 : static_cast<int32_t>(std::min(hw, 16u));
  std::vector<int32_t> cut(parts + 1, n);
  cut[0] = 0;
  for (int32_t i = 1; i < parts; ++i) {
    const int64_t target = m * i / parts;
    cut[i] = std::max(cut[i - 1], static_cast<int32_t>(std::lower_bound(row, row + n + 1, target) - row));
  }
  std::vector<std::thread> pool;
  for (int32_t i = 1; i < parts; ++i) {
    try {
      pool.emplace_back([&, i] { fn(cut[i], cut[i + 1]); });
    } catch (...) {
      fn(cut[i], cut[i + 1]);
    }
  }
  fn(cut[0], cut[1]);
  for (auto& th : pool) th.join();
}

// The caller's CSR renumbered by a seeded random permutation: vmap[v] is the
// new id of vertex v; vertex vmap[v] gets v's arcs in their order, heads and
// reverse indices translated.
void scatter_graph(const dmf_csr* g, std::vector<int32_t>& vmap, std::vector<int32_t>& row2,
                   std::vector<int32_t>& col2, std::vector<int32_t>& cap2, std::vector<int32_t>& rev2) {
  const int32_t n = g->n, m = g->m;
  std::vector<int32_t> inv(n);
  for (int32_t i = 0; i < n; ++i) inv[i] = i;
  uint64_t st = 0x5DEECE66Dull ^ (static_cast<uint64_t>(n) << 32) ^ static_cast<uint64_t>(m);
  auto next = [&st] {  // splitmix64
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  for (int32_t i = n - 1; i > 0; --i) {  // Fisher-Yates
    const int32_t j = static_cast<int32_t>((static_cast<__uint128_t>(next()) * (i + 1)) >> 64);
    std::swap(inv[i], inv[j]);
  }
  vmap.assign(n, 0);
  for (int32_t x = 0; x < n; ++x) vmap[inv[x]] = x;
  row2.assign(static_cast<size_t>(n) + 1, 0);
  for (int32_t x = 0; x < n; ++x) row2[x + 1] = row2[x] + (g->row[inv[x] + 1] - g->row[inv[x]]);
  col2.assign(m, 0);
  cap2.assign(m, 0);
  rev2.assign(m, 0);
  std::vector<int32_t> o2a(m);  // caller arc -> engine arc
  for_vertex_ranges(row2.data(), n, [&](int32_t x0, int32_t x1) {
    for (int32_t x = x0; x < x1; ++x) {
      const int32_t v = inv[x], base = g->row[v], deg = g->row[v + 1] - base;
      for (int32_t j = 0; j < deg; ++j) {
        o2a[base + j] = row2[x] + j;
        col2[row2[x] + j] = vmap[g->col[base + j]];
        cap2[row2[x] + j] = g->cap[base + j];
      }
    }
  });
  for_vertex_ranges(row2.data(), n, [&](int32_t x0, int32_t x1) {
    for (int32_t x = x0; x < x1; ++x) {
      const int32_t base = g->row[inv[x]], deg = row2[x + 1] - row2[x];
      for (int32_t j = 0; j < deg; ++j) rev2[row2[x] + j] = o2a[g->rev[base + j]];
    }
  });
}

// dmf_engine_read_state for a scattered engine: arrays back in the caller's
// vertex and arc numbering.
int read_state_scattered(dmf_engine* E, int32_t* res, int32_t* cap, int64_t* excess, int32_t* height) {
  const int32_t n = E->n;
  const std::vector<int32_t>&vm = E->h_vmap, &rc = E->h_row_caller, &re = E->h_row;
  std::vector<int32_t> a, b;
  std::vector<long long> x;
  try {
    if (res) a.resize(E->m);
    if (cap) b.resize(E->m);
    if (excess || height) x.resize(n);
  } catch (...) {
    return DMF_ENOMEM;
  }
  if (res) HIP_OK(hipMemcpyAsync(a.data(), E->D.res, sizeof(int32_t) * E->m, hipMemcpyDeviceToHost, E->stream));
  if (cap) HIP_OK(hipMemcpyAsync(b.data(), E->D.cap, sizeof(int32_t) * E->m, hipMemcpyDeviceToHost, E->stream));
  HIP_OK(hipStreamSynchronize(E->stream));
  for (int32_t v = 0; v < n; ++v) {
    const int32_t from = re[vm[v]], to = rc[v], deg = rc[v + 1] - rc[v];
    for (int32_t j = 0; j < deg; ++j) {
      if (res) res[to + j] = a[from + j];
      if (cap) cap[to + j] = b[from + j];
    }
  }
  if (excess) {
    HIP_OK(hipMemcpyAsync(x.data(), E->D.excess, sizeof(long long) * n, hipMemcpyDeviceToHost, E->stream));
    HIP_OK(hipStreamSynchronize(E->stream));
    for (int32_t v = 0; v < n; ++v) excess[v] = x[vm[v]];
  }
  if (height) {
    HIP_OK(hipMemcpyAsync(x.data(), E->D.height, sizeof(int32_t) * n, hipMemcpyDeviceToHost, E->stream));
    HIP_OK(hipStreamSynchronize(E->stream));
    const int32_t* h = reinterpret_cast<const int32_t*>(x.data());
    for (int32_t v = 0; v < n; ++v) height[v] = h[vm[v]];
  }
  return DMF_OK;
}

int create_engine(const dmf_csr* g_caller, const std::vector<int32_t>& srcs_caller,
                  const std::vector<int32_t>& snks_caller, const std::vector<int32_t>* inst, int32_t ninst,
                  dmf_engine** out) {
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return DMF_EDEVICE;
  dmf_engine* E = new (std::nothrow) dmf_engine();
  if (!E) return DMF_ENOMEM;
  // engine numbering (scattered for single graphs with hubs)
  std::vector<int32_t> srow, scol, scap, srev;
  std::vector<int32_t> srcs(srcs_caller), snks(snks_caller);
  dmf_csr scattered{};
  const dmf_csr* g = g_caller;
  try {
    if (!inst && want_scatter(g_caller)) {
      scatter_graph(g_caller, E->h_vmap, srow, scol, scap, srev);
      E->h_row_caller.assign(g_caller->row, g_caller->row + g_caller->n + 1);
      scattered = dmf_csr{g_caller->n, g_caller->m, srow.data(), scol.data(), scap.data(), srev.data()};
      g = &scattered;
      for (int32_t& v : srcs) v = E->h_vmap[v];
      for (int32_t& v : snks) v = E->h_vmap[v];
    }
    E->h_row.assign(g->row, g->row + g->n + 1);
  } catch (...) {  // host memory
    delete E;
    return DMF_ENOMEM;
  }
  E->n = g->n;
  E->m = g->m;
  E->s = srcs[0];
  E->t = snks[0];
  E->ninst = ninst;
  Dev& D = E->D;
  D.n = g->n;
  D.m = g->m;
  D.s = srcs[0];
  D.t = snks[0];
  D.lane_max = kLaneMax;
  D.nsrc = static_cast<int32_t>(srcs.size());
  D.nsnk = static_cast<int32_t>(snks.size());
  D.ninst = ninst;
  const size_t n = g->n, m = g->m;
  uint64_t hcap = 16;
  while (hcap < 2 * static_cast<uint64_t>(m) + 16) hcap <<= 1;
  std::vector<uint8_t> term(n, 0);
  for (int32_t v : srcs) term[v] |= 1;
  for (int32_t v : snks) term[v] |= 2;
  // Every pointer starts null and reaches E->D before the failure check, so a
  // failed allocation part way through frees what was already allocated.
  int32_t *row = nullptr, *col = nullptr, *rev = nullptr, *cap = nullptr, *res = nullptr,
          *rres = nullptr, *height = nullptr, *order = nullptr, *wl = nullptr, *wl2 = nullptr,
          *part = nullptr, *owner = nullptr;
  int32_t *d_srcs = nullptr, *d_snks = nullptr, *d_inst = nullptr;
  uint8_t* d_term = nullptr;
  long long* iflow = nullptr;
  unsigned long long* ctrs = nullptr;
  long long* excess = nullptr;
  uint8_t *side = nullptr, *pfinal = nullptr;
  Ctrl* ctrl = nullptr;
  ulonglong2* htab = nullptr;
  uint16_t* lev16 = nullptr;
  uint8_t* tgt = nullptr;
  int32_t* cmark = nullptr;
  int32_t* order2 = nullptr;
  int rc = DMF_OK;
  const bool alloc_failed = (rc = dalloc(E, &row, n + 1)) || (rc = dalloc(E, &col, m)) || (rc = dalloc(E, &rev, m)) ||
      (rc = dalloc(E, &cap, m)) || (rc = dalloc(E, &E->d_cap0, m)) || (rc = dalloc(E, &res, m)) ||
      (rc = dalloc(E, &rres, m)) || (rc = dalloc(E, &excess, n)) || (rc = dalloc(E, &height, n)) ||
      (rc = dalloc(E, &order, n)) || (rc = dalloc(E, &wl, n)) || (rc = dalloc(E, &wl2, n)) ||
      (rc = dalloc(E, &part, n)) || (rc = dalloc(E, &side, n)) || (rc = dalloc(E, &pfinal, n)) ||
      (rc = dalloc(E, &ctrl, 1)) || (rc = dalloc(E, &htab, hcap)) ||
      (rc = dalloc(E, &owner, m)) || (rc = dalloc(E, &ctrs, static_cast<size_t>(C_N) * kSlots)) ||
      (rc = dalloc(E, &d_term, n)) || (rc = dalloc(E, &d_srcs, srcs.size())) ||
      (rc = dalloc(E, &d_snks, snks.size())) || (rc = dalloc(E, &iflow, static_cast<size_t>(ninst))) ||
      (rc = dalloc(E, &lev16, n)) || (rc = dalloc(E, &tgt, n)) || (rc = dalloc(E, &cmark, n)) ||
      (rc = dalloc(E, &order2, n)) || (inst && (rc = dalloc(E, &d_inst, n))) ||
      (!E->h_vmap.empty() && (rc = dalloc(E, &E->d_vmap, n)));
  D.row = row; D.col = col; D.rev = rev; D.cap = cap; D.res = res; D.rres = rres;
  D.excess = excess; D.height = height; D.order = order; D.wl = wl; D.wl2 = wl2; D.part = part;
  D.side = side; D.pfinal = pfinal; D.ctrl = ctrl; D.htab = htab;
  D.hmask = hcap - 1; D.owner = owner; D.ctrs = ctrs;
  D.term = d_term; D.srcs = d_srcs; D.snks = d_snks; D.inst = d_inst; D.iflow = iflow;
  D.lev16 = lev16;
  D.tgt = tgt;
  D.cmark = cmark;
  E->ord[0] = order;
  E->ord[1] = order2;
  D.vmap = E->d_vmap;
  if (alloc_failed) {
    dmf_engine_destroy(E);
    return rc;
  }
  {
    // direction-optimizing relabel: off unless DMF_BFS_BOTTOM_UP gives the
    // divisor (on residual graphs near a maximum flow the source side is
    // unreachable and a bottom-up level scans all of it)
    const char* bu = std::getenv("DMF_BFS_BOTTOM_UP");
    D.bottom_up = bu ? std::atoi(bu) : 0;
    if (const char* ee = std::getenv("DMF_EARLY_RELABEL")) E->early_mode = std::atoi(ee);
    if (const char* cs = std::getenv("DMF_CASCADE")) E->cascade_env = std::max(0, std::atoi(cs));
    if (const char* ck = std::getenv("DMF_CASCADE_KC")) E->cascade_kc = std::max(0, std::atoi(ck));
    if (const char* cr = std::getenv("DMF_CASCADE_RELIST")) E->cascade_relist = std::atoi(cr);
    if (const char* br = std::getenv("DMF_BFS_RESTRICT")) E->restrict_mode = std::max(0, std::atoi(br));
  }
  if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess) {
    dmf_engine_destroy(E);
    return DMF_EDEVICE;
  }
  E->own_stream = true;
  if (hipHostMalloc(reinterpret_cast<void**>(&E->h_ctrl), sizeof(Ctrl), hipHostMallocDefault) != hipSuccess ||
      hipEventCreate(&E->ev0) != hipSuccess || hipEventCreate(&E->ev1) != hipSuccess ||
      dalloc(E, &E->d_dev, 1) != DMF_OK) {
    dmf_engine_destroy(E);
    return DMF_EDEVICE;
  }
  bool ok = hipMemcpyAsync(row, g->row, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, E->stream) == hipSuccess &&
            hipMemcpyAsync(col, g->col, sizeof(int32_t) * m, hipMemcpyHostToDevice, E->stream) == hipSuccess &&
            hipMemcpyAsync(rev, g->rev, sizeof(int32_t) * m, hipMemcpyHostToDevice, E->stream) == hipSuccess &&
            hipMemcpyAsync(cap, g->cap, sizeof(int32_t) * m, hipMemcpyHostToDevice, E->stream) == hipSuccess &&
            hipMemcpyAsync(E->d_cap0, g->cap, sizeof(int32_t) * m, hipMemcpyHostToDevice, E->stream) == hipSuccess &&
            hipMemcpyAsync(d_term, term.data(), n, hipMemcpyHostToDevice, E->stream) == hipSuccess &&
            hipMemcpyAsync(d_srcs, srcs.data(), sizeof(int32_t) * srcs.size(), hipMemcpyHostToDevice, E->stream) == hipSuccess &&
            hipMemcpyAsync(d_snks, snks.data(), sizeof(int32_t) * snks.size(), hipMemcpyHostToDevice, E->stream) == hipSuccess &&
            (!inst || hipMemcpyAsync(d_inst, inst->data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, E->stream) == hipSuccess) &&
            (!E->d_vmap || hipMemcpyAsync(E->d_vmap, E->h_vmap.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, E->stream) == hipSuccess) &&
            hipMemsetAsync(res, 0, sizeof(int32_t) * m, E->stream) == hipSuccess &&
            hipMemsetAsync(rres, 0, sizeof(int32_t) * m, E->stream) == hipSuccess &&
            hipMemsetAsync(excess, 0, sizeof(long long) * n, E->stream) == hipSuccess &&
            hipMemsetAsync(height, 0, sizeof(int32_t) * n, E->stream) == hipSuccess &&
            hipMemsetAsync(side, 0, n, E->stream) == hipSuccess &&
            hipMemsetAsync(cmark, 0, sizeof(int32_t) * n, E->stream) == hipSuccess &&
            hipMemsetAsync(pfinal, 0, n, E->stream) == hipSuccess &&
            hipMemsetAsync(ctrl, 0, sizeof(Ctrl), E->stream) == hipSuccess &&
            hipMemsetAsync(htab, 0xff, sizeof(ulonglong2) * hcap, E->stream) == hipSuccess &&
            hipMemsetAsync(owner, 0xff, sizeof(int32_t) * m, E->stream) == hipSuccess;
  if (!ok || setup_wide(E, kLaneMax, kWideT) != DMF_OK) {
    dmf_engine_destroy(E);
    return DMF_EDEVICE;
  }
  if (inst && ninst > 1) {
    // chunk bases of the terminals' arcs: the saturation of many sources (or
    // sinks) spreads their arcs over the grid
    std::vector<int32_t> cb;
    auto upload = [&](const std::vector<int32_t>& T, const int32_t** out, int32_t* chunks) -> bool {
      cb.assign(T.size(), 0);
      int64_t run = 0;
      for (size_t i = 0; i < T.size(); ++i) {
        cb[i] = static_cast<int32_t>(run);
        const int64_t d = static_cast<int64_t>(g->row[T[i] + 1]) - g->row[T[i]];
        run += (d + kChunk - 1) / kChunk;
      }
      int32_t* dev = nullptr;
      if (run > INT_MAX || dalloc(E, &dev, T.size()) != DMF_OK) return false;
      *out = dev;
      *chunks = static_cast<int32_t>(run);
      return hipMemcpy(dev, cb.data(), sizeof(int32_t) * T.size(), hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!upload(srcs, &D.src_cbase, &D.src_chunks) || !upload(snks, &D.snk_cbase, &D.snk_chunks)) {
      dmf_engine_destroy(E);
      return DMF_EDEVICE;
    }
  }
  {
    // the arc index is keyed by caller ids: a scattered engine builds it
    // through the inverse vertex map, held on the device for the build only
    int32_t* d_ivmap = nullptr;
    bool ok = true;
    if (!E->h_vmap.empty()) {
      std::vector<int32_t> iv(static_cast<size_t>(n));
      for (int32_t v = 0; v < n; ++v) iv[E->h_vmap[v]] = v;
      ok = hipMalloc(reinterpret_cast<void**>(&d_ivmap), sizeof(int32_t) * n) == hipSuccess &&
           hipMemcpy(d_ivmap, iv.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice) == hipSuccess;
    }
    if (ok) {
      hipLaunchKernelGGL(k_build_hash, dim3(blocks_for(m)), dim3(kBlock), 0, E->stream, D, d_ivmap);
      ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(E->stream) == hipSuccess;
    }
    if (d_ivmap) hipFree(d_ivmap);
    if (!ok) {
      dmf_engine_destroy(E);
      return DMF_EDEVICE;
    }
  }
  // Stepped grids: kGridBlocksPerCu blocks per CU (DMF_GRID_BLOCKS_PER_CU);
  // push, repair and cascade launches kPushBlocksPerCu (DMF_PUSH_BLOCKS_PER_CU).
  // Cooperative grid: as many blocks as can be co-resident.
  int dev = 0, per_cu = 0;
  hipDeviceProp_t prop{};
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    E->num_cu = prop.multiProcessorCount;
    const char* genv = std::getenv("DMF_GRID_BLOCKS_PER_CU");
    const int g = genv ? std::max(1, std::atoi(genv)) : kGridBlocksPerCu;
    E->grid_cap = static_cast<int>(std::min<int64_t>(8192, static_cast<int64_t>(std::max(1, E->num_cu)) * g));
    const char* penv = std::getenv("DMF_PUSH_BLOCKS_PER_CU");
    const int pg = penv ? std::max(1, std::atoi(penv)) : kPushBlocksPerCu;
    E->push_cap = static_cast<int>(std::min<int64_t>(8192, static_cast<int64_t>(std::max(1, E->num_cu)) * pg));
    if (prop.cooperativeLaunch &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_solve_persistent),
                                                     kBlock, 0) == hipSuccess &&
        per_cu > 0) {
      const char* capenv = std::getenv("DMF_COOP_BLOCKS_PER_CU");
      const int want = capenv ? std::max(1, std::atoi(capenv)) : 8;
      E->coop_blocks = prop.multiProcessorCount * std::min(per_cu, want);
    }
  }
  // scheduler weights (compiled in) and state
  if (hipMalloc(reinterpret_cast<void**>(&E->d_sched_w), sizeof(SchedWeights)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&E->d_sched), sizeof(SchedState)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&E->h_sched), sizeof(SchedState), hipHostMallocDefault) != hipSuccess ||
      hipMemcpy(E->d_sched_w, kSchedWeightsInit, sizeof(SchedWeights), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(E->d_sched, 0, sizeof(SchedState)) != hipSuccess) {
    dmf_engine_destroy(E);
    return DMF_EDEVICE;
  }
  E->bytes += sizeof(SchedWeights) + sizeof(SchedState);
  if (const char* tp = std::getenv("DMF_TRACE")) E->trace = std::fopen(tp, "a");
  *out = E;
  return DMF_OK;
}

// ------------------------------------------------ allocation-free one-shot solve
// dmf_static_recalculate_noalloc keeps the whole residual state of one graph of
// at most kNaMaxN vertices and kNaMaxM arcs, and kNaMaxK update records, in
// `scratch`, a fixed device array, instead of device memory it allocates.  The
// array is a host-side stand-in in host code: its device address is taken
// once per device with hipGetSymbolAddress and only that address is used.
constexpr int32_t kNaMaxN = 1 << 16;
constexpr int32_t kNaMaxM = 1 << 20;
constexpr int32_t kNaMaxK = 1 << 20;
constexpr int kNaMaxDevices = 64;

constexpr size_t na_align(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// Byte offsets of every array of the state inside the scratch array.
struct NaLayout {
  size_t row, col, rev, cap, res, rres, owner, excess, height, order, wl, wl2, part, side, pfinal, term, lev16,
      widebits, wq_vert, wq_base, wide_vert, wide_base, wave_vert, wave_base, ctrl, ctrs, iflow, srcs, snks, usrc,
      udst, uw, uarc, packed, total;
};

constexpr NaLayout na_layout() {
  NaLayout L{};
  size_t o = 0;
  auto take = [&o](size_t bytes) {
    const size_t at = o;
    o += na_align(bytes);
    return at;
  };
  const size_t N = kNaMaxN, M = kNaMaxM, K = kNaMaxK;
  L.row = take(4 * (N + 1));
  L.col = take(4 * M);
  L.rev = take(4 * M);
  L.cap = take(4 * M);
  L.res = take(4 * M);
  L.rres = take(4 * M);
  L.owner = take(4 * M);
  L.excess = take(8 * N);
  L.height = take(4 * N);
  L.order = take(4 * N);
  L.wl = take(4 * N);
  L.wl2 = take(4 * N);
  L.part = take(4 * N);
  L.side = take(N);
  L.pfinal = take(N);
  L.term = take(N);
  L.lev16 = take(2 * N);
  L.widebits = take(4 * (N / 32 + 1));
  L.wq_vert = take(4 * kNumQ * N);
  L.wq_base = take(4 * kNumQ * N);
  L.wide_vert = take(4 * N);
  L.wide_base = take(4 * N);
  L.wave_vert = take(4 * N);
  L.wave_base = take(4 * N);
  L.ctrl = take(sizeof(Ctrl));
  L.ctrs = take(8 * C_N * kSlots);
  L.iflow = take(8);
  L.srcs = take(4);
  L.snks = take(4);
  L.usrc = take(4 * K);
  L.udst = take(4 * K);
  L.uw = take(4 * K);
  L.uarc = take(4 * K);
  L.packed = take(16);
  L.total = o;
  return L;
}
constexpr NaLayout kNaLayout = na_layout();

}  // namespace

// The fixed workspace (about 46 MiB): held by every GPU the library runs on
// until the process ends.  Calls on one GPU share it, so calls that overlap on
// one GPU race; each call initialises what it reads.
static __device__ __align__(256) unsigned char scratch[kNaLayout.total];

namespace {

std::mutex g_na_mutex;
void* g_na_ptr[kNaMaxDevices] = {};

// Device address of `scratch` on the current device, taken on the first call
// on that device.
int na_scratch(void** out) {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNaMaxDevices) return DMF_EDEVICE;
  std::lock_guard<std::mutex> lock(g_na_mutex);
  if (!g_na_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      std::fprintf(stderr, "dmf: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
      return DMF_EDEVICE;
    }
    g_na_ptr[dev] = ptr;
  }
  *out = g_na_ptr[dev];
  return DMF_OK;
}

// dmf_csr_check's rules in place, without its threads and buffers.
int na_check(const dmf_csr* g) {
  if (!g || g->n < 0 || g->m < 0 || !g->row || (g->m > 0 && (!g->col || !g->cap || !g->rev))) return DMF_EINVAL;
  const int32_t n = g->n, m = g->m;
  if (g->row[0] != 0 || g->row[n] != m) return DMF_EFORMAT;
  for (int32_t v = 0; v < n; ++v) {
    if (g->row[v] > g->row[v + 1]) return DMF_EFORMAT;
    for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
      const int32_t h = g->col[e], r = g->rev[e];
      if (h < 0 || h >= n || g->cap[e] < 0 || r < 0 || r >= m) return DMF_EFORMAT;
      if (g->rev[r] != e || g->col[r] != v) return DMF_EFORMAT;
    }
  }
  return DMF_OK;
}

template <class T>
T* na_at(void* base, size_t off) {
  return reinterpret_cast<T*>(static_cast<unsigned char*>(base) + off);
}

}  // namespace

extern "C" {

int dmf_device_count(void) {
  return guarded([&]() -> int {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
  });
}

void dmf_params_default(const dmf_csr* g, dmf_params* p) {
  if (!p) return;
  p->variant = DMF_VARIANT_DATA;
  p->kc = (g && g->n > 0) ? std::max(1, g->m / g->n) : 1;
  p->relabel_interval = 1;
  p->scheduler = DMF_SCHED_FIXED;
  p->max_rounds = 0;
  p->mode = 2;  // stepped with graph-replayed BFS levels: the fastest driver
  p->lazy_start = 0;
  p->cascade = 0;
}

int dmf_engine_create(const dmf_csr* g, int32_t s, int32_t t, dmf_engine** out) {
  return guarded([&]() -> int {
    if (!g || !out || g->n < 2 || g->m < 0 || s < 0 || s >= g->n || t < 0 || t >= g->n || s == t)
      return DMF_EINVAL;
    *out = nullptr;
    if (const int rc = dmf_csr_check(g)) return rc;
    const std::vector<int32_t> srcs{s}, snks{t};
    return create_engine(g, srcs, snks, nullptr, 1, out);
  });
}

int dmf_batch_create(int32_t ninst, const dmf_csr* graphs, const int32_t* s, const int32_t* t,
                     dmf_engine** out) {
  return guarded([&]() -> int {
    if (!out || ninst < 1 || !graphs || !s || !t) return DMF_EINVAL;
    *out = nullptr;
    // disjoint union: instance i's vertices are shifted by voff[i], arcs by aoff[i]
    std::vector<int64_t> voff(ninst + 1, 0), aoff(ninst + 1, 0);
    for (int32_t i = 0; i < ninst; ++i) {
      const dmf_csr& gi = graphs[i];
      if (gi.n < 2 || gi.m < 0 || s[i] < 0 || s[i] >= gi.n || t[i] < 0 || t[i] >= gi.n || s[i] == t[i])
        return DMF_EINVAL;
      if (const int rc = dmf_csr_check(&gi)) return rc;
      voff[i + 1] = voff[i] + gi.n;
      aoff[i + 1] = aoff[i] + gi.m;
    }
    if (voff[ninst] > INT32_MAX - 1 || aoff[ninst] > INT32_MAX) return DMF_EINVAL;
    const int32_t n = static_cast<int32_t>(voff[ninst]), m = static_cast<int32_t>(aoff[ninst]);
    std::vector<int32_t> row(n + 1), col(m), cap(m), rev(m), inst(n), srcs(ninst), snks(ninst);
    for (int32_t i = 0; i < ninst; ++i) {
      const dmf_csr& gi = graphs[i];
      const int32_t vo = static_cast<int32_t>(voff[i]), ao = static_cast<int32_t>(aoff[i]);
      for (int32_t v = 0; v < gi.n; ++v) {
        row[vo + v] = ao + gi.row[v];
        inst[vo + v] = i;
      }
      for (int32_t e = 0; e < gi.m; ++e) {
        col[ao + e] = vo + gi.col[e];
        cap[ao + e] = gi.cap[e];
        rev[ao + e] = ao + gi.rev[e];
      }
      srcs[i] = vo + s[i];
      snks[i] = vo + t[i];
    }
    row[n] = m;
    dmf_csr u{n, m, row.data(), col.data(), cap.data(), rev.data()};
    int rc = create_engine(&u, srcs, snks, &inst, ninst, out);
    if (rc) return rc;
    (*out)->voff.assign(voff.begin(), voff.end());
    return DMF_OK;
  });
}

int dmf_engine_destroy(dmf_engine* E) {
  return guarded([&]() -> int {
    if (!E) return DMF_EINVAL;
    if (E->stream) hipStreamSynchronize(E->stream);
    Dev& D = E->D;
    hipFree(const_cast<int32_t*>(D.row));
    hipFree(const_cast<int32_t*>(D.col));
    hipFree(const_cast<int32_t*>(D.rev));
    hipFree(D.cap);
    hipFree(E->d_cap0);
    hipFree(D.res);
    hipFree(D.rres);
    hipFree(D.excess);
    hipFree(D.height);
    hipFree(E->ord[0]);
    hipFree(E->ord[1]);
    hipFree(D.wl);
    hipFree(D.wl2);
    hipFree(D.part);
    hipFree(D.side);
    hipFree(D.pfinal);
    hipFree(D.ctrl);
    hipFree(D.htab);
    hipFree(D.owner);
    hipFree(D.ctrs);
    hipFree(const_cast<uint8_t*>(D.term));
    hipFree(const_cast<int32_t*>(D.srcs));
    hipFree(const_cast<int32_t*>(D.snks));
    hipFree(const_cast<int32_t*>(D.src_cbase));
    hipFree(const_cast<int32_t*>(D.snk_cbase));
    hipFree(const_cast<int32_t*>(D.inst));
    hipFree(D.iflow);
    hipFree(D.lev16);
    hipFree(D.tgt);
    hipFree(D.cmark);
    hipFree(const_cast<uint32_t*>(D.widebits));
    hipFree(D.bfs_ct);
    hipFree(D.wq_vert);
    hipFree(D.wq_base);
    hipFree(const_cast<int32_t*>(D.wide_vert));
    hipFree(const_cast<int32_t*>(D.wide_base));
    hipFree(const_cast<int32_t*>(D.wave_vert));
    hipFree(const_cast<int32_t*>(D.wave_base));
    hipFree(E->d_usrc);
    hipFree(E->d_udst);
    hipFree(E->d_uw);
    hipFree(E->d_uarc);
    hipFree(E->d_urev);
    hipFree(E->d_udelta);
    hipFree(E->d_vmap);
    destroy_segments(E);
    hipFree(E->d_dev);
    if (E->h_ctrl) hipHostFree(E->h_ctrl);
    if (E->h_sched) hipHostFree(E->h_sched);
    hipFree(E->d_sched_w);
    hipFree(E->d_sched);
    if (E->trace) std::fclose(E->trace);
    if (E->ev0) hipEventDestroy(E->ev0);
    if (E->pev0) (void)hipEventDestroy(E->pev0);
    if (E->pev1) (void)hipEventDestroy(E->pev1);
    if (E->ev1) hipEventDestroy(E->ev1);
    if (E->ev_fork) (void)hipEventDestroy(E->ev_fork);
    if (E->ev_join) (void)hipEventDestroy(E->ev_join);
    if (E->pull_stream) (void)hipStreamDestroy(E->pull_stream);
    if (E->own_stream && E->stream) hipStreamDestroy(E->stream);
    delete E;
    return DMF_OK;
  });
}

int dmf_engine_set_stream(dmf_engine* E, void* hip_stream) {
  return guarded([&]() -> int {
    if (!E) return DMF_EINVAL;
    HIP_OK(hipStreamSynchronize(E->stream));
    if (E->own_stream && E->stream) hipStreamDestroy(E->stream);
    if (hip_stream) {
      E->stream = static_cast<hipStream_t>(hip_stream);
      E->own_stream = false;
    } else {
      HIP_OK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
      E->own_stream = true;
    }
    return DMF_OK;
  });
}

int dmf_engine_reset(dmf_engine* E) {
  return guarded([&]() -> int {
    if (!E) return DMF_EINVAL;
    Dev& D = E->D;
    HIP_OK(hipMemcpyAsync(D.cap, E->d_cap0, sizeof(int32_t) * E->m, hipMemcpyDeviceToDevice, E->stream));
    HIP_OK(hipMemsetAsync(D.res, 0, sizeof(int32_t) * E->m, E->stream));
    HIP_OK(hipMemsetAsync(D.rres, 0, sizeof(int32_t) * E->m, E->stream));
    HIP_OK(hipMemsetAsync(D.excess, 0, sizeof(long long) * E->n, E->stream));
    HIP_OK(hipMemsetAsync(D.side, 0, E->n, E->stream));
    HIP_OK(hipStreamSynchronize(E->stream));
    E->pp_ready = false;
    E->labels_fresh = false;
    return DMF_OK;
  });
}

int dmf_engine_static_solve(dmf_engine* E, const dmf_params* p, int64_t* flow) {
  return guarded([&]() -> int {
    if (!E) return DMF_EINVAL;
    E->D.k = 0;
    return run_solve(E, p, 0, flow);
  });
}

int dmf_engine_apply_batch_device(dmf_engine* E, int32_t k, const int32_t* d_src,
                                  const int32_t* d_dst, const int32_t* d_w, const dmf_params* p,
                                  int64_t* flow) {
  return guarded([&]() -> int {
    if (!E || k < 0 || (k > 0 && (!d_src || !d_dst || !d_w))) return DMF_EINVAL;
    int rc = ensure_update_capacity(E, std::max(k, 1));
    if (rc) return rc;
    E->D.usrc = d_src;
    E->D.udst = d_dst;
    E->D.uw = d_w;
    E->D.uarc = E->d_uarc;
    E->D.urev = E->d_urev;
    E->D.udelta = E->d_udelta;
    E->D.k = k;
    rc = run_solve(E, p, 1, flow);
    E->D.k = 0;
    return rc;
  });
}

int dmf_engine_apply_batch(dmf_engine* E, int32_t k, const int32_t* src, const int32_t* dst,
                           const int32_t* w, const dmf_params* p, int64_t* flow) {
  return guarded([&]() -> int {
    if (!E || k < 0 || (k > 0 && (!src || !dst || !w))) return DMF_EINVAL;
    int rc = ensure_update_capacity(E, std::max(k, 1));
    if (rc) return rc;
    if (k > 0) {
      HIP_OK(hipMemcpyAsync(E->d_usrc, src, sizeof(int32_t) * k, hipMemcpyHostToDevice, E->stream));
      HIP_OK(hipMemcpyAsync(E->d_udst, dst, sizeof(int32_t) * k, hipMemcpyHostToDevice, E->stream));
      HIP_OK(hipMemcpyAsync(E->d_uw, w, sizeof(int32_t) * k, hipMemcpyHostToDevice, E->stream));
    }
    return dmf_engine_apply_batch_device(E, k, E->d_usrc, E->d_udst, E->d_uw, p, flow);
  });
}

int dmf_engine_flow(dmf_engine* E, int64_t* flow) {
  return guarded([&]() -> int {
    if (!E || !flow) return DMF_EINVAL;
    hipLaunchKernelGGL(k_flow_prep, dim3(1), dim3(kBlock), 0, E->stream, E->D);
    hipLaunchKernelGGL(k_flow, dim3(blocks_for(E->n)), dim3(kBlock), 0, E->stream, E->D);
    HIP_OK(hipGetLastError());
    int rc = sync_ctrl(E);
    if (rc) return rc;
    *flow = E->h_ctrl->flow;
    return DMF_OK;
  });
}

int dmf_engine_min_cut(dmf_engine* E, uint8_t* side) {
  return guarded([&]() -> int {
    if (!E || !side) return DMF_EINVAL;
    // The search is a BFS_CUT relabel on a copy of the device state whose
    // labels, visit order and 16-bit filter are the worklists (scratch
    // between solves), so the labels and labelled set order[0, visited) that
    // a lazily started batch reuses stay as they are.
    const Dev saved = E->D;
    const dmf_counters last = E->last;
    const int last_levels = E->last_levels;
    int32_t* mark = E->D.wl2;
    E->D.height = mark;
    E->D.order = E->D.wl;
    E->D.lev16 = reinterpret_cast<uint16_t*>(E->D.part);  // 4n bytes hold n 16-bit labels
    int visited = 0;
    int rc = relabel(E, BFS_CUT, 0, &visited);
    E->D = saved;
    E->last = last;
    E->last_levels = last_levels;
    if (rc) return rc;
    std::vector<int32_t> h(E->n);
    HIP_OK(hipMemcpyAsync(h.data(), mark, sizeof(int32_t) * E->n, hipMemcpyDeviceToHost, E->stream));
    HIP_OK(hipStreamSynchronize(E->stream));
    if (E->h_vmap.empty()) {
      for (int32_t v = 0; v < E->n; ++v) side[v] = h[v] < E->n ? 1 : 0;
    } else {
      for (int32_t v = 0; v < E->n; ++v) side[v] = h[E->h_vmap[v]] < E->n ? 1 : 0;
    }
    return DMF_OK;
  });
}

int dmf_engine_complete_flow(dmf_engine* E, int64_t* flow) {
  return guarded([&]() -> int {
    if (!E) return DMF_EINVAL;
    Dev& D = E->D;
    // the solve's counters stay those of the last solve
    const dmf_counters saved = E->last;
    const uint64_t launches = E->launches;
    E->labels_fresh = false;  // heights become distances to the terminals
    const int kc = std::max(1, E->m / std::max(1, E->n));
    const int cap = static_cast<int>(std::min<int64_t>(4LL * E->n + 1000, INT32_MAX));
    int rc = complete_phase(E, true, kc, cap);
    if (!rc) rc = complete_phase(E, false, kc, cap);
    if (!rc) rc = reset_lists(E);
    if (!rc) rc = zero_dev(E, &D.ctrl->part_cnt, sizeof(int32_t));
    if (!rc) {
      LAUNCH(k_count_imbalance, blocks_for(E->n), D);
      rc = sync_ctrl(E);
    }
    int32_t left = 0;
    if (!rc) {
      left = E->h_ctrl->part_cnt;
      rc = zero_dev(E, &D.ctrl->part_cnt, sizeof(int32_t));
    }
    E->last = saved;
    E->launches = launches;
    if (rc) return rc;
    if (left) return DMF_ENOCONV;
    if (!flow) return DMF_OK;
    return dmf_engine_flow(E, flow);
  });
}

int dmf_engine_read_state(dmf_engine* E, int32_t* res, int32_t* cap, int64_t* excess,
                          int32_t* height) {
  return guarded([&]() -> int {
    if (!E) return DMF_EINVAL;
    if (!E->h_vmap.empty()) return read_state_scattered(E, res, cap, excess, height);
    if (res) HIP_OK(hipMemcpyAsync(res, E->D.res, sizeof(int32_t) * E->m, hipMemcpyDeviceToHost, E->stream));
    if (cap) HIP_OK(hipMemcpyAsync(cap, E->D.cap, sizeof(int32_t) * E->m, hipMemcpyDeviceToHost, E->stream));
    if (excess)
      HIP_OK(hipMemcpyAsync(excess, E->D.excess, sizeof(long long) * E->n, hipMemcpyDeviceToHost, E->stream));
    if (height)
      HIP_OK(hipMemcpyAsync(height, E->D.height, sizeof(int32_t) * E->n, hipMemcpyDeviceToHost, E->stream));
    HIP_OK(hipStreamSynchronize(E->stream));
    return DMF_OK;
  });
}

int dmf_engine_counters(dmf_engine* E, dmf_counters* c) {
  return guarded([&]() -> int {
    if (!E || !c) return DMF_EINVAL;
    *c = E->last;
    return DMF_OK;
  });
}

int64_t dmf_engine_device_bytes(const dmf_engine* E) { return E ? E->bytes : 0; }

int dmf_engine_set_profiling(dmf_engine* E, int on) {
  return guarded([&]() -> int {
    if (!E) return DMF_EINVAL;
    if (on && !E->pev0) {
      HIP_OK(hipEventCreate(&E->pev0));
      HIP_OK(hipEventCreate(&E->pev1));
    }
    E->profile = on != 0;
    return DMF_OK;
  });
}

int64_t dmf_engine_footprint(int32_t n, int32_t m) {
  if (n < 0 || m < 0) return 0;
  const int64_t N = n, M = m;
  uint64_t hcap = 16;
  while (hcap < 2 * static_cast<uint64_t>(M) + 16) hcap <<= 1;
  int64_t b = 4 * (N + 1);                      // row
  b += 7 * 4 * M;                               // col rev cap cap0 res rres owner
  b += 8 * N + 5 * 4 * N + 3 * N + 4 * N;       // excess; height order wl wl2 part; side pfinal term; inst or vmap
  b += 2 * N;                                   // 16-bit BFS levels
  b += N;                                       // early-exit relabel targets
  b += 4 * N;                                   // cascade list stamps
  b += 4 * N;                                   // second BFS order buffer
  b += static_cast<int64_t>(hcap) * 16;         // arc hash slots {key, value}
  b += 4 * (N / 32 + 1);                        // wide-vertex bitmap
  b += 8 * (M / 64 + 1);                        // static wide list (wide_t >= 64)
  b += 3 * 16 * (M / kChunk + M / 64 + 1);      // BFS ring chunk tables: chunks of the vertices above wide_t
  b += 48 * std::min(N + 1, M / 5 + 1);         // queues and wave list: per vertex above lane_max (>= 4)
  b += static_cast<int64_t>(C_N) * kSlots * 8;  // counters
  b += int64_t{1} << 20;                        // control block, scheduler, allocation slack
  return b;
}

int dmf_engine_reserve_updates(dmf_engine* E, int32_t k) {
  return guarded([&]() -> int {
    if (!E || k < 0) return DMF_EINVAL;
    return ensure_update_capacity(E, std::max(k, 1));
  });
}

int32_t dmf_engine_instances(const dmf_engine* E) { return E ? E->ninst : 0; }

int dmf_batch_flows(dmf_engine* E, int64_t* flows) {
  return guarded([&]() -> int {
    if (!E || !flows) return DMF_EINVAL;
    int64_t total = 0;
    int rc = dmf_engine_flow(E, &total);
    if (rc) return rc;
    HIP_OK(hipMemcpy(flows, E->D.iflow, sizeof(int64_t) * E->ninst, hipMemcpyDeviceToHost));
    if (E->ninst == 1) flows[0] = total;
    return DMF_OK;
  });
}

int dmf_batch_static_solve(dmf_engine* E, const dmf_params* p, int64_t* flows) {
  return guarded([&]() -> int {
    if (!E || !flows) return DMF_EINVAL;
    int64_t total = 0;
    int rc = dmf_engine_static_solve(E, p, &total);
    if (rc) return rc;
    return dmf_batch_flows(E, flows);
  });
}

int dmf_batch_apply(dmf_engine* E, const int32_t* counts, const int32_t* src, const int32_t* dst,
                    const int32_t* w, const dmf_params* p, int64_t* flows) {
  return guarded([&]() -> int {
    if (!E || !counts || !flows) return DMF_EINVAL;
    if (E->voff.empty()) E->voff = {0, E->n};
    int64_t total = 0;
    for (int32_t i = 0; i < E->ninst; ++i) {
      if (counts[i] < 0) return DMF_EINVAL;
      total += counts[i];
    }
    if (total > INT32_MAX || (total > 0 && (!src || !dst || !w))) return DMF_EINVAL;
    std::vector<int32_t> gs(total), gd(total);
    int64_t j = 0;
    for (int32_t i = 0; i < E->ninst; ++i) {
      const int64_t nv = E->voff[i + 1] - E->voff[i];
      for (int32_t q = 0; q < counts[i]; ++q, ++j) {
        if (src[j] < 0 || src[j] >= nv || dst[j] < 0 || dst[j] >= nv) return DMF_EINVAL;
        gs[j] = static_cast<int32_t>(E->voff[i] + src[j]);
        gd[j] = static_cast<int32_t>(E->voff[i] + dst[j]);
      }
    }
    int64_t f = 0;
    int rc = dmf_engine_apply_batch(E, static_cast<int32_t>(total), gs.data(), gd.data(), w, p, &f);
    if (rc) return rc;
    return dmf_batch_flows(E, flows);
  });
}

int dmf_engine_autotune(dmf_engine* E, const dmf_params* p, int32_t* lane_max, int32_t* wide_t) {
  return guarded([&]() -> int {
    if (!E) return DMF_EINVAL;
    static const int32_t cand[][2] = {{16, 128}, {32, 256}, {32, 1024}, {64, 512}, {16, 64}, {4, 64}, {4, 128}};
    double best_ms = 1e300;
    int best = 1;
    for (int c = 0; c < static_cast<int>(sizeof(cand) / sizeof(cand[0])); ++c) {
      int rc = setup_wide(E, cand[c][0], cand[c][1]);
      if (!rc) rc = dmf_engine_reset(E);
      int64_t f = 0;
      double ms = 0;
      for (int rep = 0; rep < 2 && !rc; ++rep) {  // first solve warms caches
        rc = dmf_engine_static_solve(E, p, &f);
        ms = E->last.device_ms;
      }
      if (rc) return rc;
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
    int rc = setup_wide(E, cand[best][0], cand[best][1]);
    if (!rc) rc = dmf_engine_reset(E);
    if (lane_max) *lane_max = cand[best][0];
    if (wide_t) *wide_t = cand[best][1];
    return rc;
  });
}

int dmf_engine_autotune_schedule(dmf_engine* E, dmf_params* p, int32_t* lane_max, int32_t* wide_t) {
  return guarded([&]() -> int {
    if (!E || !p) return DMF_EINVAL;
    int rc = dmf_engine_autotune(E, p, lane_max, wide_t);
    if (rc || p->scheduler != DMF_SCHED_FIXED) return rc;
    static const int32_t cand[] = {1, 2, 3, 5, 8};
    double best_ms = 1e300;
    int32_t best = 1;
    for (int32_t iv : cand) {
      dmf_params q = *p;
      q.relabel_interval = iv;
      double ms = 0;
      for (int rep = 0; rep < 2; ++rep) {  // the first solve warms caches
        if ((rc = dmf_engine_reset(E))) return rc;
        int64_t f = 0;
        if ((rc = dmf_engine_static_solve(E, &q, &f))) return rc;
        ms = E->last.device_ms;
      }
      if (ms < best_ms) {
        best_ms = ms;
        best = iv;
      }
    }
    p->relabel_interval = best;
    return dmf_engine_reset(E);
  });
}

int dmf_recalculate_max_flow(const dmf_csr* g, int32_t k, const int32_t* src, const int32_t* dst,
                             const int32_t* w, int32_t s, int32_t t, int32_t kc, int32_t variant,
                             int64_t* flow, double* ms_static, double* ms_batch) {
  return guarded([&]() -> int {
    dmf_engine* E = nullptr;
    int rc = dmf_engine_create(g, s, t, &E);
    if (rc) return rc;
    dmf_params p;
    dmf_params_default(g, &p);
    p.kc = kc > 0 ? kc : p.kc;
    p.variant = variant;
    int64_t f = 0;
    rc = dmf_engine_static_solve(E, &p, &f);
    if (!rc && ms_static) *ms_static = E->last.device_ms;
    if (!rc) rc = dmf_engine_apply_batch(E, k, src, dst, w, &p, &f);
    if (!rc && ms_batch) *ms_batch = E->last.device_ms;
    if (!rc && flow) *flow = f;
    dmf_engine_destroy(E);
    return rc;
  });
}

int dmf_static_recalculate(const dmf_csr* g, int32_t k, const int32_t* src, const int32_t* dst,
                           const int32_t* w, int32_t s, int32_t t, int32_t kc, int32_t variant,
                           int64_t* flow, double* ms) {
  return guarded([&]() -> int {
    dmf_engine* E = nullptr;
    int rc = dmf_engine_create(g, s, t, &E);
    if (rc) return rc;
    // update_kernel (code/dyn_data_bench.cu:300-322): set capacities, then solve.
    rc = ensure_update_capacity(E, std::max(k, 1));
    if (!rc && k > 0) {
      Dev& D = E->D;
      bool ok = hipMemcpyAsync(E->d_usrc, src, sizeof(int32_t) * k, hipMemcpyHostToDevice, E->stream) == hipSuccess &&
                hipMemcpyAsync(E->d_udst, dst, sizeof(int32_t) * k, hipMemcpyHostToDevice, E->stream) == hipSuccess &&
                hipMemcpyAsync(E->d_uw, w, sizeof(int32_t) * k, hipMemcpyHostToDevice, E->stream) == hipSuccess &&
                hipMemsetAsync(D.ctrl, 0, sizeof(Ctrl), E->stream) == hipSuccess;
      if (!ok) rc = DMF_EDEVICE;
      if (!rc) {
        D.usrc = E->d_usrc;
        D.udst = E->d_udst;
        D.uw = E->d_uw;
        D.uarc = E->d_uarc;
        D.urev = E->d_urev;
        D.udelta = E->d_udelta;
        D.k = k;
        hipLaunchKernelGGL(k_upd_lookup, dim3(blocks_for(k)), dim3(kBlock), 0, E->stream, D);
        hipLaunchKernelGGL(k_upd_apply, dim3(blocks_for(k)), dim3(kBlock), 0, E->stream, D, 1, 0);
        hipLaunchKernelGGL(k_upd_release, dim3(blocks_for(k)), dim3(kBlock), 0, E->stream, D);
        if (hipGetLastError() != hipSuccess) rc = DMF_EDEVICE;
        if (!rc) rc = sync_ctrl(E);
        if (!rc && E->h_ctrl->missing) rc = DMF_ENOARC;
        if (!rc && E->h_ctrl->invalid) rc = DMF_EINVAL;
        D.k = 0;
      }
    }
    dmf_params p;
    dmf_params_default(g, &p);
    p.kc = kc > 0 ? kc : p.kc;
    p.variant = variant;
    int64_t f = 0;
    if (!rc) rc = dmf_engine_static_solve(E, &p, &f);
    if (!rc && ms) *ms = E->last.device_ms;
    if (!rc && flow) *flow = f;
    dmf_engine_destroy(E);
    return rc;
  });
}

int dmf_static_recalculate_noalloc(const dmf_csr* g, int32_t k, const int32_t* src, const int32_t* dst,
                                   const int32_t* w, int32_t s, int32_t t, int32_t kc, int32_t variant,
                                   void* hip_stream, int64_t* flow) {
  return guarded([&]() -> int {
    if (!g || !flow || k < 0 || (k > 0 && (!src || !dst || !w)) || variant < 0 || variant > 2) return DMF_EINVAL;
    if (g->n > kNaMaxN || g->m > kNaMaxM || k > kNaMaxK) {
      std::fprintf(stderr,
                   "dmf: dmf_static_recalculate_noalloc takes at most %d vertices, %d arcs and %d updates "
                   "(got %d, %d, %d)\n",
                   kNaMaxN, kNaMaxM, kNaMaxK, g->n, g->m, k);
      return DMF_ESIZE;
    }
    if (g->n < 2 || s < 0 || s >= g->n || t < 0 || t >= g->n || s == t) return DMF_EINVAL;
    if (const int rc = na_check(g)) return rc;
    void* base = nullptr;
    if (const int rc = na_scratch(&base)) return rc;
    const NaLayout& L = kNaLayout;
    const int32_t n = g->n, m = g->m;
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    // an engine on the stack over the workspace: no stream, event, pinned
    // buffer or graph of its own; the stepped driver only
    dmf_engine E;
    E.n = n;
    E.m = m;
    E.s = s;
    E.t = t;
    E.stream = st;
    Ctrl hc{};
    E.h_ctrl = &hc;
    int dev = 0;
    hipDeviceProp_t prop{};
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipGetDeviceProperties(&prop, dev));
    E.num_cu = prop.multiProcessorCount;
    E.grid_cap = static_cast<int>(std::min<int64_t>(8192, static_cast<int64_t>(std::max(1, E.num_cu)) * kGridBlocksPerCu));
    E.push_cap = static_cast<int>(std::min<int64_t>(8192, static_cast<int64_t>(std::max(1, E.num_cu)) * kPushBlocksPerCu));
    Dev& D = E.D;
    D.n = n;
    D.m = m;
    D.s = s;
    D.t = t;
    D.nsrc = D.nsnk = 1;
    D.ninst = 1;
    D.lane_max = kLaneMax;
    D.wide_t = kWideT;
    D.wave_lane = static_wave_t(kLaneMax, kWideT);
    D.row = na_at<int32_t>(base, L.row);
    D.col = na_at<int32_t>(base, L.col);
    D.rev = na_at<int32_t>(base, L.rev);
    D.cap = na_at<int32_t>(base, L.cap);
    D.res = na_at<int32_t>(base, L.res);
    D.rres = na_at<int32_t>(base, L.rres);
    D.owner = na_at<int32_t>(base, L.owner);
    D.excess = na_at<long long>(base, L.excess);
    D.height = na_at<int32_t>(base, L.height);
    D.order = na_at<int32_t>(base, L.order);
    D.wl = na_at<int32_t>(base, L.wl);
    D.wl2 = na_at<int32_t>(base, L.wl2);
    D.part = na_at<int32_t>(base, L.part);
    D.side = na_at<uint8_t>(base, L.side);
    D.pfinal = na_at<uint8_t>(base, L.pfinal);
    uint8_t* term = na_at<uint8_t>(base, L.term);
    D.term = term;
    D.lev16 = na_at<uint16_t>(base, L.lev16);
    uint32_t* widebits = na_at<uint32_t>(base, L.widebits);
    D.widebits = widebits;
    D.wq_vert = na_at<int32_t>(base, L.wq_vert);
    D.wq_base = na_at<int32_t>(base, L.wq_base);
    D.wq_cap = n;
    int32_t* wide_vert = na_at<int32_t>(base, L.wide_vert);
    int32_t* wide_base = na_at<int32_t>(base, L.wide_base);
    int32_t* wave_vert = na_at<int32_t>(base, L.wave_vert);
    int32_t* wave_base = na_at<int32_t>(base, L.wave_base);
    D.wide_vert = wide_vert;
    D.wide_base = wide_base;
    D.wave_vert = wave_vert;
    D.wave_base = wave_base;
    D.ctrl = na_at<Ctrl>(base, L.ctrl);
    D.ctrs = na_at<unsigned long long>(base, L.ctrs);
    D.iflow = na_at<long long>(base, L.iflow);
    int32_t* srcs = na_at<int32_t>(base, L.srcs);
    int32_t* snks = na_at<int32_t>(base, L.snks);
    D.srcs = srcs;
    D.snks = snks;
    D.htab = nullptr;  // arcs are found by scanning the tail's adjacency
    D.hmask = 0;
    int32_t* usrc = na_at<int32_t>(base, L.usrc);
    int32_t* udst = na_at<int32_t>(base, L.udst);
    int32_t* uw = na_at<int32_t>(base, L.uw);
    D.usrc = usrc;
    D.udst = udst;
    D.uw = uw;
    D.uarc = na_at<int32_t>(base, L.uarc);
    unsigned long long* packed = na_at<unsigned long long>(base, L.packed);
    const int32_t s1 = s, t1 = t;
    HIP_OK(hipMemcpyAsync(const_cast<int32_t*>(D.row), g->row, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st));
    if (m > 0) {
      HIP_OK(hipMemcpyAsync(const_cast<int32_t*>(D.col), g->col, sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
      HIP_OK(hipMemcpyAsync(const_cast<int32_t*>(D.rev), g->rev, sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
      HIP_OK(hipMemcpyAsync(D.cap, g->cap, sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
      HIP_OK(hipMemsetAsync(D.owner, 0xff, sizeof(int32_t) * m, st));
    }
    if (k > 0) {
      HIP_OK(hipMemcpyAsync(usrc, src, sizeof(int32_t) * k, hipMemcpyHostToDevice, st));
      HIP_OK(hipMemcpyAsync(udst, dst, sizeof(int32_t) * k, hipMemcpyHostToDevice, st));
      HIP_OK(hipMemcpyAsync(uw, w, sizeof(int32_t) * k, hipMemcpyHostToDevice, st));
    }
    HIP_OK(hipMemcpyAsync(srcs, &s1, sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(snks, &t1, sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemsetAsync(term, 0, n, st));
    HIP_OK(hipMemsetAsync(term + s, 1, 1, st));
    HIP_OK(hipMemsetAsync(term + t, 2, 1, st));
    HIP_OK(hipMemsetAsync(D.side, 0, n, st));
    HIP_OK(hipMemsetAsync(D.pfinal, 0, n, st));
    HIP_OK(hipMemsetAsync(widebits, 0, sizeof(uint32_t) * (n / 32 + 1), st));
    HIP_OK(hipMemsetAsync(packed, 0, 16, st));
    HIP_OK(hipMemsetAsync(D.ctrl, 0, sizeof(Ctrl), st));
    HIP_OK(hipMemsetAsync(D.ctrs, 0, sizeof(unsigned long long) * C_N * kSlots, st));
    hipLaunchKernelGGL(k_build_static_lists, dim3(blocks_for(n)), dim3(kBlock), 0, st, D, wide_vert, wide_base,
                       wave_vert, wave_base, widebits, packed);
    HIP_OK(hipGetLastError());
    unsigned long long words[2] = {0, 0};
    HIP_OK(hipMemcpyAsync(words, packed, sizeof(words), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    D.wide_packed = words[0];
    D.wave_packed = words[1];
    // update_kernel (code/dyn_data_bench.cu:300-322): capacities from the records
    if (k > 0) {
      D.k = k;
      hipLaunchKernelGGL(k_upd_lookup, dim3(blocks_for(k)), dim3(kBlock), 0, st, D);
      hipLaunchKernelGGL(k_upd_apply, dim3(blocks_for(k)), dim3(kBlock), 0, st, D, 1, 0);
      hipLaunchKernelGGL(k_upd_release, dim3(blocks_for(k)), dim3(kBlock), 0, st, D);
      HIP_OK(hipGetLastError());
      if (const int rc = sync_ctrl(&E)) return rc;
      if (hc.missing) return DMF_ENOARC;
      if (hc.invalid) return DMF_EINVAL;
      D.k = 0;
      HIP_OK(hipMemsetAsync(D.ctrl, 0, sizeof(Ctrl), st));
    }
    dmf_params p;
    dmf_params_default(g, &p);
    p.kc = kc > 0 ? kc : p.kc;
    p.variant = variant;
    p.mode = 1;
    const Cfg C = make_cfg(&E, &p, 0);
    int rc = solve_stepped(&E, C);
    if (rc == DMF_EDEVICE) return rc;
    const int rc_sync = sync_ctrl(&E);
    if (rc_sync) return rc_sync;
    if (rc) return rc;
    if (hc.status) return hc.status;
    *flow = hc.flow;
    return DMF_OK;
  });
}

}  // extern "C"
