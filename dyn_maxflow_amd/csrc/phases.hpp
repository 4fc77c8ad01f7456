// Device side of the MI355X dynamic max-flow engine: state layout, wave64
// helpers, the phase functions of the three push-relabel variants, the
// stepped (one kernel per phase) kernels and the persistent cooperative
// kernel.  Included by engine.hip only.
//
// Work distribution (every phase that walks adjacency lists):
//   * degree <= lane_max          one lane scans the vertex serially;
//   * lane_max < degree <= wide_t one wave64 strides over the adjacency;
//   * degree > wide_t             the vertex goes to a "wide queue" and its
//                                 adjacency is cut into kChunk-arc chunks that
//                                 any wave of the grid picks up, so a hub of
//                                 60k arcs is spread over ~240 waves instead
//                                 of serialising one wave.
// The BFS expands the non-wide part of a frontier kGroup vertices per wave
// task as one flattened arc range (coalesced, balanced to within kGroup*wide_t
// arcs per task).
// A wide queue is a pair of arrays (vertex, first chunk) filled through one
// 64-bit atomic that reserves the slot (high half) and the chunk range (low
// half) together, so first-chunk values are monotone in slot order and a chunk
// finds its vertex by binary search.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>

#include <climits>
#include <cstddef>
#include <cstdint>

#include "../../include/dmf.h"

namespace cg = cooperative_groups;

namespace dmfdev {

constexpr int kBlock = 256;
constexpr int kWave = 64;
// Default of Dev::lane_max (thresholds are tunable per engine).  A vertex of
// degree 5-16 discharged by one lane serially held its wave for the longest
// scan; on RMAT-20 batches lane_max 4 measured learned p50 2.87 vs 3.04 ms at
// 16 and plain 4.45 vs 4.58 ms (2-8 all within 1% of 4; 32 slower).  4 keeps
// grid graphs (degree <= 4) on the lane path.
constexpr int kLaneMax = 4;
// Topology-driven push and repair visit every vertex each round: there the
// lane path keeps degrees up to 16 (wave_lane), which measured faster than 4.
constexpr int kTopoLaneMax = 16;
__host__ __device__ inline int32_t static_wave_t(int32_t lane_max, int32_t wide_t) {
  const int32_t a = lane_max > kTopoLaneMax ? lane_max : kTopoLaneMax;
  return a < wide_t ? a : wide_t;
}
constexpr int kWideT = 64;     // default of Dev::wide_t
constexpr int kChunk = 256;
constexpr int kGroup = 16;      // frontier vertices per wave task in the flattened BFS expansion
constexpr int kSlots = 64;       // work counters are spread over this many addresses
constexpr int kStageCap = 512;   // per-wave LDS staging of frontier appends
constexpr uint16_t kLevFree = 0xFFFF;  // lev16 value of an unlabelled vertex

enum CtrIdx {
  C_ARC = 0,
  C_VERT,
  C_PUSH,
  C_RELABEL,
  C_REPAIR,
  C_LEVELS,
  C_GRELABEL,
  C_ROUNDS,
  C_UPD,
  C_ATOMIC,
  C_BFSARC,   // arc visits of BFS level launches (also counted in C_ARC)
  C_BFSVERT,  // vertex visits of BFS level launches (also counted in C_VERT)
  C_TGT,      // early-exit relabel: targets marked by the BFS init
  C_TGTF,     // early-exit relabel: targets labelled so far
  C_N = 16
};

// Wide queues: 0..2 BFS frontier ring (level % 3), 3 active list, 4 pull list.
constexpr int kNumRing = 3;
constexpr int kQActive = 3;
constexpr int kQPull = 4;
constexpr int kNumQ = 5;

// Device control block: loop counters, flags and work counters.
struct Ctrl {
  int32_t qcnt[3];    // frontier sizes, ring indexed by level % 3
  int32_t wl_cnt;     // active list size
  int32_t wl2_cnt;    // pull list size (push-pull)
  int32_t part_cnt;   // stranded part size (push-pull)
  int32_t flag[2];    // topology variant: some vertex was active (slot = round & 1)
  int32_t status;     // DMF_OK or an error code set on device
  int32_t missing;    // updates naming absent arcs
  int32_t invalid;    // updates with out-of-range ids
  int32_t visited;    // vertices labelled by the last global relabel
  int32_t rounds;     // rounds of the last solve
  int32_t inc;        // updates that raise an arc's capacity (scheduler feature)
  int32_t early;      // the current relabel may stop once its targets are labelled
  int32_t bfs_cut;    // the current relabel stopped before its frontier ran out
  int32_t bfs_rbu;    // the current relabel ran a restricted bottom-up level
  int32_t bfs_inc;    // k_build_wl_after: the relabel's frontier was not empty yet
  int32_t bcur[2][2]; // graph-replayed BFS cursors {first order[] index, level},
                      // read from slot r and written to slot r^1 by each level
  // Round-graph driver: loop state of the data-driven fixed schedule kept on
  // the device (see k_round_*).
  int32_t r_done;      // the solve is over: converged, or out of rounds (r_noconv)
  int32_t r_noconv;
  int32_t r_stop;      // this round's relabel outgrew the captured level launches
  int32_t r_relabel;   // this round starts with a global relabel
  int32_t r_fresh;     // the labels come from this round's relabel
  int32_t r_skip;      // the list came out empty on stale labels: no push this round
  int32_t r_since;     // rounds since the last global relabel
  int32_t r_interval;  // rounds between global relabels
  int32_t r_rounds;    // push rounds so far
  int32_t r_max;       // round cap
  int32_t r_relabels;  // global relabels so far
  int32_t r_levels;    // their BFS levels
  int32_t r_maxlvl;    // the deepest of them
  int32_t r_cnt;       // this round's narrow list size
  unsigned long long wq[kNumQ];  // wide queues: (slots << 32) | chunks
  unsigned long long r_wpk;      // round-graph driver: this round's wide queue word
  long long flow;
};
// k_bfs_level_dev loads the ring counts and wide words as vectors
static_assert(offsetof(Ctrl, qcnt) % 16 == 0 && offsetof(Ctrl, wq) % 16 == 0 && offsetof(Ctrl, bcur) % 8 == 0,
              "control-block fields read as vectors must stay aligned");

struct Dev {
  int32_t n, m, s, t;  // s, t: the first source and sink (feature of the scheduler)
  // terminals: bit 0 source, bit 1 sink.  One instance has one of each; a batch
  // of independent instances stored as one disjoint union has one pair each.
  const uint8_t* term;
  const int32_t* srcs;
  int32_t nsrc;
  const int32_t* snks;
  int32_t nsnk;
  // With more than one source (sink): chunk c of kChunk arcs belongs to
  // terminal i when src_cbase[i] <= c < src_cbase[i+1] (the first nsrc
  // entries are given; src_chunks is the total), so saturation spreads every
  // terminal's arcs over the grid instead of looping over the terminals.
  const int32_t* src_cbase;
  int32_t src_chunks;
  const int32_t* snk_cbase;
  int32_t snk_chunks;
  const int32_t* inst;  // instance of every vertex (nullptr: a single instance)
  long long* iflow;     // per-instance flow
  int32_t ninst;
  int32_t lane_max;     // degree up to which one lane scans a vertex
  int32_t wide_t;       // degree above which a vertex is split into chunks
  int32_t wave_lane;    // the static wave list holds the vertices above this degree
                        // (topology-driven rounds take the rest one per lane)
  const int32_t* row;
  const int32_t* col;
  const int32_t* rev;
  int32_t* cap;
  int32_t* res;
  int32_t* rres;
  long long* excess;
  int32_t* height;
  int32_t* order;
  int32_t* wl;
  int32_t* wl2;
  int32_t* part;
  uint8_t* side;
  uint8_t* pfinal;
  Ctrl* ctrl;
  unsigned long long* ctrs;  // C_N x kSlots work counters
  uint16_t* lev16;           // BFS filter: 0xFFFF unlabelled in the current relabel, else min(level, 0xFFFE)
  int32_t bottom_up;         // bottom-up BFS level once frontier arcs > m / bottom_up (0 = never)
  // Restricted bottom-up levels (relabels from t inside a solve): prev[0,
  // nprev) is the labelled set of the solve's last complete relabel, the only
  // vertices a bottom-up level then tries; one is taken once the frontier's
  // arcs exceed bu_rest times the candidates not yet labelled.  nprev = 0: no
  // restriction (the rule above applies).
  const int32_t* prev;
  int32_t nprev;
  int32_t bu_rest;
  // (tail, head) -> arc index: open-addressing slots {key, arc index in the
  // low 32 bits of y}, so a probe reads key and value in one 16-byte load
  ulonglong2* htab;
  unsigned long long hmask;
  int32_t* owner;
  // wide queues (kNumQ x wq_cap) and the static list of every wide vertex
  int32_t* wq_vert;
  int32_t* wq_base;
  int32_t wq_cap;
  const int32_t* wide_vert;
  const int32_t* wide_base;
  unsigned long long wide_packed;
  // static list of every vertex above lane_max (chunked): the topology-driven
  // push and repair give each its own wave
  const int32_t* wave_vert;
  const int32_t* wave_base;
  unsigned long long wave_packed;
  const uint32_t* widebits;  // bit v set when degree(v) > wide_t
  // Chunk tables of the BFS ring queues (kNumRing x bfs_ct_cap entries, null
  // when not allocated): entry c of ring queue q is chunk c of that queue as
  // {vertex, first arc, end arc, 0}, written when the vertex is appended, so
  // a level finds a chunk's arcs with one load instead of a search over the
  // queue's slots.  A BFS appends every vertex at most once per ring slot,
  // so bfs_ct_cap = the chunks of all vertices above wide_t.
  int4* bfs_ct;
  int32_t bfs_ct_cap;
  // Engine numbering of each caller vertex id, or null when the engine keeps
  // the caller's numbering.  Update records carry caller ids.
  const int32_t* vmap;
  // Early-exit relabel: tgt[v] = 1 for the vertices active when the relabel
  // started (the BFS may stop once all of them are labelled); null when the
  // engine has no such relabel.
  uint8_t* tgt;
  // Cascade push sub-rounds: cmark[v] is the stamp of the last sub-round that
  // listed v (stamps only grow, so no clearing); null when not allocated.
  int32_t* cmark;
  // current update batch
  const int32_t* usrc;
  const int32_t* udst;
  const int32_t* uw;
  int32_t* uarc;
  // per record, written by the lookup where non-null: the reverse arc of the
  // record's arc and its capacity change (the apply and the clamp then read
  // them as streams instead of gathering rev[] and cap[] again)
  int32_t* urev;
  int32_t* udelta;
  int32_t k;
};

// Solve configuration passed to every driver.
struct Cfg {
  int32_t variant;
  int32_t kc;
  int32_t interval;
  int32_t max_rounds;
  int32_t entry;  // 0 = static (init + saturate), 1 = batch (apply + repair)
  int32_t scheduler;  // DMF_SCHED_*
  int32_t lazy;       // batch may start on the labels the previous solve ended with
  int32_t cascade;    // most cascade push sub-rounds per data-driven round (0 = none)
};

constexpr unsigned long long kEmptyKey = ~0ull;

// ------------------------------------------------------------ small helpers
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ bool is_src(const Dev& D, int v) { return D.term[v] & 1; }
__device__ __forceinline__ bool is_snk(const Dev& D, int v) { return D.term[v] & 2; }
__device__ __forceinline__ bool is_term(const Dev& D, int v) { return D.term[v] != 0; }

// A pointer loaded from device memory (rather than taken from the kernel's
// own arguments, which the compiler knows are global) is generic to the
// compiler, and every access through it becomes a flat memory instruction:
// flat loads count against the LDS/scalar wait counter as well, so each LDS or
// scalar wait also drains every gather in flight.  as_global() rebuilds the
// pointer from its address in the global address space; a plain cast there
// and back would be folded away.
template <class T>
__device__ __forceinline__ T* as_global(T* p) {
  return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)p;
}
// Every pointer of a Dev copied from device memory, as global pointers.
__device__ __forceinline__ void globalize(Dev& D) {
  D.term = as_global(D.term);
  D.srcs = as_global(D.srcs);
  D.snks = as_global(D.snks);
  D.src_cbase = as_global(D.src_cbase);
  D.snk_cbase = as_global(D.snk_cbase);
  D.inst = as_global(D.inst);
  D.iflow = as_global(D.iflow);
  D.row = as_global(D.row);
  D.col = as_global(D.col);
  D.rev = as_global(D.rev);
  D.cap = as_global(D.cap);
  D.res = as_global(D.res);
  D.rres = as_global(D.rres);
  D.excess = as_global(D.excess);
  D.height = as_global(D.height);
  D.order = as_global(D.order);
  D.wl = as_global(D.wl);
  D.wl2 = as_global(D.wl2);
  D.part = as_global(D.part);
  D.side = as_global(D.side);
  D.pfinal = as_global(D.pfinal);
  D.ctrl = as_global(D.ctrl);
  D.ctrs = as_global(D.ctrs);
  D.lev16 = as_global(D.lev16);
  D.prev = as_global(D.prev);
  D.htab = as_global(D.htab);
  D.owner = as_global(D.owner);
  D.wq_vert = as_global(D.wq_vert);
  D.wq_base = as_global(D.wq_base);
  D.wide_vert = as_global(D.wide_vert);
  D.wide_base = as_global(D.wide_base);
  D.wave_vert = as_global(D.wave_vert);
  D.wave_base = as_global(D.wave_base);
  D.widebits = as_global(D.widebits);
  D.bfs_ct = as_global(D.bfs_ct);
  D.vmap = as_global(D.vmap);
  D.tgt = as_global(D.tgt);
  D.cmark = as_global(D.cmark);
  D.usrc = as_global(D.usrc);
  D.udst = as_global(D.udst);
  D.uw = as_global(D.uw);
  D.uarc = as_global(D.uarc);
  D.urev = as_global(D.urev);
  D.udelta = as_global(D.udelta);
}

__device__ __forceinline__ int ldr(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long ldr64(const long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ldru64(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void str(int32_t* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long add64(long long* p, long long v) {
  return static_cast<long long>(
      atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v)));
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long y = __shfl_xor(x, o);
    x = y < x ? y : x;
  }
  return x;
}
__device__ __forceinline__ int wave_min_i32(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o));
  return x;
}
__device__ __forceinline__ long long wave_sum_i64(long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}
__device__ __forceinline__ long long wave_incl_scan_i64(long long x) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const long long y = __shfl_up(x, o);
    if (lane_id() >= o) x += y;
  }
  return x;
}

// Reserve slots for the lanes with pred set: one atomic per wave.  Must be
// reached by every active lane of the wave.
__device__ __forceinline__ int wave_reserve(int32_t* counter, bool pred, int* rank) {
  const unsigned long long act = __ballot(1);
  const unsigned long long b = __ballot(pred);
  const int leader = __ffsll(static_cast<long long>(act)) - 1;
  const int total = __popcll(b);
  int base = 0;
  if (lane_id() == leader && total) base = atomicAdd(counter, total);
  base = __shfl(base, leader);
  *rank = __popcll(b & ((1ull << lane_id()) - 1ull));
  return base;
}

// Per-thread work counters.  flush() reduces them over the block and adds the
// sums into one of kSlots copies of each counter: one atomic per block and
// counter, spread over kSlots addresses (same-address atomics serialise).
// Every thread of the block must call flush().
__device__ __forceinline__ void ctr_add(unsigned long long* ctrs, int idx, unsigned long long v) {
  atomicAdd(&ctrs[idx * kSlots + (blockIdx.x % kSlots)], v);
}

struct Tally {
  unsigned long long arc = 0, vert = 0, push = 0, relabel = 0, repair = 0, atom = 0;
  // bfs: the block is a BFS level launch; its arc and vertex visits are also
  // added to the BFS-only counters
  __device__ void flush(unsigned long long* ctrs, bool bfs = false) {
    __shared__ unsigned long long red[6][kBlock / kWave];
    unsigned long long v[6] = {arc, vert, push, relabel, repair, atom};
    const int wv = threadIdx.x / kWave;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      unsigned long long x = v[i];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
      if (lane_id() == 0) red[i][wv] = x;
    }
    __syncthreads();
    if (threadIdx.x < 8 && (threadIdx.x < 6 || bfs)) {
      const int idx[8] = {C_ARC, C_VERT, C_PUSH, C_RELABEL, C_REPAIR, C_ATOMIC, C_BFSARC, C_BFSVERT};
      const int src = threadIdx.x < 6 ? threadIdx.x : threadIdx.x - 6;
      unsigned long long sum = 0;
      for (int w = 0; w < kBlock / kWave; ++w) sum += red[src][w];
      if (sum) ctr_add(ctrs, idx[threadIdx.x], sum);
    }
    __syncthreads();
    arc = vert = push = relabel = repair = atom = 0;
  }
};

// Per-wave staging of list appends in LDS: a wave collects up to kStageCap ids
// and publishes them with one atomic on the shared list counter.  push() and
// flush() must be reached by all 64 lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct Stage {
  int32_t* buf;  // this wave's LDS slice, kStageCap entries
  int cnt;       // wave-uniform fill level
  __device__ void flush(int32_t* gcnt, int32_t* out) {
    if (cnt == 0) return;
    wave_lds_sync();
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(gcnt, cnt);
    base = __shfl(base, 0);
    for (int i = lane_id(); i < cnt; i += kWave) out[base + i] = buf[i];
    wave_lds_sync();
    cnt = 0;
  }
  __device__ void push(bool hit, int v, int32_t* gcnt, int32_t* out) {
    const unsigned long long b = __ballot(hit);
    if (b == 0) return;
    const int rank = __popcll(b & ((1ull << lane_id()) - 1ull));
    if (hit) buf[cnt + rank] = v;
    cnt += __popcll(b);
    if (cnt > kStageCap - kWave) flush(gcnt, out);
  }
  // Final publication by the whole block: one atomic for what all its waves
  // still hold, instead of one per wave (a large level ends with thousands of
  // waves flushing, and same-address atomics serialise).  Every thread of the
  // block must call it.
  __device__ void flush_block(int32_t* gcnt, int32_t* out) {
    __shared__ int32_t wcnt[kBlock / kWave];
    __shared__ int32_t bbase;
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) wcnt[w] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int i = 0; i < kBlock / kWave; ++i) tot += wcnt[i];
      bbase = tot ? atomicAdd(gcnt, tot) : 0;
    }
    __syncthreads();
    int off = bbase;
    for (int i = 0; i < w; ++i) off += wcnt[i];
    for (int i = lane_id(); i < cnt; i += kWave) out[off + i] = buf[i];
    cnt = 0;
  }
};

// Block staging of wide-queue appends: the block's newly labelled wide
// vertices are collected in LDS and published by wave 0 at the end with one
// atomic that reserves their slots and chunks together; chunk bases are an
// exclusive prefix sum in slot order, so base[] stays increasing as
// wave_find_slot requires.  Past kWideStage entries a vertex is appended
// directly (wq_append), which reserves its own slot and chunks in one atomic
// and keeps the same order property.
constexpr int kWideStage = 256;
struct WideStageMem {
  int32_t vert[kWideStage];    // staged vertex
  int32_t chunks[kWideStage];  // its chunk count
  int32_t lo[kWideStage];      // its arcs [lo, hi)
  int32_t hi[kWideStage];
  int32_t base[kWideStage];    // its first chunk in the queue, set by wide_flush
  int32_t cnt;
};
struct WideStage {
  WideStageMem* m;  // LDS
};

__device__ __forceinline__ int wave_incl_scan_i32(int x) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane_id() >= o) x += y;
  }
  return x;
}

__device__ __forceinline__ bool is_block0_thread0() { return blockIdx.x == 0 && threadIdx.x == 0; }

struct Range {
  int gid;  // global thread index
  int gsz;  // total threads
  int wid;  // global wave index
  int nw;   // total waves
};

__device__ __forceinline__ Range make_range() {
  Range r;
  r.gid = blockIdx.x * blockDim.x + threadIdx.x;
  r.gsz = gridDim.x * blockDim.x;
  r.wid = r.gid / kWave;
  r.nw = r.gsz / kWave;
  return r;
}

// ------------------------------------------------------------- wide queues
struct WList {
  const int32_t* vert;
  const int32_t* base;
  int slots;
  int chunks;
  int q;  // queue index, -1 for the static lists
};

__device__ __forceinline__ WList wlist(const Dev& D, int q, unsigned long long packed) {
  WList L;
  L.q = q;
  L.vert = D.wq_vert + static_cast<size_t>(q) * D.wq_cap;
  L.base = D.wq_base + static_cast<size_t>(q) * D.wq_cap;
  L.slots = static_cast<int>(packed >> 32);
  L.chunks = static_cast<int>(packed & 0xffffffffu);
  return L;
}
__device__ __forceinline__ WList wlist_waves(const Dev& D) {
  WList L;
  L.q = -1;
  L.vert = D.wave_vert;
  L.base = D.wave_base;
  L.slots = static_cast<int>(D.wave_packed >> 32);
  L.chunks = static_cast<int>(D.wave_packed & 0xffffffffu);
  return L;
}
__device__ __forceinline__ WList wlist_static(const Dev& D) {
  WList L;
  L.q = -1;
  L.vert = D.wide_vert;
  L.base = D.wide_base;
  L.slots = static_cast<int>(D.wide_packed >> 32);
  L.chunks = static_cast<int>(D.wide_packed & 0xffffffffu);
  return L;
}

__device__ __forceinline__ void wq_append(const Dev& D, int q, int v, int deg) {
  const unsigned chunks = static_cast<unsigned>((deg + kChunk - 1) / kChunk);
  const unsigned long long old = atomicAdd(&D.ctrl->wq[q], (1ull << 32) | chunks);
  const size_t slot = static_cast<size_t>(old >> 32);
  D.wq_vert[static_cast<size_t>(q) * D.wq_cap + slot] = v;
  D.wq_base[static_cast<size_t>(q) * D.wq_cap + slot] = static_cast<int>(old & 0xffffffffu);
}

// Entries [base, base + chunks) of BFS ring queue q's chunk table for vertex
// v with arcs [lo, hi).  Entries past the table (impossible while every
// vertex is appended at most once per ring slot) fail the solve instead of
// being written.  One lane.
__device__ __forceinline__ void ct_write(const Dev& D, int q, int base, int v, int lo, int hi) {
  if (!D.bfs_ct) return;
  const int chunks = (hi - lo + kChunk - 1) / kChunk;
  if (base + chunks > D.bfs_ct_cap) {
    D.ctrl->status = DMF_EDEVICE;
    return;
  }
  int4* ct = D.bfs_ct + static_cast<size_t>(q) * D.bfs_ct_cap;
  for (int k = 0; k < chunks; ++k) {
    const int a = lo + k * kChunk;
    ct[base + k] = make_int4(v, a, min(a + kChunk, hi), 0);
  }
}

// wq_append into BFS ring queue q, with the vertex's chunk-table entries.
// One lane.
__device__ __forceinline__ void bfs_wq_append(const Dev& D, int q, int v, int lo, int hi) {
  const unsigned chunks = static_cast<unsigned>((hi - lo + kChunk - 1) / kChunk);
  const unsigned long long old = atomicAdd(&D.ctrl->wq[q], (1ull << 32) | chunks);
  const size_t slot = static_cast<size_t>(old >> 32);
  const int base = static_cast<int>(old & 0xffffffffu);
  D.wq_vert[static_cast<size_t>(q) * D.wq_cap + slot] = v;
  D.wq_base[static_cast<size_t>(q) * D.wq_cap + slot] = base;
  ct_write(D, q, base, v, lo, hi);
}

// Lanes with pred set append vertex v (degree deg) to wide queue q with one
// atomic for the whole wave: slots by ballot rank, chunk bases by an exclusive
// prefix sum of the lanes' chunk counts (increasing with the slot, as
// wave_find_slot requires).  All 64 lanes.
__device__ __forceinline__ void wave_list_append(unsigned long long* word, int32_t* vert, int32_t* base, bool pred,
                                                 int v, int deg) {
  const unsigned long long b = __ballot(pred);
  if (b == 0) return;
  const int c = pred ? (deg + kChunk - 1) / kChunk : 0;
  const int incl = wave_incl_scan_i32(c);
  const int tot = __shfl(incl, kWave - 1);
  unsigned long long old = 0;
  if (lane_id() == 0)
    old = atomicAdd(word, (static_cast<unsigned long long>(__popcll(b)) << 32) | static_cast<unsigned>(tot));
  old = __shfl(old, 0);
  if (pred) {
    const size_t slot = static_cast<size_t>(old >> 32) + __popcll(b & ((1ull << lane_id()) - 1ull));
    vert[slot] = v;
    base[slot] = static_cast<int>(old & 0xffffffffu) + incl - c;
  }
}
__device__ __forceinline__ void wave_wq_append(const Dev& D, int q, bool pred, int v, int deg) {
  wave_list_append(&D.ctrl->wq[q], D.wq_vert + static_cast<size_t>(q) * D.wq_cap,
                   D.wq_base + static_cast<size_t>(q) * D.wq_cap, pred, v, deg);
}

// Lanes with pred set note wide vertex v into the block stage (past
// kWideStage entries they append it directly).  All 64 lanes.
__device__ __forceinline__ void wide_note(const Dev& D, const WideStage& S, int q, bool pred, int v) {
  const unsigned long long b = __ballot(pred);
  if (b == 0) return;
  const int leader = __ffsll(static_cast<long long>(b)) - 1;
  int base = 0;
  if (lane_id() == leader) base = atomicAdd(&S.m->cnt, __popcll(b));
  base = __shfl(base, leader);
  if (pred) {
    const int lo = D.row[v], hi = D.row[v + 1];
    const int idx = base + __popcll(b & ((1ull << lane_id()) - 1ull));
    if (idx < kWideStage) {
      S.m->vert[idx] = v;
      S.m->chunks[idx] = (hi - lo + kChunk - 1) / kChunk;
      S.m->lo[idx] = lo;
      S.m->hi[idx] = hi;
    } else if (q < kNumRing) {
      bfs_wq_append(D, q, v, lo, hi);
    } else {
      wq_append(D, q, v, hi - lo);
    }
  }
}
// Publishes the block's staged wide vertices into queue q: wave 0 reserves
// their slots and chunks with one atomic and writes the slots; for a BFS ring
// queue the whole block then writes the staged vertices' chunk-table entries,
// each thread finding the vertex of its entries by a binary search over the
// stage's chunk bases.  Every thread of the block must call it; the caller has
// made the stage visible (__syncthreads) since the last wide_note.
__device__ __forceinline__ void wide_flush(const Dev& D, const WideStage& S, int q) {
  const int nv = min(S.m->cnt, kWideStage);
  if (nv == 0) return;  // block-uniform: read from LDS after the caller's barrier
  if (threadIdx.x < kWave) {
    int sum = 0;
    for (int i = lane_id(); i < nv; i += kWave) sum += S.m->chunks[i];
    sum = __shfl(wave_incl_scan_i32(sum), kWave - 1);
    unsigned long long old = 0;
    if (lane_id() == 0)
      old = atomicAdd(&D.ctrl->wq[q], (static_cast<unsigned long long>(nv) << 32) | static_cast<unsigned>(sum));
    old = __shfl(old, 0);
    int slot = static_cast<int>(old >> 32);
    int run = static_cast<int>(old & 0xffffffffu);
    int32_t* vert = D.wq_vert + static_cast<size_t>(q) * D.wq_cap;
    int32_t* base = D.wq_base + static_cast<size_t>(q) * D.wq_cap;
    for (int i0 = 0; i0 < nv; i0 += kWave) {
      const int i = i0 + lane_id();
      const int c = i < nv ? S.m->chunks[i] : 0;
      const int incl = wave_incl_scan_i32(c);
      if (i < nv) {
        vert[slot + i] = S.m->vert[i];
        base[slot + i] = run + incl - c;
        S.m->base[i] = run + incl - c;
      }
      run += __shfl(incl, kWave - 1);
    }
  }
  if (q >= kNumRing || !D.bfs_ct) return;  // uniform
  __syncthreads();
  const int first = S.m->base[0];
  const int total = S.m->base[nv - 1] + S.m->chunks[nv - 1] - first;
  if (first + total > D.bfs_ct_cap) {  // uniform; see ct_write
    if (threadIdx.x == 0) D.ctrl->status = DMF_EDEVICE;
    return;
  }
  int4* ct = D.bfs_ct + static_cast<size_t>(q) * D.bfs_ct_cap;
  for (int e = threadIdx.x; e < total; e += kBlock) {
    int i = 0;  // last staged vertex whose first chunk is at or before entry e
#pragma unroll
    for (int sp = kWideStage / 2; sp > 0; sp >>= 1)
      if (i + sp < nv && S.m->base[i + sp] - first <= e) i += sp;
    const int a = S.m->lo[i] + (e - (S.m->base[i] - first)) * kChunk;
    ct[first + e] = make_int4(S.m->vert[i], a, min(a + kChunk, S.m->hi[i]), 0);
  }
}

// Largest slot a with base[a] <= c, found by the whole wave: each step every
// lane tests one of 64 evenly spaced candidates, so a list of up to 4096 slots
// takes two dependent loads instead of a 12-step scalar binary search.
// Must be called by all 64 lanes with the same c.
__device__ __forceinline__ int wave_find_slot(const int32_t* base, int slots, int c) {
  int lo = 0, span = slots;  // invariant: base[lo] <= c, answer in [lo, lo+span)
  while (span > kWave) {
    const int step = (span + kWave - 1) / kWave;
    const int j = lo + lane_id() * step;
    const bool le = lane_id() * step < span && base[j] <= c;
    const unsigned long long b = __ballot(le);
    const int last = 63 - __clzll(static_cast<long long>(b));
    const int end = lo + span;
    lo += last * step;
    span = min(step, end - lo);
  }
  const bool le = lane_id() < span && base[lo + lane_id()] <= c;
  const unsigned long long b = __ballot(le);
  return lo + (63 - __clzll(static_cast<long long>(b)));
}

// Chunk c of a wide list -> vertex and arc range [lo, hi); *whole is set when
// the chunk is the vertex's entire adjacency.  All 64 lanes.
__device__ __forceinline__ int chunk_range(const Dev& D, const WList& L, int c, int* lo, int* hi,
                                           bool* whole = nullptr) {
  const int a = wave_find_slot(L.base, L.slots, c);
  const int v = L.vert[a];
  const int beg = D.row[v], end = D.row[v + 1];
  *lo = beg + (c - L.base[a]) * kChunk;
  *hi = min(*lo + kChunk, end);
  if (whole) *whole = end - beg <= kChunk;
  return v;
}

// =================================================================== phases

// K1 staticMaxFlow_kernel_1 (code/dyn_data.h:25-35), arc-parallel.
__device__ void ph_init_residual(const Dev& D, const Range& r) {
  for (int e = r.gid; e < D.m; e += r.gsz) {
    D.res[e] = D.cap[e];
    D.rres[e] = D.cap[D.rev[e]];
  }
  for (int v = r.gid; v < D.n; v += r.gsz) D.excess[v] = 0;
}

// K2 staticMaxFlow_kernel_7 (code/dyn_data.h:39-56), arc-parallel over s's
// adjacency instead of one thread.  Each arc e of s has its own reverse arc, so
// only the excess updates need atomics.
__device__ void saturate_one_source(const Dev& D, const Range& r, int s, Tally& t) {
  const int b = D.row[s], deg = D.row[s + 1] - b;
  long long sent = 0;
  for (int i = r.gid - lane_id(); i < deg; i += r.gsz) {  // wave-uniform loop
    const int j = i + lane_id();
    if (j < deg) {
      const int e = b + j;
      const int d = D.res[e];
      if (d > 0) {
        const int w = D.col[e];
        add64(&D.excess[w], d);
        D.res[e] = 0;
        D.rres[e] += d;
        const int p = D.rev[e];
        D.res[p] += d;
        D.rres[p] -= d;
        sent += d;
        t.atom += 1;
      }
      t.arc += 1;
    }
  }
  sent = wave_sum_i64(sent);
  if (lane_id() == 0 && sent) add64(&D.excess[s], -sent);
}

__device__ void ph_saturate_source(const Dev& D, const Range& r, Tally& t) {
  if (D.nsrc == 1 || !D.src_cbase) {
    for (int si = 0; si < D.nsrc; ++si) saturate_one_source(D, r, D.srcs[si], t);
    return;
  }
  // many sources (a batch of instances): one chunk of one source per wave
  for (int c = r.wid; c < D.src_chunks; c += r.nw) {
    const int si = wave_find_slot(D.src_cbase, D.nsrc, c);
    const int s = D.srcs[si];
    const int b = D.row[s], deg = D.row[s + 1] - b;
    const int j0 = (c - D.src_cbase[si]) * kChunk;
    const int j1 = min(deg, j0 + kChunk);
    long long sent = 0;
    for (int j = j0 + lane_id(); j < j1; j += kWave) {
      const int e = b + j;
      const int d = D.res[e];
      if (d > 0) {
        add64(&D.excess[D.col[e]], d);
        D.res[e] = 0;
        D.rres[e] += d;
        const int p = D.rev[e];
        D.res[p] += d;
        D.rres[p] -= d;
        sent += d;
        t.atom += 1;
      }
      t.arc += 1;
    }
    sent = wave_sum_i64(sent);
    if (lane_id() == 0 && sent) add64(&D.excess[s], -sent);
  }
}
// Second half of the push-pull static saturation (code/dyn_pp.h:56-70): pull
// every residual arc into t.
__device__ void saturate_one_sink_pp(const Dev& D, const Range& r, int v, Tally& t) {
  const int b = D.row[v], deg = D.row[v + 1] - b;
  long long got = 0;
  for (int i = r.gid - lane_id(); i < deg; i += r.gsz) {
    const int j = i + lane_id();
    if (j < deg) {
      const int e = b + j;
      const int d = D.rres[e];
      if (d > 0) {
        const int w = D.col[e];
        add64(&D.excess[w], -static_cast<long long>(d));
        D.res[e] += d;
        D.rres[e] -= d;
        const int p = D.rev[e];
        D.res[p] -= d;
        D.rres[p] += d;
        got += d;
        t.atom += 1;
      }
      t.arc += 1;
    }
  }
  got = wave_sum_i64(got);
  if (lane_id() == 0 && got) add64(&D.excess[v], got);
}

__device__ void ph_saturate_sink_pp(const Dev& D, const Range& r, Tally& t) {
  if (D.nsnk == 1 || !D.snk_cbase) {
    for (int si = 0; si < D.nsnk; ++si) saturate_one_sink_pp(D, r, D.snks[si], t);
    return;
  }
  for (int c = r.wid; c < D.snk_chunks; c += r.nw) {
    const int si = wave_find_slot(D.snk_cbase, D.nsnk, c);
    const int v = D.snks[si];
    const int b = D.row[v], deg = D.row[v + 1] - b;
    const int j0 = (c - D.snk_cbase[si]) * kChunk;
    const int j1 = min(deg, j0 + kChunk);
    long long got = 0;
    for (int j = j0 + lane_id(); j < j1; j += kWave) {
      const int e = b + j;
      const int d = D.rres[e];
      if (d > 0) {
        add64(&D.excess[D.col[e]], -static_cast<long long>(d));
        D.res[e] += d;
        D.rres[e] -= d;
        const int p = D.rev[e];
        D.res[p] -= d;
        D.rres[p] += d;
        got += d;
        t.atom += 1;
      }
      t.arc += 1;
    }
    got = wave_sum_i64(got);
    if (lane_id() == 0 && got) add64(&D.excess[v], got);
  }
}

// ----------------------------------------------------------- global relabel
// Root predicates of the three global relabels, plus the two labellings of
// the preflow-to-flow completion: BFS_CT labels the distance from the sinks
// (and from stuck positive excesses) in the forward residual graph, along
// which a deficit pulls flow back; BFS_CS labels the distance to the sources
// in the residual graph, along which stuck excess is returned.  Neither passes
// through a terminal.
// BFS_CUT is the S_min search: reachability over res > 0 from the sources and
// every non-sink vertex with positive excess.
enum BfsMode { BFS_SINK = 0, BFS_PP = 1, BFS_FINAL = 2, BFS_CT = 3, BFS_CS = 4, BFS_CUT = 5 };
constexpr int kBfsModes = 6;

__device__ __forceinline__ int degree(const Dev& D, int v) { return D.row[v + 1] - D.row[v]; }

// my_init_1 (code/dyn_data.h:60-67) / staticMaxFlow_kernel_10 (code/dyn_topo.h:60-71)
// / pp_bfs_init (code/dyn_pp.h:76-83): every vertex unlabelled (height n),
// roots labelled 0 and appended to order[] (wide roots also to wide queue 0).
// ctrl->qcnt[0] and ctrl->wq[0] must be 0.
// early (BFS_SINK only): also mark as targets the vertices that are active
// now (not a terminal, excess > 0, height < n before this relabel) and count
// them.  Roots are staged per wave in LDS and published per block (wide ones
// through the block's wide stage): thousands of deficit roots appended one
// atomic per wave serialised on the frontier counter.  Every thread of the
// block must call it.
__device__ void ph_bfs_init(const Dev& D, const Range& r, int mode, Tally& t, bool early = false) {
  __shared__ int32_t sbuf[kBlock / kWave][kStageCap];
  __shared__ WideStageMem wsm;
  Stage st{sbuf[threadIdx.x / kWave], 0};
  const WideStage WS{&wsm};
  if (threadIdx.x == 0) wsm.cnt = 0;
  __syncthreads();
  const int n = D.n;
  long long targets = 0;
  for (int base = r.gid - lane_id(); base < n; base += r.gsz) {
    const int v = base + lane_id();
    bool root = false;
    if (v < n) {
      const long long ex = D.excess[v];
      if (early) {
        const bool tg = !is_term(D, v) && ex > 0 && D.height[v] < n;
        D.tgt[v] = tg ? 1 : 0;
        targets += tg ? 1 : 0;
      }
      if (mode == BFS_SINK) {
        root = is_snk(D, v) || (!is_src(D, v) && ex < 0);
      } else if (mode == BFS_CT) {
        root = is_snk(D, v) || (!is_term(D, v) && ex > 0);
      } else if (mode == BFS_CS) {
        root = is_src(D, v);
      } else if (mode == BFS_CUT) {
        root = is_src(D, v) || (!is_snk(D, v) && ex > 0);
      } else {
        const bool inB = D.side[v] != 0;
        root = is_term(D, v) || (inB && ex < 0) || (!inB && ex > 0);
      }
      D.height[v] = root ? 0 : n;
      D.lev16[v] = root ? 0 : kLevFree;
      t.vert += 1;
    }
    wide_note(D, WS, 0, root && degree(D, v) > D.wide_t, v);
    st.push(root, v, &D.ctrl->qcnt[0], D.order);
  }
  if (early) {
    targets = wave_sum_i64(targets);
    if (lane_id() == 0 && targets) atomicAdd(&D.ctrs[C_TGT * kSlots + (r.wid % kSlots)], targets);
  }
  st.flush_block(&D.ctrl->qcnt[0], D.order);  // its barriers also publish the wide stage
  wide_flush(D, WS, 0);
}

// Whether every target of the current early-exit relabel is labelled.  All 64
// lanes; the counts of earlier launches are complete at a launch boundary.
__device__ __forceinline__ bool targets_labelled(const Dev& D) {
  const long long a = static_cast<long long>(D.ctrs[C_TGT * kSlots + lane_id()]);
  const long long b = static_cast<long long>(D.ctrs[C_TGTF * kSlots + lane_id()]);
  return wave_sum_i64(b) >= wave_sum_i64(a);
}
// The same decision taken once for the whole block (wave 0 reads the counts),
// so that every wave of a block returns or expands together.  Every thread of
// the block must call it.
__device__ __forceinline__ bool block_targets_labelled(const Dev& D) {
  __shared__ int done;
  if (threadIdx.x < kWave) {
    const bool d = targets_labelled(D);
    if (threadIdx.x == 0) done = d ? 1 : 0;
  }
  __syncthreads();
  return done != 0;
}

// final_bfs_init (code/dyn_pp.h:86-93): roots are the part's deficits.
__device__ void ph_bfs_init_final(const Dev& D, const Range& r, int cnt, Tally& t) {
  for (int base = r.gid - lane_id(); base < cnt; base += r.gsz) {
    const int i = base + lane_id();
    bool root = false;
    int v = -1;
    if (i < cnt) {
      v = D.part[i];
      root = D.excess[v] < 0;
      D.height[v] = root ? 0 : D.n;
      D.lev16[v] = root ? 0 : kLevFree;
      t.vert += 1;
      if (root) {
        const int a = D.row[v], b = D.row[v + 1];
        if (b - a > D.wide_t) bfs_wq_append(D, 0, v, a, b);
      }
    }
    int rank;
    const int off = wave_reserve(&D.ctrl->qcnt[0], root, &rank);
    if (root) D.order[off + rank] = v;
  }
}

// A vertex is unlabelled while its height is n; a level claims it by
// compare-and-swap from n (one winner per vertex).  Claims stay on distinct
// words: a packed one-bit-per-vertex mark array was measured slower, its
// atomics colliding on the words 32 vertices share.  The test before the
// claim reads the 16-bit copy lev16 (half the bytes of height[], so more of it
// stays in cache); the winner writes it.  A stale copy only costs a claim
// that fails.  (The claim itself is issued in ph_bfs_level's step.)
__device__ __forceinline__ bool bfs_unlabelled(const Dev& D, int w) { return D.lev16[w] == kLevFree; }

// Whether a BFS level may label head w through arc e of frontier vertex v is
// split in two, so that the residual can be loaded a tile ahead of its use:
// bfs_resid reads the residual that decides the arc (BFS_SINK and BFS_FINAL:
// the reverse residual; BFS_PP: reverse on the B side, forward on the A side),
// bfs_head_ok checks the head (BFS_PP: never a terminal; BFS_FINAL: a non-B
// vertex of the stranded part).
__device__ __forceinline__ int bfs_resid(const Dev& D, int mode, bool vB, int e) {
  if (mode == BFS_PP) {
    // both residuals are read and the value selected: selecting between the
    // two array fields per lane would force the by-value kernel argument into
    // per-lane scratch memory
    const int32_t back = D.rres[e], fwd = D.res[e];
    return vB ? back : fwd;
  }
  if (mode == BFS_CT || mode == BFS_CUT) return D.res[e];  // mode is uniform over the launch
  return D.rres[e];
}
__device__ __forceinline__ bool bfs_head_ok(const Dev& D, int mode, int w) {
  if (mode == BFS_SINK || mode == BFS_CUT) return true;
  if (mode == BFS_PP || mode == BFS_CT || mode == BFS_CS) return !is_term(D, w);
  return !D.side[w] && D.pfinal[w];
}

// kBfsTiles 64-arc tiles with their heads and residuals loaded: one step of a
// wave's BFS expansion.
constexpr int kBfsTiles = 2;
struct BfsTiles {
  int w[kBfsTiles], r[kBfsTiles];
  bool in[kBfsTiles];
};

// Top-down BFS level (staticMaxFlow_kernel_11, code/dyn_data.h:72-87; pp_bfs,
// code/dyn_pp.h:99-120; final_bfs, code/dyn_pp.h:123-140).  Frontier is
// order[lo, lo+cnt) plus the chunks of wide queue W; newly labelled vertices
// get height lvl+1 (one winner per vertex) and are appended after the frontier
// through per-wave LDS staging, published per block at the end (wide ones
// through the block's wide stage).
//  * Wide frontier vertices: their chunks are spread over every wave of the
//    grid.  A chunk's vertex and arc range come from the ring queue's chunk
//    table with one load, issued a chunk ahead (without a table: a search
//    over the queue's slots, then the vertex's row).
//  * Narrow ones: the 16 a wave takes at a time are expanded as one flattened
//    arc range, so consecutive lanes read consecutive arcs (coalesced
//    col/residual loads) and each lane finds its arc's tail by a 4-step binary
//    search over the group's degree prefix sums.
// A step covers kBfsTiles tiles: the next step's heads and residuals are
// loaded before the current step's labels are tested.  Claims are pipelined
// one step deep: a step's compare-and-swaps (and the wide-vertex bits of their
// heads, read beside them) are issued without waiting, and their results are
// taken after the next step's label reads are out, so a claim's round trip
// overlaps the next step's gathers instead of lengthening the wave's chain of
// dependent loads.  A labelled head's lev16 is written one step later, which
// only lets a few more claims fail.
// early: count the targets this level labels (early-exit relabel).
__device__ void ph_bfs_level(const Dev& D, const Range& r, int mode, int lo, int cnt, int lvl,
                             int32_t* next_cnt, const WList& W, int qnext, Tally& t, bool early = false) {
  __shared__ int32_t sbuf[kBlock / kWave][kStageCap];
  __shared__ WideStageMem wsm;
  Stage st{sbuf[threadIdx.x / kWave], 0};
  const WideStage WS{&wsm};
  if (threadIdx.x == 0) wsm.cnt = 0;
  __syncthreads();
  const int nb = lvl + 1;
  const uint16_t nb16 = static_cast<uint16_t>(nb < kLevFree ? nb : kLevFree - 1);
  int32_t* out = D.order + lo + cnt;
  long long found = 0;
  // claims in flight: head, old height (-1: none), its widebits word, target flag
  int pw[kBfsTiles], pold[kBfsTiles];
  uint32_t pwb[kBfsTiles];
  int ptg[kBfsTiles];
#pragma unroll
  for (int j = 0; j < kBfsTiles; ++j) {
    pw[j] = 0;
    pold[j] = -1;
    pwb[j] = 0;
    ptg[j] = 0;
  }
  // takes the results of the claims in flight (all 64 lanes)
  auto settle = [&]() {
#pragma unroll
    for (int j = 0; j < kBfsTiles; ++j) {
      const bool hit = pold[j] == D.n;
      if (hit) D.lev16[pw[j]] = nb16;
      if (hit && early) found += ptg[j];
      wide_note(D, WS, qnext, hit && ((pwb[j] >> (pw[j] & 31)) & 1u), pw[j]);
      st.push(hit, pw[j], next_cnt, out);
      pold[j] = -1;
    }
  };
  auto expand = [&](const BfsTiles& T) {
    bool c[kBfsTiles];
#pragma unroll
    for (int j = 0; j < kBfsTiles; ++j) c[j] = T.in[j] && T.r[j] > 0 && bfs_head_ok(D, mode, T.w[j]);
#pragma unroll
    for (int j = 0; j < kBfsTiles; ++j) c[j] = c[j] && bfs_unlabelled(D, T.w[j]);
    settle();
#pragma unroll
    for (int j = 0; j < kBfsTiles; ++j) {
      const int w = T.w[j];
      t.arc += T.in[j] ? 1 : 0;
      t.atom += c[j] ? 1 : 0;
      pw[j] = w;
      pold[j] = c[j] ? atomicCAS(&D.height[w], D.n, nb) : -1;
      pwb[j] = c[j] ? D.widebits[static_cast<unsigned>(w) >> 5] : 0u;
      if (early) ptg[j] = c[j] ? D.tgt[w] : 0;
    }
  };
  constexpr int kStep = kBfsTiles * kWave;
  // heads and residuals of arcs [j0, min(j0 + kStep, b)) of wide vertex v
  auto wfetch = [&](int v, int j0, int b) {
    const bool vB = mode == BFS_PP && D.side[v] != 0;
    BfsTiles T;
#pragma unroll
    for (int j = 0; j < kBfsTiles; ++j) {
      const int e = j0 + j * kWave + lane_id();
      T.in[j] = e < b;
      T.w[j] = T.in[j] ? D.col[e] : 0;
      T.r[j] = T.in[j] ? bfs_resid(D, mode, vB, e) : 0;
    }
    return T;
  };
  const int4* ct = (D.bfs_ct && W.q >= 0 && W.q < kNumRing)
                       ? D.bfs_ct + static_cast<size_t>(W.q) * D.bfs_ct_cap : nullptr;
  if (ct) {
    // the wave's chunks c = wid, wid + nw, ... as one stream of steps: the
    // next step's loads are issued before the current step is expanded also
    // across chunk boundaries, and the table entry of the chunk after the
    // next is loaded a chunk ahead
    int c = r.wid;
    if (c < W.chunks) {
      int4 ch = ct[c];
      int4 nx = c + r.nw < W.chunks ? ct[c + r.nw] : make_int4(0, 0, 0, 0);
      int j0 = ch.y;
      BfsTiles cur = wfetch(ch.x, j0, ch.z);
      for (;;) {
        int nj0 = j0 + kStep;
        bool more = true;
        if (nj0 >= ch.z) {
          c += r.nw;
          more = c < W.chunks;
          if (more) {
            ch = nx;
            nj0 = ch.y;
            if (c + r.nw < W.chunks) nx = ct[c + r.nw];
          }
        }
        const BfsTiles now = cur;
        if (more) cur = wfetch(ch.x, nj0, ch.z);
        expand(now);
        if (!more) break;
        j0 = nj0;
      }
    }
  } else {
    for (int c = r.wid; c < W.chunks; c += r.nw) {
      int a, b;
      const int v = chunk_range(D, W, c, &a, &b);
      BfsTiles cur = wfetch(v, a, b);
      for (int j0 = a; j0 < b; j0 += kStep) {
        const BfsTiles now = cur;
        if (j0 + kStep < b) cur = wfetch(v, j0 + kStep, b);
        expand(now);
      }
    }
  }
  // Narrow groups: lane i < kGroup holds the group's i-th vertex.  The
  // groups c = wid, wid + nw, ... form one stream of steps like the wide
  // chunks: the next group's first loads are issued before the current
  // group's last step is expanded, the adjacency bounds of the group after it
  // are loaded one group ahead, and its order[] entries two groups ahead.
  struct Grp {
    int b, incl, excl, total, vB;
  };
  const int gstride = r.nw * kGroup;
  int gbase = r.wid * kGroup;  // base of the group whose bounds are in (cv, cb, ce)
  auto order_at = [&](int g) {
    const int i = g + lane_id();
    return lane_id() < kGroup && i < cnt ? D.order[lo + i] : -1;
  };
  int cv = order_at(gbase);
  int cb = cv >= 0 ? D.row[cv] : 0, ce = cv >= 0 ? D.row[cv + 1] : 0;
  int nv = order_at(gbase + gstride);
  // the next group holding narrow arcs (total 0: none left); all 64 lanes
  auto next_group = [&]() {
    Grp G;
    for (;;) {
      if (gbase >= cnt) {
        G.b = G.incl = G.excl = G.total = G.vB = 0;
        return G;
      }
      const int v = cv, b = cb, e = ce;
      cb = nv >= 0 ? D.row[nv] : 0;
      ce = nv >= 0 ? D.row[nv + 1] : 0;
      cv = nv;
      nv = order_at(gbase + 2 * gstride);
      gbase += gstride;
      int deg = 0;
      G.b = 0;
      G.vB = 0;
      if (v >= 0) {
        deg = e - b;
        if (deg > D.wide_t) deg = 0;  // handled through its chunks
        G.b = b;
        G.vB = (mode == BFS_PP && D.side[v] != 0) ? 1 : 0;
        t.vert += 1;
      }
      G.incl = wave_incl_scan_i32(deg);
      G.total = __shfl(G.incl, kWave - 1);
      G.excl = G.incl - deg;
      if (G.total > 0) return G;
    }
  };
  // heads and residuals of flattened arcs [k0, k0 + kStep) of group G; all
  // 64 lanes (shuffles), k0 wave-uniform
  auto gfetch = [&](const Grp& G, int k0) {
    BfsTiles T;
#pragma unroll
    for (int j = 0; j < kBfsTiles; ++j) {
      const int p = k0 + j * kWave + lane_id();
      int pos = 0;  // owner lane: number of group lanes whose prefix end <= p
#pragma unroll
      for (int sh = kGroup / 2; sh > 0; sh >>= 1)
        if (__shfl(G.incl, pos + sh - 1) <= p) pos += sh;
      pos &= kGroup - 1;
      const int e = __shfl(G.b, pos) + (p - __shfl(G.excl, pos));
      const bool ok = __shfl(G.vB, pos) != 0;
      T.in[j] = p < G.total;
      T.w[j] = T.in[j] ? D.col[e] : 0;
      T.r[j] = T.in[j] ? bfs_resid(D, mode, ok, e) : 0;
    }
    return T;
  };
  {
    Grp G = next_group();
    if (G.total > 0) {
      int k0 = 0;
      BfsTiles cur = gfetch(G, 0);
      for (;;) {
        int nk0 = k0 + kStep;
        bool more = true;
        if (nk0 >= G.total) {
          G = next_group();
          nk0 = 0;
          more = G.total > 0;
        }
        const BfsTiles now = cur;
        if (more) cur = gfetch(G, nk0);
        expand(now);
        if (!more) break;
        k0 = nk0;
      }
    }
  }
  settle();
  if (early) {
    found = wave_sum_i64(found);
    if (lane_id() == 0 && found) atomicAdd(&D.ctrs[C_TGTF * kSlots + (r.wid % kSlots)], found);
  }
  st.flush_block(next_cnt, out);  // its barriers also publish the wide stage
  wide_flush(D, WS, qnext);
}

// Bottom-up BFS level over the reverse residual graph (same levels as the
// top-down rule of staticMaxFlow_kernel_11): every unlabelled candidate w looks
// for one arc w->x with residual > 0 into the current frontier (height ==
// lvl) and stops at the first; no compare-and-swap is needed because w is
// written only by its own lane.  Candidates are every vertex (cand null) or
// cand[0, ncand).  Used for the global relabel from t when the frontier is
// large.  early: count the targets this level labels (early-exit relabel).
__device__ void ph_bfs_bottomup(const Dev& D, const Range& r, int lvl, int32_t* next_cnt, int qnext,
                                int32_t* out, Tally& t, const int32_t* cand = nullptr, int ncand = 0,
                                bool early = false) {
  __shared__ int32_t sbuf[kBlock / kWave][kStageCap];
  Stage st{sbuf[threadIdx.x / kWave], 0};
  const int n = D.n;
  const int32_t* __restrict__ col = D.col;
  const int32_t* __restrict__ res = D.res;
  const int32_t* height = D.height;
  const int total = cand ? ncand : n;
  long long found = 0;
  if (cand && is_block0_thread0()) D.ctrl->bfs_rbu = 1;
  for (int base = r.wid * kWave; base < total; base += r.nw * kWave) {
    const int i = base + lane_id();
    int w = -1;
    bool cand_ok = false;
    int b = 0, deg = 0;
    if (i < total) {
      w = cand ? cand[i] : i;
      if (D.lev16[w] == kLevFree && height[w] == n) {
        b = D.row[w];
        deg = D.row[w + 1] - b;
        cand_ok = deg > 0;
        t.vert += 1;
      }
    }
    bool hit = false;
    // candidates above lane_max: the whole wave scans, two 64-arc tiles per
    // step, and stops at the first tile holding a parent
    unsigned long long big = __ballot(cand_ok && deg > D.lane_max);
    while (big) {
      const int leader = __ffsll(static_cast<long long>(big)) - 1;
      big &= big - 1;
      const int bb = __shfl(b, leader);
      const int bd = __shfl(deg, leader);
      bool fnd = false;
      for (int j0 = 0; j0 < bd && !fnd; j0 += 2 * kWave) {
        const int jA = j0 + lane_id(), jB = jA + kWave;
        const bool inA = jA < bd, inB = jB < bd;
        const int xA = inA ? col[bb + jA] : 0, rA = inA ? res[bb + jA] : 0;
        const int xB = inB ? col[bb + jB] : 0, rB = inB ? res[bb + jB] : 0;
        const int hA = rA > 0 ? height[xA] : -1;
        const int hB = rB > 0 ? height[xB] : -1;
        t.arc += (inA ? 1 : 0) + (inB ? 1 : 0);
        fnd = __ballot(hA == lvl || hB == lvl) != 0;
      }
      if (lane_id() == leader) hit = fnd;
    }
    // narrow candidates: each lane scans its own arcs, four independent
    // (col, res) pairs and height gathers in flight per step
    if (cand_ok && deg <= D.lane_max) {
      const int end = b + deg;
      for (int e = b; e < end && !hit; e += 4) {
        const bool i1 = e + 1 < end, i2 = e + 2 < end, i3 = e + 3 < end;
        const int x0 = col[e], r0 = res[e];
        const int x1 = i1 ? col[e + 1] : 0, r1 = i1 ? res[e + 1] : 0;
        const int x2 = i2 ? col[e + 2] : 0, r2 = i2 ? res[e + 2] : 0;
        const int x3 = i3 ? col[e + 3] : 0, r3 = i3 ? res[e + 3] : 0;
        const int h0 = r0 > 0 ? height[x0] : -1;
        const int h1 = r1 > 0 ? height[x1] : -1;
        const int h2 = r2 > 0 ? height[x2] : -1;
        const int h3 = r3 > 0 ? height[x3] : -1;
        t.arc += 1 + (i1 ? 1 : 0) + (i2 ? 1 : 0) + (i3 ? 1 : 0);
        hit = h0 == lvl || h1 == lvl || h2 == lvl || h3 == lvl;
      }
    }
    if (hit) {
      // w is owned by this thread alone: a plain store labels it
      D.height[w] = lvl + 1;
      D.lev16[w] = static_cast<uint16_t>(lvl + 1 < kLevFree ? lvl + 1 : kLevFree - 1);
      if (deg > D.wide_t) bfs_wq_append(D, qnext, w, b, b + deg);
      if (early) found += D.tgt[w];
    }
    st.push(hit, w, next_cnt, out);
  }
  if (early) {
    found = wave_sum_i64(found);
    if (lane_id() == 0 && found) atomicAdd(&D.ctrs[C_TGTF * kSlots + (r.wid % kSlots)], found);
  }
  st.flush(next_cnt, out);
}

// Direction choice for one level of the relabel from t, after Beamer, Asanovic
// and Patterson, "Direction-optimizing breadth-first search" (SC 2012):
// bottom-up while the frontier's arcs exceed m / D.bottom_up.  The frontier's
// arcs are estimated as its narrow vertices times the mean degree plus its
// wide chunks times kChunk.  R-MAT frontiers are small in vertices but hold
// the hubs' arcs, so the estimate counts arcs, not vertices.
// Restricted (D.nprev > 0, the candidates are the last complete relabel's
// labelled set): bottom-up once the frontier holds at least 1/bu_rest of the
// candidates not yet labelled (labelled so far: lo + cnt, the roots and the
// levels up to the current one) and more arcs than there are such
// candidates.  Past that point most candidates left lie on the next level and
// find a parent in their first few arcs; before it, most of them would scan
// every arc and find none, which costs more than expanding the frontier even
// when hubs make it heavy.  On a residual graph near a maximum flow the
// unrestricted rule loses: the source side, unreachable, is scanned in full
// by every bottom-up level; the restricted candidates leave it out.
__host__ __device__ __forceinline__ bool use_bottom_up(const Dev& D, int mode, int lo, int cnt, long long chunks) {
  if (mode != BFS_SINK) return false;
  const long long mean = D.n > 0 ? D.m / D.n : 0;
  const long long arcs = static_cast<long long>(cnt) * mean + chunks * kChunk;
  if (D.nprev > 0) {
    if (D.bu_rest <= 0) return false;
    const long long left = static_cast<long long>(D.nprev) - lo - cnt;
    return left > 0 && static_cast<long long>(cnt) * D.bu_rest >= left && arcs > left;
  }
  if (D.bottom_up <= 0) return false;
  return arcs * D.bottom_up > D.m;
}


// ---------------------------------------------------------------- worklists
// create_wl (code/dyn_data.h:217-238) over the labelled vertices order[0,cnt):
// excess > 0, v not s or t, height < n.  Narrow and mid-degree vertices go to
// wl, wide ones to wide queue kQActive.
// Vertices above lane_max go to the chunked queues, one wave per chunk: a
// vertex that fits one chunk is discharged by its own wave (kc cycles with
// local relabels), a larger one chunk by chunk.  Those up to lane_max are
// listed for one lane each.
__device__ void ph_build_wl(const Dev& D, const Range& r, int cnt, Tally& t) {
  // cnt < 0: every vertex in id order, one coalesced pass (for a labelled
  // set that is a large part of the graph).  The set listed is the same: a
  // vertex below height n is either labelled by the last relabel or was
  // given a label by k_cap_unlabelled, which only labels vertices without
  // excess (or, after a lazy start, has since received excess downhill and is
  // active with a valid label).  Appends are staged in LDS and published per
  // block; every thread of the block must call it.
  __shared__ int32_t sbuf[kBlock / kWave][kStageCap];
  __shared__ WideStageMem wsm;
  Stage st{sbuf[threadIdx.x / kWave], 0};
  const WideStage WS{&wsm};
  if (threadIdx.x == 0) wsm.cnt = 0;
  __syncthreads();
  const int total = cnt < 0 ? D.n : cnt;
  for (int base = r.gid - lane_id(); base < total; base += r.gsz) {
    const int i = base + lane_id();
    bool act = false, wide = false;
    int v = -1;
    if (i < total) {
      v = cnt < 0 ? i : D.order[i];
      act = !is_term(D, v) && D.excess[v] > 0 && D.height[v] < D.n;
      t.vert += 1;
      if (act) {
        wide = degree(D, v) > D.lane_max;
        act = !wide;
      }
    }
    wide_note(D, WS, kQActive, wide, v);
    st.push(act, v, &D.ctrl->wl_cnt, D.wl);
  }
  st.flush_block(&D.ctrl->wl_cnt, D.wl);  // its barriers also publish the wide stage
  wide_flush(D, WS, kQActive);
}

// create_pp_wl (code/dyn_pp.h:499-545): B vertices with excess > 0 to wl,
// A vertices with excess < 0 to wl2 (those above lane_max to their chunked
// queues).
__device__ void ph_build_pp_wl(const Dev& D, const Range& r, int cnt, Tally& t) {
  for (int base = r.gid - lane_id(); base < cnt; base += r.gsz) {
    const int i = base + lane_id();
    bool a = false, b = false, wa = false, wb = false;
    int v = -1, d = 0;
    if (i < cnt) {
      v = D.order[i];
      if (!is_term(D, v) && D.height[v] < D.n) {
        const long long ex = D.excess[v];
        if (D.side[v]) a = ex > 0;
        else b = ex < 0;
      }
      t.vert += 1;
      if (a || b) {
        d = degree(D, v);
        if (d > D.lane_max) {
          wa = a;
          wb = b;
          a = b = false;
        }
      }
    }
    wave_wq_append(D, kQActive, wa, v, d);
    wave_wq_append(D, kQPull, wb, v, d);
    int rank;
    int off = wave_reserve(&D.ctrl->wl_cnt, a, &rank);
    if (a) D.wl[off + rank] = v;
    off = wave_reserve(&D.ctrl->wl2_cnt, b, &rank);
    if (b) D.wl2[off + rank] = v;
  }
}

// Worklist of the preflow-to-flow completion over the labelled vertices
// order[0,cnt): with deficits set, non-terminal vertices with excess < 0 (to
// wl2 and the pull queue), else those with excess > 0 (to wl and the active
// queue); vertices above lane_max go to the chunked queue.
__device__ void ph_build_imbalance_wl(const Dev& D, const Range& r, int cnt, bool deficits, Tally& t) {
  int32_t* list = deficits ? D.wl2 : D.wl;
  int32_t* lcnt = deficits ? &D.ctrl->wl2_cnt : &D.ctrl->wl_cnt;
  const int q = deficits ? kQPull : kQActive;
  for (int base = r.gid - lane_id(); base < cnt; base += r.gsz) {
    const int i = base + lane_id();
    bool act = false, wide = false;
    int v = -1, d = 0;
    if (i < cnt) {
      v = D.order[i];
      const long long ex = D.excess[v];
      act = !is_term(D, v) && (deficits ? ex < 0 : ex > 0) && D.height[v] < D.n;
      t.vert += 1;
      if (act) {
        d = degree(D, v);
        wide = d > D.lane_max;
        act = !wide;
      }
    }
    wave_wq_append(D, q, wide, v, d);
    int rank;
    const int off = wave_reserve(lcnt, act, &rank);
    if (act) list[off + rank] = v;
  }
}

// Non-terminal vertices whose excess is not zero (after a completion: none).
__device__ void ph_count_imbalance(const Dev& D, const Range& r) {
  int bad = 0;
  for (int v = r.gid; v < D.n; v += r.gsz) bad += (!is_term(D, v) && D.excess[v] != 0) ? 1 : 0;
  for (int o = 32; o > 0; o >>= 1) bad += __shfl_xor(bad, o);
  if (lane_id() == 0 && bad) atomicAdd(&D.ctrl->part_cnt, bad);
}

// create_final_wl (code/dyn_pp.h:619-642) over the part list.
__device__ void ph_build_final_wl(const Dev& D, const Range& r, int cnt, Tally& t) {
  for (int base = r.gid - lane_id(); base < cnt; base += r.gsz) {
    const int i = base + lane_id();
    bool act = false, wide = false;
    int v = -1, d = 0;
    if (i < cnt) {
      v = D.part[i];
      act = D.excess[v] > 0 && D.height[v] < D.n;
      t.vert += 1;
      if (act) {
        d = degree(D, v);
        wide = d > D.lane_max;
        act = !wide;
      }
    }
    wave_wq_append(D, kQActive, wide, v, d);
    int rank;
    const int off = wave_reserve(&D.ctrl->wl_cnt, act, &rank);
    if (act) D.wl[off + rank] = v;
  }
}

// ---------------------------------------------------------- push / pull
// Neighbour filters: all, B only (pp push side), A only (pp pull side), and
// the completion's: never pull from a source, never push into a sink (either
// would change the flow value).
enum PushMode { PM_ALL = 0, PM_ONLY_B = 1, PM_ONLY_A = 2, PM_NO_SRC = 3, PM_NO_SNK = 4 };

__device__ __forceinline__ bool push_filter(const Dev& D, int pmode, int w) {
  if (pmode == PM_ONLY_B) return D.side[w] != 0;
  if (pmode == PM_ONLY_A) return D.side[w] == 0;
  if (pmode == PM_NO_SRC) return !is_src(D, w);
  if (pmode == PM_NO_SNK) return !is_snk(D, w);
  return true;
}

// Cascade push: after a round's push, sub-rounds on the device push again
// from the vertices that gained excess or kept some, without a host round
// trip, a worklist scan or a relabel in between.  Their lists form a ring of
// three slots: slot s is the narrow list cascade_list(D, s) with size
// ctrl->qcnt[s] plus the chunked wide queue s (the BFS ring, idle between
// relabels).  Cascade{slot, stamp}: the slot the current pass appends to and
// its stamp (slot < 0: no cascade).
// relist: 0 = a vertex that keeps excess is not listed again (it waits for
// the next round), 1 = only if it pushed in this pass, 2 = always.
struct Cascade {
  int slot;
  int stamp;
  int relist;
};
__device__ __forceinline__ int32_t* cascade_list(const Dev& D, int slot) {
  return slot == 0 ? D.wl2 : slot == 1 ? D.part : D.wl;
}
// Lanes with cand set note vertex x for the next sub-round, once per stamp.
// Works on whatever lanes are active: the narrow appends of those lanes take
// one atomic, the wide ones reserve their chunks in the slot's queue.
__device__ __forceinline__ void cascade_note(const Dev& D, const Cascade& cz, bool cand, int x) {
  if (cand) cand = atomicMax(&D.cmark[x], cz.stamp) < cz.stamp;
  int deg = 0;
  if (cand) deg = degree(D, x);
  const bool wide = cand && deg > D.lane_max;
  const bool narrow = cand && !wide;
  const unsigned long long b = __ballot(narrow);
  if (b) {
    const int leader = __ffsll(static_cast<long long>(b)) - 1;
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(&D.ctrl->qcnt[cz.slot], __popcll(b));
    base = __shfl(base, leader);
    if (narrow) cascade_list(D, cz.slot)[base + __popcll(b & ((1ull << lane_id()) - 1ull))] = x;
  }
  if (wide) wq_append(D, cz.slot, x, deg);
}
// The same for up to N candidates per lane, with every stamp claim and
// degree load of the N issued before any result is waited for, and one list
// reservation for all of them: a lane that pushed along several arcs in one
// step waits for one round trip per stage, not one per arc.
template <int N>
__device__ __forceinline__ void cascade_note_n(const Dev& D, const Cascade& cz, const bool (&cand)[N],
                                               const int (&x)[N]) {
  int m[N];
#pragma unroll
  for (int j = 0; j < N; ++j) m[j] = cand[j] ? atomicMax(&D.cmark[x[j]], cz.stamp) : cz.stamp;
  int deg[N];
#pragma unroll
  for (int j = 0; j < N; ++j) deg[j] = m[j] < cz.stamp ? degree(D, x[j]) : -1;
  unsigned long long bn[N];
  unsigned long long any = 0;
  int total = 0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    bn[j] = __ballot(deg[j] >= 0 && deg[j] <= D.lane_max);
    any |= bn[j];
    total += __popcll(bn[j]);
  }
  if (any) {
    const int leader = __ffsll(static_cast<long long>(any)) - 1;
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(&D.ctrl->qcnt[cz.slot], total);
    base = __shfl(base, leader);
    int32_t* list = cascade_list(D, cz.slot);
    const unsigned long long below = (1ull << lane_id()) - 1ull;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      if ((bn[j] >> lane_id()) & 1ull) list[base + __popcll(bn[j] & below)] = x[j];
      base += __popcll(bn[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < N; ++j)
    if (deg[j] > D.lane_max) wq_append(D, cz.slot, x[j], deg[j]);
}

// Lane path (degree <= lane_max): up to kc discharge steps.  A step scans the
// arcs once and pushes along every residual arc to a lower neighbour (the arcs
// the reference's step of code/dyn_data.h:94-129 may use, one per scan there)
// until the excess is spent; if excess remains, v relabels to (lowest residual
// neighbour)+1, the reference's relabel rule.  Only this lane lowers excess[v]
// and the residuals of v's arcs, so the excess read at the start bounds the
// step's pushes and is debited with one atomic at the end.
template <bool PULL>
__device__ void lane_push_cycles(const Dev& D, int v, int b, int deg, int kc, int pmode, Tally& t,
                                 const Cascade& cz = Cascade{-1, 0, 0}) {
  const int n = D.n;
  bool pushed = false;
  for (int c = 0; c < kc; ++c) {
    const long long ex = ldr64(&D.excess[v]);
    const int hv = ldr(&D.height[v]);
    const long long budget = PULL ? -ex : ex;
    if (!(budget > 0 && (PULL || hv < n))) break;
    long long moved = 0;
    int mo = INT_MAX;
    const int end = b + deg;
    // four arcs at a time: their heads, residuals and head heights are loaded
    // before any of them is used, so one lane keeps several loads in flight
    for (int e0 = b; e0 < end && moved < budget; e0 += 4) {
      int wv[4], rv[4], hv4[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) wv[j] = e0 + j < end ? D.col[e0 + j] : -1;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool ok = wv[j] >= 0 && push_filter(D, pmode, wv[j]);
        rv[j] = ok ? (PULL ? ldr(&D.rres[e0 + j]) : ldr(&D.res[e0 + j])) : 0;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) hv4[j] = rv[j] > 0 ? D.height[wv[j]] : INT_MAX;
      // cascade: the step's pushes are all issued before their old excesses
      // are read, then the heads that turned active are noted together
      bool got[4] = {false, false, false, false};
      long long old4[4] = {0, 0, 0, 0};
      int d4[4] = {0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int e = e0 + j;
        const int w = wv[j];
        const int resid = rv[j];
        if (resid <= 0 || moved >= budget) continue;
        const int hw = hv4[j];
        if (hw >= hv) {
          mo = min(mo, hw);
          continue;
        }
        const long long left = budget - moved;
        const int d = static_cast<int>(left < resid ? left : resid);
        const int p = D.rev[e];
        if (PULL) {
          atomicSub(&D.rres[e], d);
          atomicAdd(&D.res[e], d);
          atomicAdd(&D.rres[p], d);
          atomicSub(&D.res[p], d);
          add64(&D.excess[w], -static_cast<long long>(d));
        } else {
          atomicSub(&D.res[e], d);
          atomicAdd(&D.rres[e], d);
          atomicAdd(&D.res[p], d);
          atomicSub(&D.rres[p], d);
          old4[j] = add64(&D.excess[w], d);
          d4[j] = d;
          got[j] = true;
        }
        moved += d;
        t.push += 1;
        t.atom += 5;
      }
      if (!PULL && cz.slot >= 0) {
        // heads that turned active: the next cascade sub-round takes them
        bool cand[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) cand[j] = got[j] && old4[j] <= 0 && old4[j] + d4[j] > 0 && !is_term(D, wv[j]);
        cascade_note_n<4>(D, cz, cand, wv);
      }
    }
    t.arc += deg;
    if (moved) {
      add64(&D.excess[v], PULL ? moved : -moved);
      t.atom += 1;
      pushed = true;
    }
    if (moved == budget) continue;  // spent; re-check for refills
    if (mo == INT_MAX) break;       // no residual arc left
    str(&D.height[v], mo + 1);
    t.relabel += 1;
  }
  if (!PULL && cz.slot >= 0) {  // excess left after the cycles: the next sub-round tries again
    const bool again = cz.relist == 2 || (cz.relist == 1 && pushed);
    cascade_note(D, cz, again && ldr64(&D.excess[v]) > 0 && ldr(&D.height[v]) < D.n, v);
  }
}

// Wave discharge of v over arcs [lo, hi): every residual arc to a lower
// neighbour (the only arcs the reference's step may push along) receives flow
// in CSR order until the budget runs out; the budget is claimed from excess[v]
// per 64-arc tile with one atomic, so the chunks of one wide vertex can run in
// different waves at once.  Push moves flow v->w; pull (push-pull's deficit
// side) moves flow w->v.  Returns true when the budget ran out; *min_other
// receives the lowest height among residual arcs that were not lower.
template <bool PULL>
__device__ bool wave_discharge(const Dev& D, int v, int hv, int lo, int hi, int pmode,
                               int* min_other, Tally& t, const Cascade& cz = Cascade{-1, 0, 0}) {
  int mo = INT_MAX;
  bool out = false;
  for (int j0 = lo; j0 < hi; j0 += kWave) {
    const int e = j0 + lane_id();
    long long want = 0;
    int w = -1;
    if (e < hi) {
      w = D.col[e];
      t.arc += 1;
      if (push_filter(D, pmode, w)) {
        const int resid = PULL ? ldr(&D.rres[e]) : ldr(&D.res[e]);
        if (resid > 0) {
          const int hw = D.height[w];
          if (hw < hv) want = resid;
          else mo = min(mo, hw);
        }
      }
    }
    const long long incl = wave_incl_scan_i64(want);
    const long long total = __shfl(incl, kWave - 1);
    if (total == 0) continue;
    long long avail = 0;
    if (lane_id() == 0) {
      // claim up to `total` units of excess (pull: of deficit)
      const long long old = add64(&D.excess[v], PULL ? total : -total);
      const long long have = PULL ? -old : old;
      avail = have <= 0 ? 0 : (have < total ? have : total);
      if (avail < total) add64(&D.excess[v], PULL ? -(total - avail) : (total - avail));
      t.atom += avail < total ? 2 : 1;
    }
    avail = __shfl(avail, 0);
    const long long before = incl - want;
    long long d = avail - before;
    d = d < 0 ? 0 : (d > want ? want : d);
    if (d > 0) {
      const int di = static_cast<int>(d);
      const int p = D.rev[e];
      if (PULL) {
        atomicSub(&D.rres[e], di);
        atomicAdd(&D.res[e], di);
        atomicAdd(&D.rres[p], di);
        atomicSub(&D.res[p], di);
        add64(&D.excess[w], -d);
      } else {
        atomicSub(&D.res[e], di);
        atomicAdd(&D.rres[e], di);
        atomicAdd(&D.res[p], di);
        atomicSub(&D.rres[p], di);
        const long long old = add64(&D.excess[w], d);
        if (cz.slot >= 0) cascade_note(D, cz, old <= 0 && old + d > 0 && !is_term(D, w), w);
      }
      t.push += 1;
      t.atom += 5;
    }
    if (avail < total) {
      out = true;
      break;
    }
  }
  *min_other = wave_min_i32(mo);
  return out;
}

// Wave path for mid-degree vertices: up to kc discharges over the whole
// adjacency, relabelling to (lowest residual neighbour)+1 when a discharge
// leaves excess behind — the reference's relabel rule.
template <bool PULL>
__device__ void wave_push_cycles(const Dev& D, int v, int b, int deg, int kc, int pmode, Tally& t,
                                 const Cascade& cz = Cascade{-1, 0, 0}) {
  const int n = D.n;
  for (int c = 0; c < kc; ++c) {
    long long ex = 0;
    int hv = 0;
    if (lane_id() == 0) {
      ex = ldr64(&D.excess[v]);
      hv = ldr(&D.height[v]);
    }
    ex = __shfl(ex, 0);
    hv = __shfl(hv, 0);
    const long long amt = PULL ? -ex : ex;
    if (!(amt > 0 && (PULL || hv < n))) break;
    int mo;
    const bool spent = wave_discharge<PULL>(D, v, hv, b, b + deg, pmode, &mo, t, cz);
    if (spent) continue;  // re-check: pushes from neighbours may have refilled it
    if (mo == INT_MAX) break;  // no residual arc at all
    if (lane_id() == 0) {
      const long long now = ldr64(&D.excess[v]);
      if ((PULL ? -now : now) > 0) {
        str(&D.height[v], mo + 1);
        t.relabel += 1;
      }
    }
  }
}

// K6 push/relabel (code/dyn_data.h:89-131; pp_push_kernel code/dyn_pp.h:191-232;
// pull_kernel code/dyn_pp.h:237-280; push_kernel code/dyn_pp.h:145-185; topo
// code/dyn_topo.h:90-136).  list == nullptr means every vertex (topology-driven):
// inactive vertices, s and t are skipped and any active vertex sets *flag.
// Wide vertices come through W (one discharge pass over each chunk).
template <bool PULL>
__device__ void ph_push(const Dev& D, const Range& r, const int32_t* list, int cnt, int kc, int pmode,
                        int32_t* flag, const WList& W, Tally& t, const Cascade& cz = Cascade{-1, 0, 0}) {
  // a wide vertex that keeps excess after its wave's work is listed again
  // (lane 0 reads it; every lane of the wave calls cascade_note)
  auto note_self = [&](int v) {
    if (PULL || cz.slot < 0 || cz.relist == 0) return;
    bool c = false;
    if (lane_id() == 0) c = ldr64(&D.excess[v]) > 0 && ldr(&D.height[v]) < D.n;
    cascade_note(D, cz, c, v);
  };
  const int n = D.n;
  // Wide chunks: the excess read here only filters out inactive vertices (the
  // claim in wave_discharge is the exact test) and a wide vertex's height does
  // not change during a push phase, so both are plain, cached loads.  Uncached
  // reads of one hub's words by each of its chunks' waves would serialise at
  // one memory channel.
  for (int c = r.wid; c < W.chunks; c += r.nw) {
    int a, b;
    bool whole;
    const int v = chunk_range(D, W, c, &a, &b, &whole);
    if (whole) {
      // a vertex in one chunk: this wave alone discharges it, with kc cycles
      // and local relabels
      if (!list) {  // topology-driven: skip inactive vertices, report active ones
        long long ex0 = 0;
        int h0 = 0;
        if (lane_id() == 0) {
          ex0 = D.excess[v];
          h0 = D.height[v];
        }
        ex0 = __shfl(ex0, 0);
        h0 = __shfl(h0, 0);
        if (is_term(D, v) || (PULL ? -ex0 : ex0) <= 0 || (!PULL && h0 >= n)) continue;
        if (lane_id() == 0 && *flag == 0) *flag = 1;
      }
      wave_push_cycles<PULL>(D, v, a, b - a, kc, pmode, t, cz);
      note_self(v);
      continue;
    }
    long long ex = 0;
    int hv = 0;
    if (lane_id() == 0) {
      ex = D.excess[v];
      hv = D.height[v];
    }
    ex = __shfl(ex, 0);
    hv = __shfl(hv, 0);
    const long long amt = PULL ? -ex : ex;
    bool act = amt > 0 && (PULL || hv < n);
    if (!list) act = act && !is_term(D, v);
    if (!act) continue;
    if (!list && lane_id() == 0 && *flag == 0) *flag = 1;
    int mo;
    wave_discharge<PULL>(D, v, hv, a, b, pmode, &mo, t, cz);
    note_self(v);
  }
  for (int base = r.wid * kWave; base < cnt; base += r.nw * kWave) {
    const int i = base + lane_id();
    int v = -1, b = 0, deg = 0;
    if (i < cnt) {
      v = list ? list[i] : i;
      b = D.row[v];
      deg = D.row[v + 1] - b;
      if (!list) {  // vertices above lane_max come through W
        const bool act = deg <= D.lane_max && !is_term(D, v) && D.excess[v] > 0 && D.height[v] < n;
        if (!act) v = -1;
        else if (*flag == 0) *flag = 1;
      }
      if (v >= 0) t.vert += 1;
    }
    if (v >= 0) lane_push_cycles<PULL>(D, v, b, deg, kc, pmode, t, cz);
  }
}

// --------------------------------------------------------------- repair
// Label repair of arc e = v->w (v listed, h = h(v)), as in
// staticMaxFlow_kernel_17 (code/dyn_data.h:140-150): residual w->v exists but
// h(w) > h(v)+1, so saturate w->v; with the pull rule, remove_pull
// (code/dyn_pp.h:381-403): residual v->w exists but h(w) > h(v)+1, so
// saturate v->w.  Heights are not written during this phase, so at most one
// endpoint of an arc pair acts.
// Repair filters: all neighbours, B neighbours (pp_remove_push), A neighbours
// with the pull rule (remove_pull), part neighbours (remove_push_final).
enum RepairMode { RM_ALL = 0, RM_B = 1, RM_PULL_A = 2, RM_PFINAL = 3 };

// Repair of arc e = v->w whose residual d > 0 was read and whose head is too
// high.
__device__ __forceinline__ void repair_apply(const Dev& D, int rmode, int v, int e, int w, int d, Tally& t) {
  const int p = D.rev[e];
  if (rmode == RM_PULL_A) {
    D.rres[e] += d;
    D.res[e] = 0;
    D.rres[p] -= d;
    D.res[p] += d;
    add64(&D.excess[v], -static_cast<long long>(d));
    add64(&D.excess[w], d);
  } else {
    D.rres[e] = 0;
    D.res[e] += d;
    D.rres[p] += d;
    D.res[p] -= d;
    add64(&D.excess[w], -static_cast<long long>(d));
    add64(&D.excess[v], d);
  }
  t.repair += 1;
  t.atom += 2;
}

// The residual a repair of arc e = v->w looks at (0 when the mode's filter
// excludes w).
__device__ __forceinline__ int repair_resid(const Dev& D, int rmode, int e, int w) {
  if (w < 0) return 0;
  if (rmode == RM_B && !D.side[w]) return 0;
  if (rmode == RM_PULL_A && D.side[w]) return 0;
  if (rmode == RM_PFINAL && !D.pfinal[w]) return 0;
  return rmode == RM_PULL_A ? D.res[e] : D.rres[e];
}

// Lane path of the repair: four arcs at a time, heads, residuals and head
// heights loaded before any of them is tested (one lane keeps several loads in
// flight); each arc is then tested and repaired.
__device__ __forceinline__ void repair_lane(const Dev& D, int rmode, int v, int h, int b, int end, Tally& t) {
  for (int e0 = b; e0 < end; e0 += 4) {
    int wv[4], dv[4], hw[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) wv[j] = e0 + j < end ? D.col[e0 + j] : -1;
#pragma unroll
    for (int j = 0; j < 4; ++j) dv[j] = repair_resid(D, rmode, e0 + j, wv[j]);
#pragma unroll
    for (int j = 0; j < 4; ++j) hw[j] = dv[j] > 0 ? D.height[wv[j]] : 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (wv[j] >= 0) t.arc += 1;
      if (dv[j] > 0 && hw[j] > h + 1) repair_apply(D, rmode, v, e0 + j, wv[j], dv[j], t);
    }
  }
}

// Wave path of the repair over arcs [lo, hi) of v: kRepTiles 64-arc tiles per
// step, all their heads, residuals and head heights loaded before any is
// tested.
constexpr int kRepTiles = 4;
__device__ __forceinline__ void repair_wave(const Dev& D, int rmode, int v, int h, int lo, int hi, Tally& t) {
  for (int j0 = lo; j0 < hi; j0 += kRepTiles * kWave) {
    int wv[kRepTiles], dv[kRepTiles], hw[kRepTiles];
#pragma unroll
    for (int j = 0; j < kRepTiles; ++j) {
      const int e = j0 + j * kWave + lane_id();
      wv[j] = e < hi ? D.col[e] : -1;
    }
#pragma unroll
    for (int j = 0; j < kRepTiles; ++j) dv[j] = repair_resid(D, rmode, j0 + j * kWave + lane_id(), wv[j]);
#pragma unroll
    for (int j = 0; j < kRepTiles; ++j) hw[j] = dv[j] > 0 ? D.height[wv[j]] : 0;
#pragma unroll
    for (int j = 0; j < kRepTiles; ++j) {
      if (wv[j] >= 0) t.arc += 1;
      if (dv[j] > 0 && hw[j] > h + 1) repair_apply(D, rmode, v, j0 + j * kWave + lane_id(), wv[j], dv[j], t);
    }
  }
}

__device__ void ph_repair(const Dev& D, const Range& r, const int32_t* list, int cnt, int rmode,
                          const WList& W, Tally& t) {
  for (int c = r.wid; c < W.chunks; c += r.nw) {
    int a, b;
    const int v = chunk_range(D, W, c, &a, &b);
    const int h = D.height[v];
    repair_wave(D, rmode, v, h, a, b, t);
  }
  for (int base = r.wid * kWave; base < cnt; base += r.nw * kWave) {
    const int i = base + lane_id();
    int v = -1, b = 0, deg = 0, h = 0;
    if (i < cnt) {
      v = list ? list[i] : i;
      b = D.row[v];
      deg = D.row[v + 1] - b;
      if (deg > D.lane_max) v = -1;  // through W, one wave per chunk
      else {
        h = D.height[v];
        t.vert += 1;
      }
    }
    if (v >= 0) repair_lane(D, rmode, v, h, b, b + deg, t);
  }
}

// ------------------------------------------------------- push-pull phases
// pp incremental_kernel_20 (code/dyn_pp.h:405-436): clamp, saturate the A->B
// direction of arcs crossing the stored cut, recompute excess = net inflow.
// Wide vertices accumulate their chunks' sums atomically; their excess was
// zeroed in the preceding phase (ph_upd_apply with zero_wide).
__device__ __forceinline__ long long pp_arc_fix(const Dev& D, int e, uint8_t sv) {
  int rs = D.res[e], rr = D.rres[e];
  if (rs < 0) {
    rr += rs;
    rs = 0;
  }
  if (rr < 0) {
    rs += rr;
    rr = 0;
  }
  if (sv != D.side[D.col[e]]) {
    if (sv) {
      rs += rr;
      rr = 0;
    } else {
      rr += rs;
      rs = 0;
    }
  }
  D.res[e] = rs;
  D.rres[e] = rr;
  return static_cast<long long>(rs) - D.cap[e];
}

__device__ void ph_pp_cut(const Dev& D, const Range& r, Tally& t) {
  const int n = D.n;
  const WList W = wlist_static(D);
  for (int c = r.wid; c < W.chunks; c += r.nw) {
    int a, b;
    const int v = chunk_range(D, W, c, &a, &b);
    const uint8_t sv = D.side[v];
    long long acc = 0;
    for (int e = a + lane_id(); e < b; e += kWave) {
      acc += pp_arc_fix(D, e, sv);
      t.arc += 1;
    }
    acc = wave_sum_i64(acc);
    if (lane_id() == 0 && acc) add64(&D.excess[v], acc);
  }
  for (int base = r.wid * kWave; base < n; base += r.nw * kWave) {
    const int v = base + lane_id();
    int b = 0, deg = 0;
    uint8_t self = 0;
    bool mine = false;
    if (v < n) {
      b = D.row[v];
      deg = D.row[v + 1] - b;
      self = D.side[v];
      mine = deg <= D.wide_t;
      t.vert += 1;
    }
    unsigned long long big = __ballot(mine && deg > D.lane_max);
    while (big) {
      const int leader = __ffsll(static_cast<long long>(big)) - 1;
      big &= big - 1;
      const int bv = __shfl(v, leader);
      const int bb = __shfl(b, leader);
      const int bd = __shfl(deg, leader);
      const uint8_t bs = static_cast<uint8_t>(__shfl(static_cast<int>(self), leader));
      long long acc = 0;
      for (int e = bb + lane_id(); e < bb + bd; e += kWave) {
        acc += pp_arc_fix(D, e, bs);
        t.arc += 1;
      }
      acc = wave_sum_i64(acc);
      if (lane_id() == 0) D.excess[bv] = acc;
    }
    if (mine && deg <= D.lane_max) {
      long long acc = 0;
      for (int e = b; e < b + deg; ++e) acc += pp_arc_fix(D, e, self);
      t.arc += deg;
      D.excess[v] = acc;
    }
  }
}

// find_cut (code/dyn_pp.h:440-449): B = labelled by the last relabel.
__device__ void ph_find_cut(const Dev& D, const Range& r) {
  for (int v = r.gid; v < D.n; v += r.gsz) D.side[v] = D.height[v] < D.n ? 1 : 0;
}

// find_part (code/dyn_pp.h:330-356): vertices the last two-sided relabel left
// unlabelled move to A and form the part list; every height is reset.
__device__ void ph_find_part(const Dev& D, const Range& r, Tally& t) {
  const int n = D.n;
  for (int base = r.gid - lane_id(); base < n; base += r.gsz) {
    const int v = base + lane_id();
    bool in = false;
    if (v < n) {
      in = D.height[v] == n;
      D.pfinal[v] = in ? 1 : 0;
      if (in) D.side[v] = 0;
      D.height[v] = n;
      t.vert += 1;
    }
    int rank;
    const int off = wave_reserve(&D.ctrl->part_cnt, in, &rank);
    if (in) D.part[off + rank] = v;
  }
}

// find_final_cut (code/dyn_pp.h:452-460).
__device__ void ph_find_final_cut(const Dev& D, const Range& r) {
  for (int v = r.gid; v < D.n; v += r.gsz)
    if (!D.side[v] && D.height[v] < D.n) D.side[v] = 1;
}

// recalculate_max_flow_kernel_39 (code/dyn_data.h:208-215) as a reduction that
// leaves excess untouched: flow = excess[t] + sum of excess[v] < 0, v != s.
__device__ void ph_flow(const Dev& D, const Range& r) {
  long long acc = 0;
  for (int v = r.gid; v < D.n; v += r.gsz) {
    const long long ex = D.excess[v];
    long long c = 0;
    if (is_snk(D, v)) c += ex;
    if (ex < 0 && !is_src(D, v)) c += ex;
    if (c && D.inst) add64(&D.iflow[D.inst[v]], c);
    acc += c;
  }
  acc = wave_sum_i64(acc);
  if (lane_id() == 0 && acc) add64(&D.ctrl->flow, acc);
}

// ---------------------------------------------------------- update phases
__device__ __forceinline__ unsigned long long hash_key(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}
__device__ __forceinline__ unsigned long long pair_key(int u, int v) {
  return (static_cast<unsigned long long>(static_cast<unsigned>(u)) << 32) |
         static_cast<unsigned>(v);
}

// Arc (a, b) in the caller's vertex numbering.  The hash index is keyed by
// caller ids (k_build_hash maps every arc's ends back through the inverse of
// vmap), so a probe needs no vmap gathers.
__device__ int lookup_arc(const Dev& D, int a, int b) {
  if (!D.htab) {
    // no arc index (the allocation-free solve): scan the tail's arcs, the
    // last match wins as in OnAdd_kernel's scan (code/dyn_data.h:191-201)
    const int u = D.vmap ? D.vmap[a] : a;
    const int v = D.vmap ? D.vmap[b] : b;
    int found = -1;
    for (int e = D.row[u]; e < D.row[u + 1]; ++e)
      if (D.col[e] == v) found = e;
    return found;
  }
  const unsigned long long key = pair_key(a, b);
  unsigned long long h = hash_key(key) & D.hmask;
  for (;;) {
    const ulonglong2 slot = D.htab[h];
    if (slot.x == key) return static_cast<int>(static_cast<unsigned>(slot.y));
    if (slot.x == kEmptyKey) return -1;
    h = (h + 1) & D.hmask;
  }
}

// OnAdd_kernel arc search (code/dyn_data.h:191-201) by hash lookup, plus
// last-record-wins resolution of duplicate arcs within the batch.
__device__ void ph_upd_lookup(const Dev& D, const Range& r) {
  int missing = 0, invalid = 0, inc = 0;
  for (int i = r.gid; i < D.k; i += r.gsz) {
    const int a = D.usrc[i], b = D.udst[i];
    int arc = -2;
    if (a < 0 || a >= D.n || b < 0 || b >= D.n || D.uw[i] < 0) {  // capacities are >= 0
      ++invalid;
    } else if (a != b) {  // self-loops skipped (code/update.hpp:82, code/dyn_data.h:194)
      arc = lookup_arc(D, a, b);
      if (arc < 0) {
        ++missing;
      } else {
        atomicMax(&D.owner[arc], i);
        const int c0 = D.cap[arc];  // the apply changes it only after this kernel
        inc += D.uw[i] > c0;
        if (D.urev) {
          D.urev[i] = D.rev[arc];
          D.udelta[i] = D.uw[i] - c0;
        }
      }
    }
    D.uarc[i] = arc;
  }
  if (missing) atomicAdd(&D.ctrl->missing, missing);
  if (invalid) atomicAdd(&D.ctrl->invalid, invalid);
  if (inc) atomicAdd(&D.ctrl->inc, inc);
}

// OnAdd_kernel update (code/dyn_data.h:202-205) for the winning record of each
// arc: residual and its mirror move with the capacity.  capacity_only mirrors
// update_kernel (code/dyn_data_bench.cu:300-322).  zero_wide clears the excess
// of every wide vertex for the push-pull recompute that follows.
// The winning record also releases its arc (owner back to -1) and every
// record that lost its arc is marked in uarc (-1), so the clamp that follows
// needs no owner reads and no release pass runs after a successful apply.  A
// losing record reads owner as the winner's index or, once released, -1:
// neither equals its own index.
__device__ void ph_upd_apply(const Dev& D, const Range& r, bool capacity_only, bool zero_wide) {
  if (D.ctrl->missing || D.ctrl->invalid) return;
  for (int i = r.gid; i < D.k; i += r.gsz) {
    const int arc = D.uarc[i];
    if (arc < 0) continue;
    if (D.owner[arc] != i) {
      D.uarc[i] = -1;
      continue;
    }
    const int nw = D.uw[i];
    if (!capacity_only) {
      const int delta = D.udelta ? D.udelta[i] : nw - D.cap[arc];
      D.res[arc] += delta;
      D.rres[D.urev ? D.urev[i] : D.rev[arc]] += delta;
    }
    D.cap[arc] = nw;
    D.owner[arc] = -1;
  }
  if (zero_wide) {
    const int slots = static_cast<int>(D.wide_packed >> 32);
    for (int i = r.gid; i < slots; i += r.gsz) D.excess[D.wide_vert[i]] = 0;
  }
}

// Decrement repair restricted to the updated arcs (the only ones whose
// residual can be negative): incremental_kernel_20 (code/dyn_data.h:157-172)
// cancels the flow above the new capacity, and the net-inflow recompute of
// incremental_kernel_22 (code/dyn_data.h:173-183) reduces to moving that
// amount between the two endpoint excesses.
__device__ void ph_upd_clamp(const Dev& D, const Range& r, Tally& t) {
  if (D.ctrl->missing || D.ctrl->invalid) return;
  for (int i = r.gid; i < D.k; i += r.gsz) {
    const int arc = D.uarc[i];
    if (arc < 0) continue;  // the apply marked every record that lost its arc
    const int rs = D.res[arc];
    t.arc += 1;
    if (rs < 0) {
      const int a = -rs;
      const int p = D.urev ? D.urev[i] : D.rev[arc];
      D.res[arc] = 0;
      D.rres[p] = 0;
      D.rres[arc] -= a;
      D.res[p] -= a;
      add64(&D.excess[D.col[p]], a);  // the arc's tail
      add64(&D.excess[D.col[arc]], -static_cast<long long>(a));
      t.atom += 2;
    }
  }
}

__device__ void ph_upd_release(const Dev& D, const Range& r) {
  for (int i = r.gid; i < D.k; i += r.gsz) {
    const int arc = D.uarc[i];
    if (arc >= 0) D.owner[arc] = -1;
  }
}

// ============================================================ stepped kernels
// Control-block clears done by kernels rather than hipMemsetAsync: each is one
// small launch in place of several runtime fill dispatches.
// k_begin: the whole control block and the work counters, before a solve.
__global__ void __launch_bounds__(kBlock) k_begin(Dev D) {
  int32_t* c = reinterpret_cast<int32_t*>(D.ctrl);
  for (int i = threadIdx.x; i < static_cast<int>(sizeof(Ctrl) / sizeof(int32_t)); i += kBlock) c[i] = 0;
  for (int i = threadIdx.x; i < C_N * kSlots; i += kBlock) D.ctrs[i] = 0;
}
// k_relabel_prep: the frontier ring, the BFS wide queues and the replay
// cursor, before a global relabel; with out set, also the device copy of Dev
// that graph-replayed levels read.
// early: the relabel may stop once every target is labelled (its target
// counters are zeroed here).  Launched with one wave.
__global__ void k_relabel_prep(Dev D, Dev* out, int early) {
  if (threadIdx.x == 0) {
    Ctrl* c = D.ctrl;
    c->qcnt[0] = c->qcnt[1] = c->qcnt[2] = 0;
    c->wq[0] = c->wq[1] = c->wq[2] = 0;
    c->bcur[0][0] = c->bcur[0][1] = c->bcur[1][0] = c->bcur[1][1] = 0;
    c->early = early;
    c->bfs_cut = 0;
    c->bfs_rbu = 0;
    if (out) *out = D;
  }
  if (threadIdx.x < kSlots) {
    D.ctrs[C_TGT * kSlots + threadIdx.x] = 0;
    D.ctrs[C_TGTF * kSlots + threadIdx.x] = 0;
  }
}
// After an early-exit relabel whose deepest label is cap: vertices it left
// unlabelled and that hold no excess get height cap, so that the labels stay
// valid (h(w) <= h(v) + 1 for every residual arc w->v; unlabelled vertices lie
// at least one level beyond the last level that was expanded in full) and
// the label repair does not take them for violations.  Unlabelled vertices
// with excess keep height n until a full relabel decides whether they can
// still reach a sink.
__global__ void __launch_bounds__(kBlock) k_cap_unlabelled(Dev D, int cap) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < D.n; v += gridDim.x * blockDim.x)
    if (D.height[v] == D.n && D.excess[v] <= 0) D.height[v] = cap;
}
// Worklist behind a relabel's first replayed segment, without a host check in
// between: each kernel finds from the level cursor (slot 0 at a segment's end)
// whether the BFS is over, i.e. the frontier there is empty, and does nothing
// while it is not; k_build_wl_after records the answer (bfs_inc) and the
// labelled count, and the host, seeing bfs_inc, runs further segments and
// launches the two again.  The ring counts are loaded beside the cursor (one
// round trip).  k_build_wl_after's block 0 clears the ring once the BFS is
// over, which leaves the count at the cursor zero, so every block agrees.
__device__ __forceinline__ bool bfs_over(const Ctrl* c, int* labelled) {
  const int2 cur = *reinterpret_cast<const int2*>(&c->bcur[0][0]);
  const int4 qc = *reinterpret_cast<const int4*>(&c->qcnt[0]);
  const int slot = cur.y % 3;
  *labelled = cur.x;
  return (slot == 0 ? qc.x : (slot == 1 ? qc.y : qc.z)) == 0;
}
// k_cap_unlabelled after an early-exit relabel that stopped short (bfs_cut),
// with the cap taken from the level cursor.
__global__ void __launch_bounds__(kBlock) k_cap_after(Dev D) {
  const Ctrl* c = D.ctrl;
  int labelled;
  if (!bfs_over(c, &labelled) || !c->bfs_cut) return;
  const int cap = max(0, c->bcur[0][1] - 1);
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < D.n; v += gridDim.x * blockDim.x)
    if (D.height[v] == D.n && D.excess[v] <= 0) D.height[v] = cap;
}
// k_build_wl over the labelled count (every vertex when that is a quarter of
// the graph or more).
__global__ void __launch_bounds__(kBlock) k_build_wl_after(Dev D) {
  int cnt;
  const bool over = bfs_over(D.ctrl, &cnt);  // block-uniform
  if (is_block0_thread0()) {
    Ctrl* c = D.ctrl;
    c->bfs_inc = over ? 0 : 1;
    if (over) {
      c->visited = cnt;
      c->qcnt[0] = c->qcnt[1] = c->qcnt[2] = 0;
      c->wq[0] = c->wq[1] = c->wq[2] = 0;
    }
  }
  if (!over) return;
  Tally t;
  ph_build_wl(D, make_range(), cnt > D.n / 4 ? -1 : cnt, t);
  t.flush(D.ctrs);
}
// k_list_snap: the round's list sizes, for push, repair and cascade launches
// queued behind a relabel before the host has read them.  The round is
// skipped (r_skip) when the BFS is not over (bfs_inc: the list was not built)
// or the list is empty; otherwise the list counters are cleared for the next
// round's builder, as a host-sized push launch does.  One wave.
__global__ void k_list_snap(Dev D) {
  if (threadIdx.x != 0) return;
  Ctrl* c = D.ctrl;
  const int inc = c->bfs_inc;
  const int cnt = c->wl_cnt;
  const unsigned long long wpk = c->wq[kQActive];
  const bool skip = inc || (cnt == 0 && wpk == 0);
  c->r_skip = skip ? 1 : 0;
  c->r_cnt = inc ? 0 : cnt;
  c->r_wpk = inc ? 0ull : wpk;
  if (!inc) {
    c->wl_cnt = c->wl2_cnt = 0;
    c->wq[kQActive] = c->wq[kQPull] = 0;
  }
}
// k_flow_prep: list counters, flags, the flow and the per-instance flows,
// before the flow reduction.
__global__ void __launch_bounds__(kBlock) k_flow_prep(Dev D) {
  Ctrl* c = D.ctrl;
  if (threadIdx.x == 0) {
    c->wl_cnt = c->wl2_cnt = c->part_cnt = 0;
    c->wq[kQActive] = c->wq[kQPull] = 0;
    c->flag[0] = c->flag[1] = 0;
    c->flow = 0;
  }
  for (int i = threadIdx.x; i < D.ninst; i += kBlock) D.iflow[i] = 0;
}

__global__ void __launch_bounds__(kBlock) k_init_residual(Dev D) { ph_init_residual(D, make_range()); }
__global__ void __launch_bounds__(kBlock) k_saturate_source(Dev D) {
  Tally t;
  ph_saturate_source(D, make_range(), t);
  t.flush(D.ctrs);
}
__global__ void __launch_bounds__(kBlock) k_saturate_sink_pp(Dev D) {
  Tally t;
  ph_saturate_sink_pp(D, make_range(), t);
  t.flush(D.ctrs);
}
__global__ void __launch_bounds__(kBlock) k_bfs_init(Dev D, int mode, int early) {
  Tally t;
  ph_bfs_init(D, make_range(), mode, t, early != 0);
  t.flush(D.ctrs);
}
__global__ void __launch_bounds__(kBlock) k_bfs_init_final(Dev D, int cnt) {
  Tally t;
  ph_bfs_init_final(D, make_range(), cnt, t);
  t.flush(D.ctrs);
}
// One level of the stepped relabel.  Block 0 also clears ring slot
// (lvl+2)%3: it held the size of level lvl-1, which the host has read, and
// level lvl+1 appends to it.
__global__ void __launch_bounds__(kBlock) k_bfs_level(Dev D, int mode, int lo, int cnt, int lvl,
                                                      unsigned long long wpk, int bottomup) {
  Tally t;
  const Range r = make_range();
  if (is_block0_thread0()) {
    D.ctrl->qcnt[(lvl + 2) % 3] = 0;
    D.ctrl->wq[(lvl + 2) % 3] = 0;
  }
  // early-exit relabel whose targets are all labelled: the level is not
  // expanded, nothing is appended, and the host sees an empty next level
  const bool early = D.ctrl->early != 0;
  if (early && block_targets_labelled(D)) {
    if (threadIdx.x == 0) D.ctrl->bfs_cut = 1;
    return;
  }
  if (bottomup) {
    ph_bfs_bottomup(D, r, lvl, &D.ctrl->qcnt[(lvl + 1) % 3], (lvl + 1) % 3, D.order + lo + cnt, t,
                    D.nprev > 0 ? D.prev : nullptr, D.nprev, early);
  } else {
    const WList W = wlist(D, lvl % 3, wpk);
    ph_bfs_level(D, r, mode, lo, cnt, lvl, &D.ctrl->qcnt[(lvl + 1) % 3], W, (lvl + 1) % 3, t, early);
  }
  t.flush(D.ctrs, true);
}
// Graph-replayed BFS levels.  A captured segment is a fixed chain of
// k_bfs_level_dev launches whose cursor slot alternates 0,1,0,1,...; the
// frontier of the current level is found from the device cursor rather than
// from kernel arguments, and once a level comes out empty every later launch
// of the segment returns at once.  The engine state is read through a device
// copy of Dev (written by k_relabel_prep) because captured kernel arguments
// are frozen at capture time.
// Reads cursor slot r.  Block 0 writes the next level's cursor into slot r^1,
// which no block of this launch reads, and clears ring slot (lvl+2)%3 (the
// size of level lvl-1, already folded into the cursor; level lvl+1 appends
// to it).
// The cursor, the frontier size and the wide-queue word of the current level
// are read with ordinary (scalar, cached) loads: none of them is written
// during this launch, and the previous launch's writes are visible at the
// launch boundary.  Uncached loads of the same words by every wave of a large
// grid would serialise at one memory channel.
// The control block comes in as an argument (its address is fixed for the
// engine's life) and the three ring counts and wide words are loaded whole,
// next to the cursor, so a launch learns whether its level is empty after one
// memory round trip instead of three (Dev copy, cursor, ring slot); the Dev
// copy is read only by launches that have a frontier.  Empty trailing levels
// are a third of the launches of a replayed segment.
__global__ void __launch_bounds__(kBlock) k_bfs_level_dev(const Dev* __restrict__ pD, Ctrl* ctrl, int mode, int r) {
  const int2 cur = *reinterpret_cast<const int2*>(&ctrl->bcur[r][0]);
  const int4 qc = *reinterpret_cast<const int4*>(&ctrl->qcnt[0]);  // qcnt[0..2], wl_cnt
  const ulonglong2 w01 = *reinterpret_cast<const ulonglong2*>(&ctrl->wq[0]);
  const unsigned long long w2 = ctrl->wq[2];
  const int early_flag = ctrl->early;
  const int lo = cur.x;
  const int lvl = cur.y;
  const int slot = lvl % 3;
  const int cnt = slot == 0 ? qc.x : (slot == 1 ? qc.y : qc.z);
  const unsigned long long wpk = slot == 0 ? w01.x : (slot == 1 ? w01.y : w2);
  if (is_block0_thread0()) {
    ctrl->bcur[r ^ 1][0] = lo + cnt;
    ctrl->bcur[r ^ 1][1] = cnt ? lvl + 1 : lvl;
    if (cnt) {
      ctrl->qcnt[(lvl + 2) % 3] = 0;
      ctrl->wq[(lvl + 2) % 3] = 0;
    }
  }
  if (cnt == 0) return;  // uniform: the BFS is over
  Dev D = *pD;
  globalize(D);
  // early-exit relabel whose targets are all labelled: this level is not
  // expanded, so the next one comes out empty.  A block that starts late may
  // see counts that blocks still running have raised and stop while others
  // expand; a partly expanded level still labels exact distances.
  const bool early = early_flag != 0;
  if (early && block_targets_labelled(D)) {
    if (threadIdx.x == 0) D.ctrl->bfs_cut = 1;
    return;
  }
  Tally t;
  if (use_bottom_up(D, mode, lo, cnt, static_cast<long long>(wpk & 0xffffffffu))) {
    ph_bfs_bottomup(D, make_range(), lvl, &D.ctrl->qcnt[(lvl + 1) % 3], (lvl + 1) % 3, D.order + lo + cnt, t,
                    D.nprev > 0 ? D.prev : nullptr, D.nprev, early);
  } else {
    const WList W = wlist(D, lvl % 3, wpk);
    ph_bfs_level(D, make_range(), mode, lo, cnt, lvl, &D.ctrl->qcnt[(lvl + 1) % 3], W, (lvl + 1) % 3, t, early);
  }
  t.flush(D.ctrs, true);
}
// Block 0 also clears the cascade ring (the BFS ring, which this round's
// relabel, if any, has finished with): the round's push lists into slot 0.
__global__ void __launch_bounds__(kBlock) k_build_wl(Dev D, int cnt) {
  if (is_block0_thread0()) {
    Ctrl* c = D.ctrl;
    c->qcnt[0] = c->qcnt[1] = c->qcnt[2] = 0;
    c->wq[0] = c->wq[1] = c->wq[2] = 0;
  }
  Tally t;
  ph_build_wl(D, make_range(), cnt, t);
  t.flush(D.ctrs);
}
__global__ void __launch_bounds__(kBlock) k_build_pp_wl(Dev D, int cnt) {
  Tally t;
  ph_build_pp_wl(D, make_range(), cnt, t);
  t.flush(D.ctrs);
}
__global__ void __launch_bounds__(kBlock) k_build_final_wl(Dev D, int cnt) {
  Tally t;
  ph_build_final_wl(D, make_range(), cnt, t);
  t.flush(D.ctrs);
}
// which: 0 = all vertices (topology), 1 = wl, 2 = wl2 (pull).
// The list sizes were read by the host before this launch and come in as
// arguments, so block 0 clears the list counters for the next round's builder.
// Topology-driven rounds alternate flag slots: this round reports activity in
// flag[fslot] and clears flag[fslot ^ 1], which the host read a round ago.
// cslot >= 0 (which == 1 only): also list, in cascade slot cslot with stamp
// cstamp, the vertices the first cascade sub-round takes.
// cnt < 0 (which == 1): the list sizes come from the snapshot k_list_snap
// took (r_cnt, r_wpk), and the launch does nothing when it marked the round
// skipped (r_skip).
__global__ void __launch_bounds__(kBlock) k_push(Dev D, int which, int cnt, int kc, int pmode,
                                                 unsigned long long wpk, int fslot, int cslot = -1,
                                                 int cstamp = 0, int crelist = 0) {
  if (cnt < 0) {
    const Ctrl* cc = D.ctrl;
    if (cc->r_skip) return;  // uniform
    cnt = cc->r_cnt;
    wpk = cc->r_wpk;
  }
  Tally t;
  const Range r = make_range();
  if (is_block0_thread0()) {
    Ctrl* c = D.ctrl;
    if (which == 0) {
      c->flag[fslot ^ 1] = 0;
    } else {
      c->wl_cnt = c->wl2_cnt = 0;
      c->wq[kQActive] = c->wq[kQPull] = 0;
    }
  }
  if (which == 0) {
    ph_push<false>(D, r, nullptr, cnt, kc, pmode, &D.ctrl->flag[fslot], wlist_waves(D), t);
  } else if (which == 1) {
    ph_push<false>(D, r, D.wl, cnt, kc, pmode, nullptr, wlist(D, kQActive, wpk), t, Cascade{cslot, cstamp, crelist});
  } else {
    ph_push<true>(D, r, D.wl2, cnt, kc, pmode, nullptr, wlist(D, kQPull, wpk), t);
  }
  t.flush(D.ctrs);
}
__global__ void __launch_bounds__(kBlock) k_repair(Dev D, int which, int cnt, int rmode,
                                                   unsigned long long wpk) {
  if (cnt < 0) {  // as k_push
    const Ctrl* cc = D.ctrl;
    if (cc->r_skip) return;
    cnt = cc->r_cnt;
    wpk = cc->r_wpk;
  }
  Tally t;
  const Range r = make_range();
  if (which == 0) ph_repair(D, r, nullptr, cnt, rmode, wlist_waves(D), t);
  else if (which == 1) ph_repair(D, r, D.wl, cnt, rmode, wlist(D, kQActive, wpk), t);
  else ph_repair(D, r, D.wl2, cnt, rmode, wlist(D, kQPull, wpk), t);
  t.flush(D.ctrs);
}
// ============================================================ round graphs
// The round-graph driver replays a whole round of the data-driven fixed
// schedule from one captured hipGraph: k_round_begin, k_round_bfs_init, a
// fixed number of k_bfs_level_dev launches, k_round_mid, k_round_wl,
// k_round_list, k_round_push, k_round_repair.  The loop conditions the
// stepped driver checks on the host live in the control block (r_*): every
// kernel reads them at its start and returns at once when the solve is over
// (r_done) or when the host has to finish the round (r_stop), so the round's
// launches are the stepped driver's, with one host check per replay instead
// of one per BFS segment and one per worklist.

// Once per solve, before the first replay: loop state, and the device copy of
// Dev that k_bfs_level_dev reads.  One wave.
__global__ void k_round_init(Dev D, Dev* out, int interval, int max_rounds) {
  if (threadIdx.x != 0) return;
  Ctrl* c = D.ctrl;
  c->r_done = c->r_noconv = c->r_stop = c->r_relabel = c->r_fresh = c->r_skip = 0;
  c->r_since = interval;  // the first round relabels
  c->r_interval = interval;
  c->r_rounds = 0;
  c->r_max = max_rounds;
  c->r_relabels = c->r_levels = c->r_maxlvl = 0;
  c->r_cnt = 0;
  c->r_wpk = 0;
  c->early = c->bfs_cut = c->bfs_rbu = 0;
  *out = D;
}
// Round start: whether this round relabels (every r_interval rounds, or at
// once after a list that came out empty on stale labels); a relabel clears
// the frontier ring and the level cursor.  One wave.
__global__ void k_round_begin(Dev D) {
  Ctrl* c = D.ctrl;
  if (threadIdx.x != 0 || c->r_done || c->r_stop) return;
  const int rl = c->r_since >= c->r_interval ? 1 : 0;
  c->r_relabel = rl;
  if (rl) {
    c->qcnt[0] = c->qcnt[1] = c->qcnt[2] = 0;
    c->wq[0] = c->wq[1] = c->wq[2] = 0;
    c->bcur[0][0] = c->bcur[0][1] = c->bcur[1][0] = c->bcur[1][1] = 0;
  }
}
// Roots of the round's relabel from t (my_init_1, code/dyn_data.h:60-67).
__global__ void __launch_bounds__(kBlock) k_round_bfs_init(Dev D) {
  const Ctrl* c = D.ctrl;
  if (c->r_done || c->r_stop || !c->r_relabel) return;  // block-uniform
  Tally t;
  ph_bfs_init(D, make_range(), BFS_SINK, t);
  t.flush(D.ctrs);
}
// After the level launches: a relabel whose frontier is not yet empty stops
// the replay (r_stop; the host runs further levels, then k_round_resume and
// the round's remaining kernels); a finished one gives the labelled set
// order[0, visited).  Clears the worklist counters and the BFS ring for the
// worklist builder.  One wave.
__global__ void k_round_mid(Dev D) {
  Ctrl* c = D.ctrl;
  if (threadIdx.x != 0 || c->r_done || c->r_stop) return;
  if (c->r_relabel) {
    const int lvl = c->bcur[0][1];
    if (c->qcnt[lvl % 3] != 0) {
      c->r_stop = 1;
      return;
    }
    c->visited = c->bcur[0][0];
    c->r_levels += lvl;
    c->r_maxlvl = max(c->r_maxlvl, lvl);
    c->r_relabels += 1;
    c->r_since = 0;
    c->r_fresh = 1;
  } else {
    c->r_fresh = 0;
  }
  c->wl_cnt = 0;
  c->wq[kQActive] = 0;
  c->qcnt[0] = c->qcnt[1] = c->qcnt[2] = 0;
  c->wq[0] = c->wq[1] = c->wq[2] = 0;
}
// Clears r_stop once the host has finished the relabel's levels.  One wave.
__global__ void k_round_resume(Dev D) {
  if (threadIdx.x == 0) D.ctrl->r_stop = 0;
}
// create_wl (code/dyn_data.h:217-238) over the labelled set of the last
// relabel, or over every vertex when that set is a quarter of the graph or
// more (one coalesced pass instead of a gather through order[]).
__global__ void __launch_bounds__(kBlock) k_round_wl(Dev D) {
  const Ctrl* c = D.ctrl;
  if (c->r_done || c->r_stop) return;  // block-uniform
  const int cnt = c->visited;
  Tally t;
  ph_build_wl(D, make_range(), cnt > D.n / 4 ? -1 : cnt, t);
  t.flush(D.ctrs);
}
// The list decides the round: empty after this round's relabel ends the
// solve (converged); empty on stale labels makes the next round relabel
// without pushing in this one; past the round cap the solve ends without
// convergence.  Otherwise the round's list is kept for push and repair.
// One wave.
__global__ void k_round_list(Dev D) {
  Ctrl* c = D.ctrl;
  if (threadIdx.x != 0 || c->r_done || c->r_stop) return;
  const int cnt = c->wl_cnt;
  const unsigned long long wpk = c->wq[kQActive];
  c->r_cnt = cnt;
  c->r_wpk = wpk;
  c->r_skip = 0;
  if (cnt == 0 && wpk == 0) {
    if (c->r_fresh) {
      c->r_done = 1;
    } else {
      c->r_since = c->r_interval;
      c->r_skip = 1;
    }
    return;
  }
  if (c->r_rounds >= c->r_max) {
    c->r_done = 1;
    c->r_noconv = 1;
    return;
  }
  c->r_rounds += 1;
  c->r_since += 1;
}
// staticMaxFlow_kernel_14 (code/dyn_data.h:89-131) over the round's list.
__global__ void __launch_bounds__(kBlock) k_round_push(Dev D, int kc) {
  const Ctrl* c = D.ctrl;
  if (c->r_done || c->r_stop || c->r_skip) return;  // block-uniform
  const int cnt = c->r_cnt;
  const unsigned long long wpk = c->r_wpk;
  Tally t;
  ph_push<false>(D, make_range(), D.wl, cnt, kc, PM_ALL, nullptr, wlist(D, kQActive, wpk), t);
  t.flush(D.ctrs);
}
// staticMaxFlow_kernel_17 (code/dyn_data.h:132-153) over the round's list.
__global__ void __launch_bounds__(kBlock) k_round_repair(Dev D) {
  const Ctrl* c = D.ctrl;
  if (c->r_done || c->r_stop || c->r_skip) return;  // block-uniform
  const int cnt = c->r_cnt;
  const unsigned long long wpk = c->r_wpk;
  Tally t;
  ph_repair(D, make_range(), D.wl, cnt, RM_ALL, wlist(D, kQActive, wpk), t);
  t.flush(D.ctrs);
}
__global__ void __launch_bounds__(kBlock) k_pp_cut(Dev D) {
  Tally t;
  ph_pp_cut(D, make_range(), t);
  t.flush(D.ctrs);
}
__global__ void __launch_bounds__(kBlock) k_find_cut(Dev D) { ph_find_cut(D, make_range()); }
__global__ void __launch_bounds__(kBlock) k_find_part(Dev D) {
  Tally t;
  ph_find_part(D, make_range(), t);
  t.flush(D.ctrs);
}
__global__ void __launch_bounds__(kBlock) k_find_final_cut(Dev D) { ph_find_final_cut(D, make_range()); }
__global__ void __launch_bounds__(kBlock) k_flow(Dev D) { ph_flow(D, make_range()); }
__global__ void __launch_bounds__(kBlock) k_upd_lookup(Dev D) { ph_upd_lookup(D, make_range()); }
__global__ void __launch_bounds__(kBlock) k_upd_apply(Dev D, int capacity_only, int zero_wide) {
  ph_upd_apply(D, make_range(), capacity_only != 0, zero_wide != 0);
}
__global__ void __launch_bounds__(kBlock) k_upd_clamp(Dev D) {
  Tally t;
  ph_upd_clamp(D, make_range(), t);
  t.flush(D.ctrs);
}
__global__ void __launch_bounds__(kBlock) k_upd_release(Dev D) { ph_upd_release(D, make_range()); }
__global__ void __launch_bounds__(kBlock) k_build_imbalance_wl(Dev D, int cnt, int deficits) {
  Tally t;
  ph_build_imbalance_wl(D, make_range(), cnt, deficits != 0, t);
  t.flush(D.ctrs);
}
__global__ void __launch_bounds__(kBlock) k_count_imbalance(Dev D) { ph_count_imbalance(D, make_range()); }

// Static lists of the vertices above the degree thresholds (the wide list of
// degree > wide_t with its bitmap, and the wave list of degree > wave_lane),
// built on the device for engines that have no host copy of the CSR.
// packed[0] / packed[1] receive the wide / wave list words (slots << 32 |
// chunks); both and widebits must be zero on entry.
__global__ void __launch_bounds__(kBlock) k_build_static_lists(Dev D, int32_t* wide_vert, int32_t* wide_base,
                                                               int32_t* wave_vert, int32_t* wave_base,
                                                               uint32_t* widebits, unsigned long long* packed) {
  const Range r = make_range();
  const int wave_t = D.wave_lane;
  for (int base = r.gid - lane_id(); base < D.n; base += r.gsz) {
    const int v = base + lane_id();
    const int d = v < D.n ? degree(D, v) : 0;
    wave_list_append(&packed[0], wide_vert, wide_base, d > D.wide_t, v, d);
    wave_list_append(&packed[1], wave_vert, wave_base, d > wave_t, v, d);
    if (v < D.n && d > D.wide_t) atomicOr(&widebits[static_cast<unsigned>(v) >> 5], 1u << (v & 31));
  }
}

// Builds the (tail, head) -> arc hash index, one thread per arc; for a repeated
// (tail, head) pair the highest arc index wins, as the reference's last-match scan.
// ivmap (engine id -> caller id; null when the engine keeps the caller's
// numbering) keys the index by the ids update records carry.
__global__ void __launch_bounds__(kBlock) k_build_hash(Dev D, const int32_t* __restrict__ ivmap) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < D.m; e += gridDim.x * blockDim.x) {
    int a = 0, b = D.n - 1;  // tail = largest v with row[v] <= e
    while (a < b) {
      const int mid = (a + b + 1) >> 1;
      if (D.row[mid] <= e) a = mid;
      else b = mid - 1;
    }
    const int head = D.col[e];
    const unsigned long long key = ivmap ? pair_key(ivmap[a], ivmap[head]) : pair_key(a, head);
    unsigned long long h = hash_key(key) & D.hmask;
    for (;;) {
      const unsigned long long prev = atomicCAS(&D.htab[h].x, kEmptyKey, key);
      if (prev == kEmptyKey || prev == key) {
        // the value word starts all ones (-1 in its low half, little-endian)
        atomicMax(reinterpret_cast<int*>(&D.htab[h].y), e);
        break;
      }
      h = (h + 1) & D.hmask;
    }
  }
}

// ======================================================== persistent kernel
// One cooperative launch runs a whole solve: the batch prelude (update
// lookup/apply/clamp, source re-saturation), the relabel/push/repair rounds of
// the chosen variant, the push-pull finishing phases and the flow reduction.
// Control decisions read device counters after a grid barrier, so every block
// takes the same branch.  A counter is cleared only after a barrier that every
// reader of it has passed, and before the barrier preceding its next writer.

// Global relabel with the given root mode; returns the number of vertices
// labelled (order[0, visited)).  On entry ctrl->qcnt[] and ctrl->wq[0..2] are
// 0; on exit they are 0 again.
__device__ int relabel_persistent(const Dev& D, cg::grid_group& G, const Range& r, int mode,
                                  int part_cnt, Tally& t) {
  if (mode == BFS_FINAL) ph_bfs_init_final(D, r, part_cnt, t);
  else ph_bfs_init(D, r, mode, t);
  G.sync();
  int lo = 0, lvl = 0;
  for (;;) {
    const int cur = lvl % 3;
    const int cnt = ldr(&D.ctrl->qcnt[cur]);
    if (cnt == 0) break;
    if (use_bottom_up(D, mode, lo, cnt, static_cast<long long>(ldru64(&D.ctrl->wq[cur]) & 0xffffffffu))) {
      ph_bfs_bottomup(D, r, lvl, &D.ctrl->qcnt[(lvl + 1) % 3], (lvl + 1) % 3, D.order + lo + cnt, t);
    } else {
      const WList W = wlist(D, cur, ldru64(&D.ctrl->wq[cur]));
      ph_bfs_level(D, r, mode, lo, cnt, lvl, &D.ctrl->qcnt[(lvl + 1) % 3], W, (lvl + 1) % 3, t);
    }
    if (is_block0_thread0()) {
      // slot (lvl+2)%3 was read during level lvl-1 and is next written during
      // level lvl+1
      D.ctrl->qcnt[(lvl + 2) % 3] = 0;
      D.ctrl->wq[(lvl + 2) % 3] = 0;
      ctr_add(D.ctrs, C_LEVELS, 1);
    }
    lo += cnt;
    ++lvl;
    G.sync();
  }
  G.sync();
  if (is_block0_thread0()) {
    D.ctrl->qcnt[0] = D.ctrl->qcnt[1] = D.ctrl->qcnt[2] = 0;
    D.ctrl->wq[0] = D.ctrl->wq[1] = D.ctrl->wq[2] = 0;
    D.ctrl->visited = lo;
    ctr_add(D.ctrs, C_GRELABEL, 1);
  }
  return lo;
}

__global__ void __launch_bounds__(kBlock) k_solve_persistent(Dev D, Cfg C) {
  cg::grid_group G = cg::this_grid();
  const Range r = make_range();
  Tally t;
  const int n = D.n;
  const bool pp = C.variant == DMF_VARIANT_PP;
  const bool topo = C.variant == DMF_VARIANT_TOPO;

  // ---- prelude
  if (C.entry == 0) {
    ph_init_residual(D, r);
    G.sync();
    ph_saturate_source(D, r, t);
    G.sync();
    if (pp) {
      ph_saturate_sink_pp(D, r, t);
      G.sync();
    }
  } else if (C.entry == 1) {
    ph_upd_lookup(D, r);
    G.sync();
    if (D.ctrl->missing || D.ctrl->invalid) {
      ph_upd_release(D, r);  // leave the state untouched, owners cleared
      t.flush(D.ctrs);
      return;  // uniform: every block read the same counters after the barrier
    }
    ph_upd_apply(D, r, false, pp);
    G.sync();
    if (pp) {
      ph_pp_cut(D, r, t);
    } else {
      ph_upd_clamp(D, r, t);
      G.sync();
      ph_saturate_source(D, r, t);
    }
    if (is_block0_thread0()) ctr_add(D.ctrs, C_UPD, static_cast<unsigned long long>(D.k));
    G.sync();
  }

  int round = 0;
  bool conv = true;
  if (!pp || C.entry == 0) {
    // ---- data / topo rounds (also the push-pull static solve)
    int since = C.interval;  // rounds since the last global relabel
    int visited = 0;
    for (;;) {
      bool fresh = false;
      if (since >= C.interval) {
        visited = relabel_persistent(D, G, r, BFS_SINK, 0, t);
        since = 0;
        fresh = true;
      }
      int cnt = 0;
      unsigned long long wpk = 0;
      if (!topo) {
        ph_build_wl(D, r, visited, t);
        G.sync();
        cnt = ldr(&D.ctrl->wl_cnt);
        wpk = ldru64(&D.ctrl->wq[kQActive]);
        if (cnt == 0 && wpk == 0) {
          if (fresh) break;
          since = C.interval;  // confirm convergence with fresh labels
          G.sync();
          continue;
        }
      }
      if (round >= C.max_rounds) {
        conv = false;
        break;
      }
      if (topo) {
        // flag slots alternate by round: the slot for the next round was last
        // read one round ago, so clearing it here cannot race with a reader.
        int32_t* fl = &D.ctrl->flag[round & 1];
        ph_push<false>(D, r, nullptr, n, C.kc, PM_ALL, fl, wlist_waves(D), t);
        G.sync();
        const int flag = ldr(fl);
        if (is_block0_thread0()) D.ctrl->flag[(round + 1) & 1] = 0;
        ph_repair(D, r, nullptr, n, RM_ALL, wlist_waves(D), t);
        G.sync();
        if (!flag) break;
      } else {
        const WList W = wlist(D, kQActive, wpk);
        ph_push<false>(D, r, D.wl, cnt, C.kc, PM_ALL, nullptr, W, t);
        G.sync();
        // every block read wl_cnt / wq before the push barrier; the next
        // append comes after the repair barrier
        if (is_block0_thread0()) {
          D.ctrl->wl_cnt = 0;
          D.ctrl->wq[kQActive] = 0;
        }
        ph_repair(D, r, D.wl, cnt, RM_ALL, W, t);
        G.sync();
      }
      ++round;
      ++since;
    }
    if (pp && conv) {
      ph_find_cut(D, r);
      G.sync();
    }
  } else {
    // ---- push-pull batch: part_pp (code/dyn_pp_bench.cu:6-88).  The
    // reference relabels every round; with a longer interval the labels of the
    // last two-sided BFS stay valid (pushes and pulls relabel locally and the
    // repair passes restore validity), and convergence is only accepted on
    // fresh labels, so the final BFS that part_final reads is always current.
    int since = C.interval, visited = 0;
    for (;;) {
      bool fresh = false;
      if (since >= C.interval) {
        visited = relabel_persistent(D, G, r, BFS_PP, 0, t);
        since = 0;
        fresh = true;
      }
      ph_build_pp_wl(D, r, visited, t);
      G.sync();
      const int np = ldr(&D.ctrl->wl_cnt), nq = ldr(&D.ctrl->wl2_cnt);
      const unsigned long long wp = ldru64(&D.ctrl->wq[kQActive]);
      const unsigned long long wq = ldru64(&D.ctrl->wq[kQPull]);
      if (np + nq == 0 && wp == 0 && wq == 0) {
        if (fresh) break;
        since = C.interval;
        G.sync();
        continue;
      }
      if (round >= C.max_rounds) {
        conv = false;
        break;
      }
      const WList WP = wlist(D, kQActive, wp), WQ = wlist(D, kQPull, wq);
      ph_push<false>(D, r, D.wl, np, C.kc, PM_ONLY_B, nullptr, WP, t);
      ph_push<true>(D, r, D.wl2, nq, C.kc, PM_ONLY_A, nullptr, WQ, t);
      G.sync();
      // every block read the list counters before the push barrier; the next
      // append comes after the repair barrier
      if (is_block0_thread0()) {
        D.ctrl->wl_cnt = D.ctrl->wl2_cnt = 0;
        D.ctrl->wq[kQActive] = D.ctrl->wq[kQPull] = 0;
      }
      ph_repair(D, r, D.wl, np, RM_B, WP, t);
      ph_repair(D, r, D.wl2, nq, RM_PULL_A, WQ, t);
      G.sync();
      ++round;
      ++since;
    }
    // ---- part_final (code/dyn_pp_bench.cu:91-152)
    if (conv) {
      ph_find_part(D, r, t);
      G.sync();
      const int parts = ldr(&D.ctrl->part_cnt);
      if (parts > 0) {
        for (;;) {
          relabel_persistent(D, G, r, BFS_FINAL, parts, t);
          ph_build_final_wl(D, r, parts, t);
          G.sync();
          const int cnt = ldr(&D.ctrl->wl_cnt);
          const unsigned long long wpk = ldru64(&D.ctrl->wq[kQActive]);
          if (cnt == 0 && wpk == 0) break;
          if (round >= C.max_rounds) {
            conv = false;
            break;
          }
          const WList W = wlist(D, kQActive, wpk);
          ph_push<false>(D, r, D.wl, cnt, C.kc, PM_ALL, nullptr, W, t);
          G.sync();
          ph_repair(D, r, D.wl, cnt, RM_PFINAL, W, t);
          G.sync();
          if (is_block0_thread0()) {
            D.ctrl->wl_cnt = 0;
            D.ctrl->wq[kQActive] = 0;
          }
          ++round;
        }
        if (conv) ph_find_final_cut(D, r);
      }
      G.sync();
    }
  }
  for (int i = r.gid; i < D.ninst; i += r.gsz) D.iflow[i] = 0;
  G.sync();
  if (is_block0_thread0()) {
    D.ctrl->wl_cnt = D.ctrl->wl2_cnt = D.ctrl->part_cnt = 0;
    D.ctrl->wq[kQActive] = D.ctrl->wq[kQPull] = 0;
    D.ctrl->flag[0] = D.ctrl->flag[1] = 0;
    D.ctrl->rounds = round;
    ctr_add(D.ctrs, C_ROUNDS, static_cast<unsigned long long>(round));
    D.ctrl->flow = 0;
    if (!conv) D.ctrl->status = DMF_ENOCONV;
  }
  G.sync();
  ph_flow(D, r);
  t.flush(D.ctrs);
}

}  // namespace dmfdev
