// MI355X (gfx950) dynamic max-flow engine: device phases, stepped and
// persistent drivers, and the C-ABI declared in include/dmf.h.
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
//   hkey/hval                  open-addressing (tail,head) -> arc index, so an
//                              update finds its arc in O(1) instead of the
//                              reference's scan of the tail's adjacency
//
// Every phase is a __device__ function over a grid-stride range so that the
// same code runs either as one kernel per phase (stepped driver, host-checked
// loop conditions) or inside one cooperative persistent kernel per solve with
// grid barriers between phases (no host round trips).
//
// Work distribution: a worklist / frontier is consumed 64 entries per wave.
// Vertices with degree <= kLaneMax are scanned serially by one lane; larger
// ones are scanned by the whole wave64, lanes striding over the adjacency and
// combining with cross-lane reductions.
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/dmf.h"

namespace cg = cooperative_groups;

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kLaneMax = 32;

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
  C_N = 16
};

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
  int32_t pad[3];
  unsigned long long ctr[C_N];
  long long flow;
};

struct Dev {
  int32_t n, m, s, t;
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
  unsigned long long* hkey;
  int32_t* hval;
  unsigned long long hmask;
  int32_t* owner;
  // current update batch
  const int32_t* usrc;
  const int32_t* udst;
  const int32_t* uw;
  int32_t* uarc;
  int32_t k;
};

// Solve configuration passed to every driver.
struct Cfg {
  int32_t variant;
  int32_t kc;
  int32_t interval;
  int32_t max_rounds;
  int32_t entry;  // 0 = static (init + saturate), 1 = batch (apply + repair), 2 = loop only
};

constexpr unsigned long long kEmptyKey = ~0ull;

// ------------------------------------------------------------ small helpers
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ int ldr(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long ldr64(const long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void str(int32_t* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void add64(long long* p, long long v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v));
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long y = __shfl_xor(x, o);
    x = y < x ? y : x;
  }
  return x;
}
__device__ __forceinline__ long long wave_sum_i64(long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
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

// Per-thread work counters, flushed once per phase with one atomic per wave.
struct Tally {
  unsigned long long arc = 0, vert = 0, push = 0, relabel = 0, repair = 0, atom = 0;
  __device__ void flush(Ctrl* c) {
    unsigned long long v[6] = {arc, vert, push, relabel, repair, atom};
    const int idx[6] = {C_ARC, C_VERT, C_PUSH, C_RELABEL, C_REPAIR, C_ATOMIC};
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      unsigned long long x = v[i];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
      if (lane_id() == 0 && x) atomicAdd(&c->ctr[idx[i]], x);
    }
  }
};

struct Range {
  int gid;    // global thread index
  int gsz;    // total threads
  int wid;    // global wave index
  int nw;     // total waves
};

__device__ __forceinline__ Range make_range() {
  Range r;
  r.gid = blockIdx.x * blockDim.x + threadIdx.x;
  r.gsz = gridDim.x * blockDim.x;
  r.wid = r.gid / kWave;
  r.nw = r.gsz / kWave;
  return r;
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
__device__ void ph_saturate_source(const Dev& D, const Range& r, Tally& t) {
  const int s = D.s;
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

// Second half of the push-pull static saturation (code/dyn_pp.h:56-70): pull
// every residual arc into t.
__device__ void ph_saturate_sink_pp(const Dev& D, const Range& r, Tally& t) {
  const int v = D.t;
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

// Root predicates of the three global relabels.
enum BfsMode { BFS_SINK = 0, BFS_PP = 1, BFS_FINAL = 2 };

// my_init_1 (code/dyn_data.h:60-67) / staticMaxFlow_kernel_10 (code/dyn_topo.h:60-71)
// / pp_bfs_init (code/dyn_pp.h:76-83): every vertex unlabelled (height n),
// roots labelled 0 and appended to order[].  ctrl->qcnt[0] must be 0.
__device__ void ph_bfs_init(const Dev& D, const Range& r, int mode, Tally& t) {
  const int n = D.n;
  for (int base = r.gid - lane_id(); base < n; base += r.gsz) {
    const int v = base + lane_id();
    bool root = false;
    if (v < n) {
      const long long ex = D.excess[v];
      if (mode == BFS_SINK) {
        root = (v == D.t) || (v != D.s && ex < 0);
      } else {
        const bool inB = D.side[v] != 0;
        root = (v == D.s) || (v == D.t) || (inB && ex < 0) || (!inB && ex > 0);
      }
      D.height[v] = root ? 0 : n;
      t.vert += 1;
    }
    int rank;
    const int off = wave_reserve(&D.ctrl->qcnt[0], root, &rank);
    if (root) D.order[off + rank] = v;
  }
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
      t.vert += 1;
    }
    int rank;
    const int off = wave_reserve(&D.ctrl->qcnt[0], root, &rank);
    if (root) D.order[off + rank] = v;
  }
}

// Whether BFS level expansion may label w through arc e of frontier vertex v.
__device__ __forceinline__ bool bfs_arc_ok(const Dev& D, int mode, bool vB, int e, int w) {
  if (mode == BFS_SINK) return D.rres[e] > 0;
  if (mode == BFS_PP) {
    if (w == D.s || w == D.t) return false;
    return vB ? (D.rres[e] > 0) : (D.res[e] > 0);
  }
  // BFS_FINAL: non-B neighbours of the stranded part only
  if (D.side[w] || !D.pfinal[w]) return false;
  return D.rres[e] > 0;
}

// One BFS level (staticMaxFlow_kernel_11, code/dyn_data.h:72-87; pp_bfs,
// code/dyn_pp.h:99-120; final_bfs, code/dyn_pp.h:123-140).  Frontier is
// order[lo, lo+cnt); newly labelled vertices get height lvl+1 (one winner per
// vertex through compare-and-swap) and are appended after the frontier.
__device__ void ph_bfs_level(const Dev& D, const Range& r, int mode, int lo, int cnt, int lvl,
                             int32_t* next_cnt, Tally& t) {
  const int n = D.n;
  const int nb = lvl + 1;
  int32_t* out = D.order + lo + cnt;
  for (int base = r.wid * kWave; base < cnt; base += r.nw * kWave) {
    const int i = base + lane_id();
    int v = -1, b = 0, deg = 0;
    bool vB = false;
    if (i < cnt) {
      v = D.order[lo + i];
      b = D.row[v];
      deg = D.row[v + 1] - b;
      vB = mode == BFS_PP && D.side[v] != 0;
      t.vert += 1;
    }
    // wide vertices: the whole wave walks the adjacency
    unsigned long long big = __ballot(deg > kLaneMax);
    while (big) {
      const int leader = __ffsll(static_cast<long long>(big)) - 1;
      big &= big - 1;
      const int bv = __shfl(v, leader);
      const int bb = __shfl(b, leader);
      const int bd = __shfl(deg, leader);
      const bool bB = __shfl(static_cast<int>(vB), leader) != 0;
      for (int j0 = 0; j0 < bd; j0 += kWave) {
        const int j = j0 + lane_id();
        bool hit = false;
        int w = -1;
        if (j < bd) {
          const int e = bb + j;
          w = D.col[e];
          t.arc += 1;
          if (bfs_arc_ok(D, mode, bB, e, w) && D.height[w] == n) {
            hit = atomicCAS(&D.height[w], n, nb) == n;
            t.atom += 1;
          }
        }
        int rank;
        const int off = wave_reserve(next_cnt, hit, &rank);
        if (hit) out[off + rank] = w;
      }
      (void)bv;
    }
    // narrow vertices: one lane each, lanes step through arcs in lockstep
    const int my_deg = (deg <= kLaneMax) ? deg : 0;
    int maxd = my_deg;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) maxd = max(maxd, __shfl_xor(maxd, o));
    for (int j = 0; j < maxd; ++j) {
      bool hit = false;
      int w = -1;
      if (j < my_deg) {
        const int e = b + j;
        w = D.col[e];
        t.arc += 1;
        if (bfs_arc_ok(D, mode, vB, e, w) && D.height[w] == n) {
          hit = atomicCAS(&D.height[w], n, nb) == n;
          t.atom += 1;
        }
      }
      int rank;
      const int off = wave_reserve(next_cnt, hit, &rank);
      if (hit) out[off + rank] = w;
    }
  }
}

// create_wl (code/dyn_data.h:217-238) over the labelled vertices order[0,cnt):
// excess > 0, v not s or t (labelled implies height < n).  Also used with the
// current heights when the global relabel was skipped this round.
__device__ void ph_build_wl(const Dev& D, const Range& r, int cnt, Tally& t) {
  for (int base = r.gid - lane_id(); base < cnt; base += r.gsz) {
    const int i = base + lane_id();
    bool act = false;
    int v = -1;
    if (i < cnt) {
      v = D.order[i];
      act = v != D.s && v != D.t && D.excess[v] > 0 && D.height[v] < D.n;
      t.vert += 1;
    }
    int rank;
    const int off = wave_reserve(&D.ctrl->wl_cnt, act, &rank);
    if (act) D.wl[off + rank] = v;
  }
}

// create_pp_wl (code/dyn_pp.h:499-545): B vertices with excess > 0 to wl,
// A vertices with excess < 0 to wl2.
__device__ void ph_build_pp_wl(const Dev& D, const Range& r, int cnt, Tally& t) {
  for (int base = r.gid - lane_id(); base < cnt; base += r.gsz) {
    const int i = base + lane_id();
    bool a = false, b = false;
    int v = -1;
    if (i < cnt) {
      v = D.order[i];
      if (v != D.s && v != D.t && D.height[v] < D.n) {
        const long long ex = D.excess[v];
        if (D.side[v]) a = ex > 0; else b = ex < 0;
      }
      t.vert += 1;
    }
    int rank;
    int off = wave_reserve(&D.ctrl->wl_cnt, a, &rank);
    if (a) D.wl[off + rank] = v;
    off = wave_reserve(&D.ctrl->wl2_cnt, b, &rank);
    if (b) D.wl2[off + rank] = v;
  }
}

// create_final_wl (code/dyn_pp.h:619-642) over the part list.
__device__ void ph_build_final_wl(const Dev& D, const Range& r, int cnt, Tally& t) {
  for (int base = r.gid - lane_id(); base < cnt; base += r.gsz) {
    const int i = base + lane_id();
    bool act = false;
    int v = -1;
    if (i < cnt) {
      v = D.part[i];
      act = D.excess[v] > 0 && D.height[v] < D.n;
      t.vert += 1;
    }
    int rank;
    const int off = wave_reserve(&D.ctrl->wl_cnt, act, &rank);
    if (act) D.wl[off + rank] = v;
  }
}

// Neighbour filters of the push kernels.
enum PushMode { PM_ALL = 0, PM_ONLY_B = 1 };

// One push-or-relabel step of staticMaxFlow_kernel_14 (code/dyn_data.h:94-129)
// for vertex v, given the chosen arc.  Executed by one lane.
__device__ __forceinline__ void do_push_or_relabel(const Dev& D, int v, long long ex1, int hv,
                                                   int hh, int fe, Tally& t) {
  if (fe < 0) return;
  if (hv > hh) {
    const int fec = ldr(&D.res[fe]);
    const int d = static_cast<int>(ex1 < fec ? ex1 : fec);
    if (d <= 0) return;
    const int w = D.col[fe];
    const int p = D.rev[fe];
    add64(&D.excess[v], -static_cast<long long>(d));
    add64(&D.excess[w], d);
    atomicSub(&D.res[fe], d);
    atomicAdd(&D.rres[fe], d);
    atomicAdd(&D.res[p], d);
    atomicSub(&D.rres[p], d);
    t.push += 1;
    t.atom += 6;
  } else {
    str(&D.height[v], hh + 1);
    t.relabel += 1;
  }
}

// K6 push/relabel (code/dyn_data.h:89-131; pp_push_kernel code/dyn_pp.h:191-232;
// push_kernel code/dyn_pp.h:145-185; topo code/dyn_topo.h:90-136).
// Up to kc steps per vertex: find the lowest-height neighbour over residual
// arcs (first minimum in CSR order), push min(excess, residual) to it if it is
// lower, else relabel to its height + 1.  list == nullptr means every vertex
// (topology-driven), where s, t and inactive vertices are skipped and any
// active vertex sets ctrl->flag.
__device__ void ph_push(const Dev& D, const Range& r, const int32_t* list, int cnt, int kc,
                        int pmode, int32_t* flag, Tally& t) {
  const int n = D.n;
  for (int base = r.wid * kWave; base < cnt; base += r.nw * kWave) {
    const int i = base + lane_id();
    int v = -1, b = 0, deg = 0;
    if (i < cnt) {
      v = list ? list[i] : i;
      if (!list) {
        const bool act = v != D.s && v != D.t && D.excess[v] > 0 && D.height[v] < n;
        if (!act) v = -1; else if (*flag == 0) *flag = 1;
      }
      if (v >= 0) {
        b = D.row[v];
        deg = D.row[v + 1] - b;
        t.vert += 1;
      }
    }
    // wide vertices, whole wave per step
    unsigned long long big = __ballot(v >= 0 && deg > kLaneMax);
    while (big) {
      const int leader = __ffsll(static_cast<long long>(big)) - 1;
      big &= big - 1;
      const int bv = __shfl(v, leader);
      const int bb = __shfl(b, leader);
      const int bd = __shfl(deg, leader);
      for (int c = 0; c < kc; ++c) {
        long long ex1 = 0;
        int hv = 0;
        if (lane_id() == 0) {
          ex1 = ldr64(&D.excess[bv]);
          hv = ldr(&D.height[bv]);
        }
        ex1 = __shfl(ex1, 0);
        hv = __shfl(hv, 0);
        if (!(ex1 > 0 && hv < n)) break;
        unsigned long long best = kEmptyKey;
        for (int j = lane_id(); j < bd; j += kWave) {
          const int e = bb + j;
          const int w = D.col[e];
          if (pmode == PM_ONLY_B && !D.side[w]) continue;
          const int h1 = ldr(&D.height[w]);
          if (ldr(&D.res[e]) > 0) {
            const unsigned long long key =
                (static_cast<unsigned long long>(static_cast<unsigned>(h1)) << 32) |
                static_cast<unsigned>(j);
            best = key < best ? key : best;
          }
        }
        t.arc += (lane_id() < bd) ? (bd - lane_id() + kWave - 1) / kWave : 0;
        best = wave_min_u64(best);
        if (best == kEmptyKey) continue;  // no residual arc: nothing changes
        if (lane_id() == 0) {
          const int hh = static_cast<int>(best >> 32);
          const int fe = bb + static_cast<int>(best & 0xffffffffu);
          do_push_or_relabel(D, bv, ex1, hv, hh, fe, t);
        }
      }
    }
    // narrow vertices, one lane each
    if (v >= 0 && deg <= kLaneMax) {
      for (int c = 0; c < kc; ++c) {
        const long long ex1 = ldr64(&D.excess[v]);
        const int hv = ldr(&D.height[v]);
        if (!(ex1 > 0 && hv < n)) break;
        int hh = INT_MAX, fe = -1;
        for (int e = b; e < b + deg; ++e) {
          const int w = D.col[e];
          if (pmode == PM_ONLY_B && !D.side[w]) continue;
          const int h1 = ldr(&D.height[w]);
          if (h1 < hh && ldr(&D.res[e]) > 0) {
            hh = h1;
            fe = e;
          }
        }
        t.arc += deg;
        do_push_or_relabel(D, v, ex1, hv, hh, fe, t);
      }
    }
  }
}

// pull_kernel (code/dyn_pp.h:237-280): a deficit vertex in A pulls
// min(-excess, residual) from its lowest A neighbour that can send to it, or
// relabels.
__device__ __forceinline__ void do_pull_or_relabel(const Dev& D, int v, long long need, int hv,
                                                   int hh, int fe, Tally& t) {
  if (fe < 0) return;
  if (hv > hh) {
    const int rec = ldr(&D.rres[fe]);
    const int d = static_cast<int>(need < rec ? need : rec);
    if (d <= 0) return;
    const int w = D.col[fe];
    const int p = D.rev[fe];
    add64(&D.excess[w], -static_cast<long long>(d));
    add64(&D.excess[v], d);
    atomicSub(&D.rres[fe], d);
    atomicAdd(&D.res[fe], d);
    atomicAdd(&D.rres[p], d);
    atomicSub(&D.res[p], d);
    t.push += 1;
    t.atom += 6;
  } else {
    str(&D.height[v], hh + 1);
    t.relabel += 1;
  }
}

__device__ void ph_pull(const Dev& D, const Range& r, const int32_t* list, int cnt, int kc,
                        Tally& t) {
  for (int base = r.wid * kWave; base < cnt; base += r.nw * kWave) {
    const int i = base + lane_id();
    int v = -1, b = 0, deg = 0;
    if (i < cnt) {
      v = list[i];
      b = D.row[v];
      deg = D.row[v + 1] - b;
      t.vert += 1;
    }
    unsigned long long big = __ballot(v >= 0 && deg > kLaneMax);
    while (big) {
      const int leader = __ffsll(static_cast<long long>(big)) - 1;
      big &= big - 1;
      const int bv = __shfl(v, leader);
      const int bb = __shfl(b, leader);
      const int bd = __shfl(deg, leader);
      for (int c = 0; c < kc; ++c) {
        long long need = 0;
        int hv = 0;
        if (lane_id() == 0) {
          need = -ldr64(&D.excess[bv]);
          hv = ldr(&D.height[bv]);
        }
        need = __shfl(need, 0);
        hv = __shfl(hv, 0);
        if (!(need > 0)) break;
        unsigned long long best = kEmptyKey;
        for (int j = lane_id(); j < bd; j += kWave) {
          const int e = bb + j;
          const int w = D.col[e];
          if (D.side[w]) continue;
          const int h1 = ldr(&D.height[w]);
          if (ldr(&D.rres[e]) > 0) {
            const unsigned long long key =
                (static_cast<unsigned long long>(static_cast<unsigned>(h1)) << 32) |
                static_cast<unsigned>(j);
            best = key < best ? key : best;
          }
        }
        t.arc += (lane_id() < bd) ? (bd - lane_id() + kWave - 1) / kWave : 0;
        best = wave_min_u64(best);
        if (best == kEmptyKey) continue;
        if (lane_id() == 0) {
          const int hh = static_cast<int>(best >> 32);
          const int fe = bb + static_cast<int>(best & 0xffffffffu);
          do_pull_or_relabel(D, bv, need, hv, hh, fe, t);
        }
      }
    }
    if (v >= 0 && deg <= kLaneMax) {
      for (int c = 0; c < kc; ++c) {
        const long long need = -ldr64(&D.excess[v]);
        const int hv = ldr(&D.height[v]);
        if (!(need > 0)) break;
        int hh = INT_MAX, fe = -1;
        for (int e = b; e < b + deg; ++e) {
          const int w = D.col[e];
          if (D.side[w]) continue;
          const int h1 = ldr(&D.height[w]);
          if (h1 < hh && ldr(&D.rres[e]) > 0) {
            hh = h1;
            fe = e;
          }
        }
        t.arc += deg;
        do_pull_or_relabel(D, v, need, hv, hh, fe, t);
      }
    }
  }
}

// Label repair on one arc (staticMaxFlow_kernel_17, code/dyn_data.h:140-150):
// residual w->v exists but h(w) > h(v)+1, so saturate w->v.  Heights are not
// written during this phase, so at most one endpoint of an arc pair acts.
__device__ __forceinline__ void repair_arc(const Dev& D, int v, int h, int e, int w, Tally& t) {
  const int d = D.rres[e];
  if (d > 0 && D.height[w] > h + 1) {
    D.rres[e] = 0;
    D.res[e] += d;
    const int p = D.rev[e];
    D.rres[p] += d;
    D.res[p] -= d;
    add64(&D.excess[w], -static_cast<long long>(d));
    add64(&D.excess[v], d);
    t.repair += 1;
    t.atom += 2;
  }
}

// remove_pull (code/dyn_pp.h:381-403): residual v->w exists but h(w) > h(v)+1,
// so saturate v->w.
__device__ __forceinline__ void repair_pull_arc(const Dev& D, int v, int h, int e, int w,
                                                Tally& t) {
  const int d = D.res[e];
  if (d > 0 && D.height[w] > h + 1) {
    D.rres[e] += d;
    D.res[e] = 0;
    const int p = D.rev[e];
    D.rres[p] -= d;
    D.res[p] += d;
    add64(&D.excess[v], -static_cast<long long>(d));
    add64(&D.excess[w], d);
    t.repair += 1;
    t.atom += 2;
  }
}

// Repair filters: all neighbours, B neighbours (pp_remove_push), A neighbours
// with the pull rule (remove_pull), part neighbours (remove_push_final).
enum RepairMode { RM_ALL = 0, RM_B = 1, RM_PULL_A = 2, RM_PFINAL = 3 };

__device__ __forceinline__ bool repair_filter(const Dev& D, int rmode, int w) {
  if (rmode == RM_B) return D.side[w] != 0;
  if (rmode == RM_PULL_A) return D.side[w] == 0;
  if (rmode == RM_PFINAL) return D.pfinal[w] != 0;
  return true;
}

__device__ void ph_repair(const Dev& D, const Range& r, const int32_t* list, int cnt, int rmode,
                          Tally& t) {
  for (int base = r.wid * kWave; base < cnt; base += r.nw * kWave) {
    const int i = base + lane_id();
    int v = -1, b = 0, deg = 0, h = 0;
    if (i < cnt) {
      v = list ? list[i] : i;
      b = D.row[v];
      deg = D.row[v + 1] - b;
      h = D.height[v];
      t.vert += 1;
    }
    unsigned long long big = __ballot(v >= 0 && deg > kLaneMax);
    while (big) {
      const int leader = __ffsll(static_cast<long long>(big)) - 1;
      big &= big - 1;
      const int bv = __shfl(v, leader);
      const int bb = __shfl(b, leader);
      const int bd = __shfl(deg, leader);
      const int bh = __shfl(h, leader);
      for (int j = lane_id(); j < bd; j += kWave) {
        const int e = bb + j;
        const int w = D.col[e];
        t.arc += 1;
        if (!repair_filter(D, rmode, w)) continue;
        if (rmode == RM_PULL_A) repair_pull_arc(D, bv, bh, e, w, t);
        else repair_arc(D, bv, bh, e, w, t);
      }
    }
    if (v >= 0 && deg <= kLaneMax) {
      for (int e = b; e < b + deg; ++e) {
        const int w = D.col[e];
        t.arc += 1;
        if (!repair_filter(D, rmode, w)) continue;
        if (rmode == RM_PULL_A) repair_pull_arc(D, v, h, e, w, t);
        else repair_arc(D, v, h, e, w, t);
      }
    }
  }
}

// pp incremental_kernel_20 (code/dyn_pp.h:405-436): clamp, saturate the A->B
// direction of arcs crossing the stored cut, recompute excess = net inflow.
__device__ void ph_pp_cut(const Dev& D, const Range& r, Tally& t) {
  const int n = D.n;
  for (int base = r.wid * kWave; base < n; base += r.nw * kWave) {
    const int v = base + lane_id();
    int b = 0, deg = 0;
    uint8_t self = 0;
    if (v < n) {
      b = D.row[v];
      deg = D.row[v + 1] - b;
      self = D.side[v];
      t.vert += 1;
    }
    auto arc_fix = [&](int e, uint8_t sv) -> long long {
      int rs = D.res[e], rr = D.rres[e];
      if (rs < 0) { rr += rs; rs = 0; }
      if (rr < 0) { rs += rr; rr = 0; }
      if (sv != D.side[D.col[e]]) {
        if (sv) { rs += rr; rr = 0; } else { rr += rs; rs = 0; }
      }
      D.res[e] = rs;
      D.rres[e] = rr;
      return static_cast<long long>(rs) - D.cap[e];
    };
    unsigned long long big = __ballot(v < n && deg > kLaneMax);
    while (big) {
      const int leader = __ffsll(static_cast<long long>(big)) - 1;
      big &= big - 1;
      const int bv = __shfl(v, leader);
      const int bb = __shfl(b, leader);
      const int bd = __shfl(deg, leader);
      const uint8_t bs = static_cast<uint8_t>(__shfl(static_cast<int>(self), leader));
      long long acc = 0;
      for (int j = lane_id(); j < bd; j += kWave) {
        acc += arc_fix(bb + j, bs);
        t.arc += 1;
      }
      acc = wave_sum_i64(acc);
      if (lane_id() == 0) D.excess[bv] = acc;
    }
    if (v < n && deg <= kLaneMax) {
      long long acc = 0;
      for (int e = b; e < b + deg; ++e) acc += arc_fix(e, self);
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
    if (v == D.t) acc += ex;
    if (ex < 0 && v != D.s) acc += ex;
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

__device__ int lookup_arc(const Dev& D, int u, int v) {
  const unsigned long long key = pair_key(u, v);
  unsigned long long h = hash_key(key) & D.hmask;
  for (;;) {
    const unsigned long long k = D.hkey[h];
    if (k == key) return D.hval[h];
    if (k == kEmptyKey) return -1;
    h = (h + 1) & D.hmask;
  }
}

// OnAdd_kernel arc search (code/dyn_data.h:191-201) by hash lookup, plus
// last-record-wins resolution of duplicate arcs within the batch.
__device__ void ph_upd_lookup(const Dev& D, const Range& r) {
  int missing = 0, invalid = 0;
  for (int i = r.gid; i < D.k; i += r.gsz) {
    const int a = D.usrc[i], b = D.udst[i];
    int arc = -2;
    if (a < 0 || a >= D.n || b < 0 || b >= D.n) {
      ++invalid;
    } else if (a != b) {  // self-loops skipped (code/update.hpp:82, code/dyn_data.h:194)
      arc = lookup_arc(D, a, b);
      if (arc < 0) ++missing;
      else atomicMax(&D.owner[arc], i);
    }
    D.uarc[i] = arc;
  }
  if (missing) atomicAdd(&D.ctrl->missing, missing);
  if (invalid) atomicAdd(&D.ctrl->invalid, invalid);
}

// OnAdd_kernel update (code/dyn_data.h:202-205) for the winning record of each
// arc: residual and its mirror move with the capacity.  capacity_only mirrors
// update_kernel (code/dyn_data_bench.cu:300-322).
__device__ void ph_upd_apply(const Dev& D, const Range& r, bool capacity_only) {
  if (D.ctrl->missing || D.ctrl->invalid) return;
  for (int i = r.gid; i < D.k; i += r.gsz) {
    const int arc = D.uarc[i];
    if (arc < 0 || D.owner[arc] != i) continue;
    const int nw = D.uw[i];
    if (!capacity_only) {
      const int delta = nw - D.cap[arc];
      D.res[arc] += delta;
      D.rres[D.rev[arc]] += delta;
    }
    D.cap[arc] = nw;
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
    if (arc < 0 || D.owner[arc] != i) continue;
    const int rs = D.res[arc];
    t.arc += 1;
    if (rs < 0) {
      const int a = -rs;
      const int p = D.rev[arc];
      D.res[arc] = 0;
      D.rres[p] = 0;
      D.rres[arc] -= a;
      D.res[p] -= a;
      add64(&D.excess[D.usrc[i]], a);
      add64(&D.excess[D.udst[i]], -static_cast<long long>(a));
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

// Reachability from s and from positive-excess vertices over res > 0 arcs,
// one BFS level (used for S_min).  mark = height array reused by the caller.
__device__ void ph_cut_init(const Dev& D, const Range& r, int32_t* mark) {
  for (int base = r.gid - lane_id(); base < D.n; base += r.gsz) {
    const int v = base + lane_id();
    bool root = false;
    if (v < D.n) {
      root = v == D.s || (v != D.t && D.excess[v] > 0);
      mark[v] = root ? 1 : 0;
    }
    int rank;
    const int off = wave_reserve(&D.ctrl->qcnt[0], root, &rank);
    if (root) D.order[off + rank] = v;
  }
}

__device__ void ph_cut_level(const Dev& D, const Range& r, int32_t* mark, int lo, int cnt,
                             int32_t* next_cnt) {
  int32_t* out = D.order + lo + cnt;
  for (int base = r.wid * kWave; base < cnt; base += r.nw * kWave) {
    const int i = base + lane_id();
    int b = 0, deg = 0;
    if (i < cnt) {
      const int v = D.order[lo + i];
      b = D.row[v];
      deg = D.row[v + 1] - b;
    }
    unsigned long long big = __ballot(deg > kLaneMax);
    while (big) {
      const int leader = __ffsll(static_cast<long long>(big)) - 1;
      big &= big - 1;
      const int bb = __shfl(b, leader);
      const int bd = __shfl(deg, leader);
      for (int j0 = 0; j0 < bd; j0 += kWave) {
        const int j = j0 + lane_id();
        bool hit = false;
        int w = -1;
        if (j < bd) {
          const int e = bb + j;
          w = D.col[e];
          if (D.res[e] > 0 && mark[w] == 0) hit = atomicCAS(&mark[w], 0, 1) == 0;
        }
        int rank;
        const int off = wave_reserve(next_cnt, hit, &rank);
        if (hit) out[off + rank] = w;
      }
    }
    const int my_deg = deg <= kLaneMax ? deg : 0;
    int maxd = my_deg;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) maxd = max(maxd, __shfl_xor(maxd, o));
    for (int j = 0; j < maxd; ++j) {
      bool hit = false;
      int w = -1;
      if (j < my_deg) {
        const int e = b + j;
        w = D.col[e];
        if (D.res[e] > 0 && mark[w] == 0) hit = atomicCAS(&mark[w], 0, 1) == 0;
      }
      int rank;
      const int off = wave_reserve(next_cnt, hit, &rank);
      if (hit) out[off + rank] = w;
    }
  }
}

// ============================================================ stepped kernels
__global__ void __launch_bounds__(kBlock) k_init_residual(Dev D) { ph_init_residual(D, make_range()); }
__global__ void __launch_bounds__(kBlock) k_saturate_source(Dev D) {
  Tally t;
  ph_saturate_source(D, make_range(), t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_saturate_sink_pp(Dev D) {
  Tally t;
  ph_saturate_sink_pp(D, make_range(), t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_bfs_init(Dev D, int mode) {
  Tally t;
  ph_bfs_init(D, make_range(), mode, t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_bfs_init_final(Dev D, int cnt) {
  Tally t;
  ph_bfs_init_final(D, make_range(), cnt, t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_bfs_level(Dev D, int mode, int lo, int cnt, int lvl, int slot) {
  Tally t;
  ph_bfs_level(D, make_range(), mode, lo, cnt, lvl, &D.ctrl->qcnt[slot], t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_build_wl(Dev D, int cnt) {
  Tally t;
  ph_build_wl(D, make_range(), cnt, t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_build_pp_wl(Dev D, int cnt) {
  Tally t;
  ph_build_pp_wl(D, make_range(), cnt, t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_build_final_wl(Dev D, int cnt) {
  Tally t;
  ph_build_final_wl(D, make_range(), cnt, t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_push(Dev D, int use_list, int cnt, int kc, int pmode) {
  Tally t;
  ph_push(D, make_range(), use_list ? D.wl : nullptr, cnt, kc, pmode, &D.ctrl->flag[0], t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_pull(Dev D, int cnt, int kc) {
  Tally t;
  ph_pull(D, make_range(), D.wl2, cnt, kc, t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_repair(Dev D, int which, int cnt, int rmode) {
  Tally t;
  const int32_t* list = which == 0 ? nullptr : (which == 1 ? D.wl : D.wl2);
  ph_repair(D, make_range(), list, cnt, rmode, t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_pp_cut(Dev D) {
  Tally t;
  ph_pp_cut(D, make_range(), t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_find_cut(Dev D) { ph_find_cut(D, make_range()); }
__global__ void __launch_bounds__(kBlock) k_find_part(Dev D) {
  Tally t;
  ph_find_part(D, make_range(), t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_find_final_cut(Dev D) { ph_find_final_cut(D, make_range()); }
__global__ void __launch_bounds__(kBlock) k_flow(Dev D) { ph_flow(D, make_range()); }
__global__ void __launch_bounds__(kBlock) k_upd_lookup(Dev D) { ph_upd_lookup(D, make_range()); }
__global__ void __launch_bounds__(kBlock) k_upd_apply(Dev D, int capacity_only) {
  ph_upd_apply(D, make_range(), capacity_only != 0);
}
__global__ void __launch_bounds__(kBlock) k_upd_clamp(Dev D) {
  Tally t;
  ph_upd_clamp(D, make_range(), t);
  t.flush(D.ctrl);
}
__global__ void __launch_bounds__(kBlock) k_upd_release(Dev D) { ph_upd_release(D, make_range()); }
__global__ void __launch_bounds__(kBlock) k_cut_init(Dev D, int32_t* mark) { ph_cut_init(D, make_range(), mark); }
__global__ void __launch_bounds__(kBlock) k_cut_level(Dev D, int32_t* mark, int lo, int cnt, int slot) {
  ph_cut_level(D, make_range(), mark, lo, cnt, &D.ctrl->qcnt[slot]);
}
__global__ void __launch_bounds__(kBlock) k_set_int(int32_t* p, int v) { *p = v; }

// Builds the (tail, head) -> arc hash index.
__global__ void __launch_bounds__(kBlock) k_build_hash(Dev D) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < D.n; v += gridDim.x * blockDim.x) {
    for (int e = D.row[v]; e < D.row[v + 1]; ++e) {
      const unsigned long long key = pair_key(v, D.col[e]);
      unsigned long long h = hash_key(key) & D.hmask;
      for (;;) {
        const unsigned long long prev = atomicCAS(&D.hkey[h], kEmptyKey, key);
        if (prev == kEmptyKey || prev == key) {
          D.hval[h] = e;  // duplicate (tail,head): last arc wins, as a linear scan would
          break;
        }
        h = (h + 1) & D.hmask;
      }
    }
  }
}

// ======================================================== persistent kernel
// One cooperative launch runs a whole solve: the batch prelude (update
// lookup/apply/clamp, source re-saturation), the relabel/push/repair rounds of
// the chosen variant, the push-pull finishing phases and the flow reduction.
// Control decisions read device counters after a grid barrier, so every block
// takes the same branch.  Counters are reset one barrier after their last read.

struct Grid {
  cg::grid_group g;
  __device__ void sync() { g.sync(); }
};

__device__ bool is_block0_thread0() { return blockIdx.x == 0 && threadIdx.x == 0; }

// Global relabel with the given root mode; returns the number of vertices
// labelled (order[0, visited)).  On entry ctrl->qcnt[] must all be 0; on exit
// they are 0 again.
__device__ int relabel_persistent(const Dev& D, Grid& G, const Range& r, int mode, int part_cnt,
                                  Tally& t) {
  if (mode == BFS_FINAL) ph_bfs_init_final(D, r, part_cnt, t);
  else ph_bfs_init(D, r, mode, t);
  G.sync();
  int lo = 0, lvl = 0;
  for (;;) {
    const int cur = lvl % 3;
    const int cnt = ldr(&D.ctrl->qcnt[cur]);
    if (cnt == 0) break;
    ph_bfs_level(D, r, mode, lo, cnt, lvl, &D.ctrl->qcnt[(lvl + 1) % 3], t);
    if (is_block0_thread0()) {
      D.ctrl->qcnt[(lvl + 2) % 3] = 0;
      D.ctrl->ctr[C_LEVELS] += 1;
    }
    lo += cnt;
    ++lvl;
    G.sync();
  }
  // qcnt[lvl%3] == 0 was read by everyone; the slot written during level lvl-1's
  // successor was cleared above.  Clear the remaining one after a barrier.
  G.sync();
  if (is_block0_thread0()) {
    D.ctrl->qcnt[0] = D.ctrl->qcnt[1] = D.ctrl->qcnt[2] = 0;
    D.ctrl->visited = lo;
    D.ctrl->ctr[C_GRELABEL] += 1;
  }
  return lo;
}

__global__ void __launch_bounds__(kBlock) k_solve_persistent(Dev D, Cfg C) {
  Grid G{cg::this_grid()};
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
      t.flush(D.ctrl);
      return;  // uniform: every block read the same counters after the barrier
    }
    ph_upd_apply(D, r, false);
    G.sync();
    if (pp) {
      ph_pp_cut(D, r, t);
    } else {
      ph_upd_clamp(D, r, t);
      G.sync();
      ph_saturate_source(D, r, t);
    }
    G.sync();
    ph_upd_release(D, r);
    if (is_block0_thread0()) D.ctrl->ctr[C_UPD] += D.k;
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
      int cnt;
      if (topo) {
        cnt = 0;  // decided after the push pass
      } else {
        ph_build_wl(D, r, visited, t);
        G.sync();
        cnt = ldr(&D.ctrl->wl_cnt);
        if (cnt == 0) {
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
        ph_push(D, r, nullptr, n, C.kc, PM_ALL, fl, t);
        G.sync();
        const int flag = ldr(fl);
        if (is_block0_thread0()) D.ctrl->flag[(round + 1) & 1] = 0;
        ph_repair(D, r, nullptr, n, RM_ALL, t);
        G.sync();
        if (!flag) break;
      } else {
        ph_push(D, r, D.wl, cnt, C.kc, PM_ALL, nullptr, t);
        G.sync();
        // every block read wl_cnt before this barrier; the next append comes
        // after the repair barrier
        if (is_block0_thread0()) D.ctrl->wl_cnt = 0;
        ph_repair(D, r, D.wl, cnt, RM_ALL, t);
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
    // ---- push-pull batch: part_pp (code/dyn_pp_bench.cu:6-88)
    for (;;) {
      const int visited = relabel_persistent(D, G, r, BFS_PP, 0, t);
      ph_build_pp_wl(D, r, visited, t);
      G.sync();
      const int np = ldr(&D.ctrl->wl_cnt), nq = ldr(&D.ctrl->wl2_cnt);
      if (np + nq == 0) break;
      if (round >= C.max_rounds) {
        conv = false;
        break;
      }
      ph_push(D, r, D.wl, np, C.kc, PM_ONLY_B, nullptr, t);
      ph_pull(D, r, D.wl2, nq, C.kc, t);
      G.sync();
      ph_repair(D, r, D.wl, np, RM_B, t);
      ph_repair(D, r, D.wl2, nq, RM_PULL_A, t);
      G.sync();
      if (is_block0_thread0()) D.ctrl->wl_cnt = D.ctrl->wl2_cnt = 0;
      ++round;
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
          if (cnt == 0) break;
          if (round >= C.max_rounds) {
            conv = false;
            break;
          }
          ph_push(D, r, D.wl, cnt, C.kc, PM_ALL, nullptr, t);
          G.sync();
          ph_repair(D, r, D.wl, cnt, RM_PFINAL, t);
          G.sync();
          if (is_block0_thread0()) D.ctrl->wl_cnt = 0;
          ++round;
        }
        if (conv) {
          ph_find_final_cut(D, r);
        }
      }
      G.sync();
    }
  }
  G.sync();
  if (is_block0_thread0()) {
    D.ctrl->wl_cnt = D.ctrl->wl2_cnt = D.ctrl->part_cnt = 0;
    D.ctrl->flag[0] = D.ctrl->flag[1] = 0;
    D.ctrl->rounds = round;
    D.ctrl->ctr[C_ROUNDS] += round;
    D.ctrl->flow = 0;
    if (!conv) D.ctrl->status = DMF_ENOCONV;
  }
  G.sync();
  ph_flow(D, r);
  t.flush(D.ctrl);
}

}  // namespace

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
  int32_t ucap = 0;
  int coop_blocks = 0;
  int num_cu = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  dmf_counters last{};
  int64_t bytes = 0;
  uint64_t launches = 0;
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

int read_int(dmf_engine* E, const int32_t* dptr, int32_t* out) {
  HIP_OK(hipMemcpyAsync(out, dptr, sizeof(int32_t), hipMemcpyDeviceToHost, E->stream));
  HIP_OK(hipStreamSynchronize(E->stream));
  return DMF_OK;
}

#define LAUNCH(kernel, blocks, ...)                                                   \
  do {                                                                                \
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), 0, E->stream, __VA_ARGS__); \
    ++E->launches;                                                                    \
    HIP_OK(hipGetLastError());                                                        \
  } while (0)

int set_dev_int(dmf_engine* E, int32_t* p, int v) {
  LAUNCH(k_set_int, 1, p, v);
  return DMF_OK;
}

// Stepped global relabel: host loop over levels.  Returns visited count.
int relabel_stepped(dmf_engine* E, int mode, int part_cnt, int* visited) {
  Dev& D = E->D;
  int rc;
  if ((rc = set_dev_int(E, &D.ctrl->qcnt[0], 0))) return rc;
  if ((rc = set_dev_int(E, &D.ctrl->qcnt[1], 0))) return rc;
  if ((rc = set_dev_int(E, &D.ctrl->qcnt[2], 0))) return rc;
  if (mode == BFS_FINAL) LAUNCH(k_bfs_init_final, blocks_for(part_cnt), D, part_cnt);
  else LAUNCH(k_bfs_init, blocks_for(E->n), D, mode);
  int lo = 0, lvl = 0;
  for (;;) {
    int32_t cnt = 0;
    if ((rc = read_int(E, &D.ctrl->qcnt[lvl % 3], &cnt))) return rc;
    if (cnt == 0) break;
    if ((rc = set_dev_int(E, &D.ctrl->qcnt[(lvl + 1) % 3], 0))) return rc;
    LAUNCH(k_bfs_level, blocks_for(cnt), D, mode, lo, cnt, lvl, (lvl + 1) % 3);
    lo += cnt;
    ++lvl;
    if (lvl > E->n + 1) return DMF_EDEVICE;
  }
  E->last.bfs_levels += lvl;
  E->last.global_relabels += 1;
  *visited = lo;
  return DMF_OK;
}

int reset_lists(dmf_engine* E) {
  int rc;
  if ((rc = set_dev_int(E, &E->D.ctrl->wl_cnt, 0))) return rc;
  if ((rc = set_dev_int(E, &E->D.ctrl->wl2_cnt, 0))) return rc;
  return DMF_OK;
}

int push_blocks(int cnt) { return blocks_for(cnt); }

// Stepped solve (one kernel per phase, host-checked conditions).
int solve_stepped(dmf_engine* E, const Cfg& C) {
  Dev& D = E->D;
  int rc;
  const bool pp = C.variant == DMF_VARIANT_PP;
  const bool topo = C.variant == DMF_VARIANT_TOPO;
  const int n = E->n;
  if (C.entry == 0) {
    LAUNCH(k_init_residual, blocks_for(std::max(E->m, E->n)), D);
    LAUNCH(k_saturate_source, blocks_for(E->n > 0 ? 1024 : 1), D);
    if (pp) LAUNCH(k_saturate_sink_pp, blocks_for(1024), D);
  } else if (C.entry == 1) {
    LAUNCH(k_upd_lookup, blocks_for(D.k), D);
    if ((rc = sync_ctrl(E))) return rc;
    if (E->h_ctrl->missing || E->h_ctrl->invalid) {  // caller reports; owners cleared
      LAUNCH(k_upd_release, blocks_for(D.k), D);
      return DMF_OK;
    }
    LAUNCH(k_upd_apply, blocks_for(D.k), D, 0);
    if (pp) {
      LAUNCH(k_pp_cut, blocks_for(static_cast<int64_t>(n)), D);
    } else {
      LAUNCH(k_upd_clamp, blocks_for(D.k), D);
      LAUNCH(k_saturate_source, blocks_for(1024), D);
    }
    LAUNCH(k_upd_release, blocks_for(D.k), D);
  }
  int round = 0;
  bool conv = true;
  if ((rc = reset_lists(E))) return rc;
  if (!pp || C.entry == 0) {
    int since = C.interval, visited = 0;
    for (;;) {
      bool fresh = false;
      if (since >= C.interval) {
        if ((rc = relabel_stepped(E, BFS_SINK, 0, &visited))) return rc;
        since = 0;
        fresh = true;
      }
      int32_t cnt = 0;
      if (!topo) {
        if ((rc = set_dev_int(E, &D.ctrl->wl_cnt, 0))) return rc;
        LAUNCH(k_build_wl, blocks_for(visited), D, visited);
        if ((rc = read_int(E, &D.ctrl->wl_cnt, &cnt))) return rc;
        if (cnt == 0) {
          if (fresh) break;
          since = C.interval;
          continue;
        }
      }
      if (round >= C.max_rounds) {
        conv = false;
        break;
      }
      if (topo) {
        if ((rc = set_dev_int(E, &D.ctrl->flag[0], 0))) return rc;
        LAUNCH(k_push, push_blocks(n), D, 0, n, C.kc, PM_ALL);
        LAUNCH(k_repair, push_blocks(n), D, 0, n, RM_ALL);
        int32_t flag = 0;
        if ((rc = read_int(E, &D.ctrl->flag[0], &flag))) return rc;
        if (!flag) break;
      } else {
        LAUNCH(k_push, push_blocks(cnt), D, 1, cnt, C.kc, PM_ALL);
        LAUNCH(k_repair, push_blocks(cnt), D, 1, cnt, RM_ALL);
      }
      ++round;
      ++since;
    }
    if (pp && conv) LAUNCH(k_find_cut, blocks_for(n), D);
  } else {
    for (;;) {
      int visited = 0;
      if ((rc = relabel_stepped(E, BFS_PP, 0, &visited))) return rc;
      if ((rc = reset_lists(E))) return rc;
      LAUNCH(k_build_pp_wl, blocks_for(visited), D, visited);
      if ((rc = sync_ctrl(E))) return rc;
      const int np = E->h_ctrl->wl_cnt, nq = E->h_ctrl->wl2_cnt;
      if (np + nq == 0) break;
      if (round >= C.max_rounds) {
        conv = false;
        break;
      }
      if (np) LAUNCH(k_push, push_blocks(np), D, 1, np, C.kc, PM_ONLY_B);
      if (nq) LAUNCH(k_pull, push_blocks(nq), D, nq, C.kc);
      if (np) LAUNCH(k_repair, push_blocks(np), D, 1, np, RM_B);
      if (nq) LAUNCH(k_repair, push_blocks(nq), D, 2, nq, RM_PULL_A);
      ++round;
    }
    if (conv) {
      if ((rc = set_dev_int(E, &D.ctrl->part_cnt, 0))) return rc;
      LAUNCH(k_find_part, blocks_for(n), D);
      int32_t parts = 0;
      if ((rc = read_int(E, &D.ctrl->part_cnt, &parts))) return rc;
      if (parts > 0) {
        for (;;) {
          int visited = 0;
          if ((rc = relabel_stepped(E, BFS_FINAL, parts, &visited))) return rc;
          if ((rc = set_dev_int(E, &D.ctrl->wl_cnt, 0))) return rc;
          LAUNCH(k_build_final_wl, blocks_for(parts), D, parts);
          int32_t cnt = 0;
          if ((rc = read_int(E, &D.ctrl->wl_cnt, &cnt))) return rc;
          if (cnt == 0) break;
          if (round >= C.max_rounds) {
            conv = false;
            break;
          }
          LAUNCH(k_push, push_blocks(cnt), D, 1, cnt, C.kc, PM_ALL);
          LAUNCH(k_repair, push_blocks(cnt), D, 1, cnt, RM_PFINAL);
          ++round;
        }
        if (conv) LAUNCH(k_find_final_cut, blocks_for(n), D);
      }
    }
  }
  if ((rc = reset_lists(E))) return rc;
  if ((rc = set_dev_int(E, &D.ctrl->part_cnt, 0))) return rc;
  if ((rc = set_dev_int(E, &D.ctrl->flag[0], 0))) return rc;
  HIP_OK(hipMemsetAsync(&D.ctrl->flow, 0, sizeof(long long), E->stream));
  LAUNCH(k_flow, blocks_for(n), D);
  E->last.rounds = round;
  if (!conv) return DMF_ENOCONV;
  return DMF_OK;
}

int solve_persistent(dmf_engine* E, const Cfg& C) {
  Dev D = E->D;
  Cfg c = C;
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
    p = &d;
  }
  c.variant = p->variant;
  c.kc = p->kc > 0 ? p->kc : 1;
  c.interval = p->relabel_interval > 0 ? p->relabel_interval : 1;
  c.max_rounds = p->max_rounds > 0 ? p->max_rounds : 4 * E->n + 1000;
  c.entry = entry;
  return c;
}

// Zero the control block's counters before a solve.
int begin_solve(dmf_engine* E) {
  HIP_OK(hipMemsetAsync(E->D.ctrl, 0, sizeof(Ctrl), E->stream));
  E->last = dmf_counters{};
  E->launches = 0;
  return DMF_OK;
}

int finish_solve(dmf_engine* E, int rc_driver, int64_t* flow) {
  int rc = sync_ctrl(E);
  if (rc) return rc;
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, E->ev0, E->ev1));
  const Ctrl& c = *E->h_ctrl;
  E->last.arc_visits = c.ctr[C_ARC];
  E->last.vertex_visits = c.ctr[C_VERT];
  E->last.pushes = c.ctr[C_PUSH];
  E->last.relabels = c.ctr[C_RELABEL];
  E->last.repairs = c.ctr[C_REPAIR];
  if (c.ctr[C_LEVELS] || c.ctr[C_GRELABEL]) {  // persistent driver counts on device
    E->last.bfs_levels = c.ctr[C_LEVELS];
    E->last.global_relabels = c.ctr[C_GRELABEL];
  }
  if (c.ctr[C_ROUNDS]) E->last.rounds = c.ctr[C_ROUNDS];
  E->last.updates = c.ctr[C_UPD] ? c.ctr[C_UPD] : E->last.updates;
  E->last.atomics = c.ctr[C_ATOMIC];
  E->last.launches = E->launches;
  E->last.device_ms = ms;
  if (c.missing) return DMF_ENOARC;
  if (c.invalid) return DMF_EINVAL;
  if (rc_driver) return rc_driver;
  if (c.status) return c.status;
  if (flow) *flow = c.flow;
  return DMF_OK;
}

int run_solve(dmf_engine* E, const dmf_params* p, int entry, int64_t* flow) {
  const Cfg C = make_cfg(E, p, entry);
  if (C.variant < 0 || C.variant > 2) return DMF_EINVAL;
  int rc = begin_solve(E);
  if (rc) return rc;
  HIP_OK(hipEventRecord(E->ev0, E->stream));
  const bool stepped = (p && p->mode == 1) || E->coop_blocks <= 0;
  int rc_driver = stepped ? solve_stepped(E, C) : solve_persistent(E, C);
  if (rc_driver == DMF_EDEVICE) return rc_driver;
  HIP_OK(hipEventRecord(E->ev1, E->stream));
  if (entry == 1) E->last.updates = static_cast<uint64_t>(E->D.k);
  return finish_solve(E, rc_driver, flow);
}

int ensure_update_capacity(dmf_engine* E, int32_t k) {
  if (k <= E->ucap) return DMF_OK;
  int32_t cap = std::max<int32_t>(k, 1024);
  hipFree(E->d_usrc);
  hipFree(E->d_udst);
  hipFree(E->d_uw);
  hipFree(E->d_uarc);
  E->bytes -= static_cast<int64_t>(E->ucap) * 16;
  E->d_usrc = E->d_udst = E->d_uw = E->d_uarc = nullptr;
  E->ucap = 0;
  int rc;
  if ((rc = dalloc(E, &E->d_usrc, cap)) || (rc = dalloc(E, &E->d_udst, cap)) ||
      (rc = dalloc(E, &E->d_uw, cap)) || (rc = dalloc(E, &E->d_uarc, cap)))
    return rc;
  E->ucap = cap;
  return DMF_OK;
}

}  // namespace

extern "C" {

int dmf_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

void dmf_params_default(const dmf_csr* g, dmf_params* p) {
  if (!p) return;
  p->variant = DMF_VARIANT_DATA;
  p->kc = (g && g->n > 0) ? std::max(1, g->m / g->n) : 1;
  p->relabel_interval = 1;
  p->scheduler = DMF_SCHED_FIXED;
  p->max_rounds = 0;
  p->mode = 0;
}

int dmf_engine_create(const dmf_csr* g, int32_t s, int32_t t, dmf_engine** out) {
  if (!g || !out || g->n < 2 || g->m < 0 || s < 0 || s >= g->n || t < 0 || t >= g->n || s == t)
    return DMF_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return DMF_EDEVICE;
  dmf_engine* E = new (std::nothrow) dmf_engine();
  if (!E) return DMF_ENOMEM;
  E->n = g->n;
  E->m = g->m;
  E->s = s;
  E->t = t;
  Dev& D = E->D;
  D.n = g->n;
  D.m = g->m;
  D.s = s;
  D.t = t;
  const size_t n = g->n, m = g->m;
  uint64_t hcap = 16;
  while (hcap < 2 * static_cast<uint64_t>(m) + 16) hcap <<= 1;
  int32_t *row, *col, *rev, *cap, *res, *rres, *height, *order, *wl, *wl2, *part, *owner, *hval;
  long long* excess;
  uint8_t *side, *pfinal;
  Ctrl* ctrl;
  unsigned long long* hkey;
  int rc = DMF_OK;
  if ((rc = dalloc(E, &row, n + 1)) || (rc = dalloc(E, &col, m)) || (rc = dalloc(E, &rev, m)) ||
      (rc = dalloc(E, &cap, m)) || (rc = dalloc(E, &E->d_cap0, m)) || (rc = dalloc(E, &res, m)) ||
      (rc = dalloc(E, &rres, m)) || (rc = dalloc(E, &excess, n)) || (rc = dalloc(E, &height, n)) ||
      (rc = dalloc(E, &order, n)) || (rc = dalloc(E, &wl, n)) || (rc = dalloc(E, &wl2, n)) ||
      (rc = dalloc(E, &part, n)) || (rc = dalloc(E, &side, n)) || (rc = dalloc(E, &pfinal, n)) ||
      (rc = dalloc(E, &ctrl, 1)) || (rc = dalloc(E, &hkey, hcap)) || (rc = dalloc(E, &hval, hcap)) ||
      (rc = dalloc(E, &owner, m))) {
    dmf_engine_destroy(E);
    return rc;
  }
  D.row = row; D.col = col; D.rev = rev; D.cap = cap; D.res = res; D.rres = rres;
  D.excess = excess; D.height = height; D.order = order; D.wl = wl; D.wl2 = wl2; D.part = part;
  D.side = side; D.pfinal = pfinal; D.ctrl = ctrl; D.hkey = hkey; D.hval = hval;
  D.hmask = hcap - 1; D.owner = owner;
  if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess) {
    dmf_engine_destroy(E);
    return DMF_EDEVICE;
  }
  E->own_stream = true;
  if (hipHostMalloc(reinterpret_cast<void**>(&E->h_ctrl), sizeof(Ctrl), hipHostMallocDefault) != hipSuccess ||
      hipEventCreate(&E->ev0) != hipSuccess || hipEventCreate(&E->ev1) != hipSuccess) {
    dmf_engine_destroy(E);
    return DMF_EDEVICE;
  }
  bool ok = hipMemcpyAsync(row, g->row, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, E->stream) == hipSuccess &&
            hipMemcpyAsync(col, g->col, sizeof(int32_t) * m, hipMemcpyHostToDevice, E->stream) == hipSuccess &&
            hipMemcpyAsync(rev, g->rev, sizeof(int32_t) * m, hipMemcpyHostToDevice, E->stream) == hipSuccess &&
            hipMemcpyAsync(cap, g->cap, sizeof(int32_t) * m, hipMemcpyHostToDevice, E->stream) == hipSuccess &&
            hipMemcpyAsync(E->d_cap0, g->cap, sizeof(int32_t) * m, hipMemcpyHostToDevice, E->stream) == hipSuccess &&
            hipMemsetAsync(res, 0, sizeof(int32_t) * m, E->stream) == hipSuccess &&
            hipMemsetAsync(rres, 0, sizeof(int32_t) * m, E->stream) == hipSuccess &&
            hipMemsetAsync(excess, 0, sizeof(long long) * n, E->stream) == hipSuccess &&
            hipMemsetAsync(height, 0, sizeof(int32_t) * n, E->stream) == hipSuccess &&
            hipMemsetAsync(side, 0, n, E->stream) == hipSuccess &&
            hipMemsetAsync(pfinal, 0, n, E->stream) == hipSuccess &&
            hipMemsetAsync(ctrl, 0, sizeof(Ctrl), E->stream) == hipSuccess &&
            hipMemsetAsync(hkey, 0xff, sizeof(unsigned long long) * hcap, E->stream) == hipSuccess &&
            hipMemsetAsync(owner, 0xff, sizeof(int32_t) * m, E->stream) == hipSuccess;
  if (!ok) {
    dmf_engine_destroy(E);
    return DMF_EDEVICE;
  }
  hipLaunchKernelGGL(k_build_hash, dim3(blocks_for(n)), dim3(kBlock), 0, E->stream, D);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(E->stream) != hipSuccess) {
    dmf_engine_destroy(E);
    return DMF_EDEVICE;
  }
  // Cooperative grid: as many blocks as can be co-resident.
  int dev = 0, per_cu = 0;
  hipDeviceProp_t prop{};
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
      prop.cooperativeLaunch &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_solve_persistent),
                                                   kBlock, 0) == hipSuccess &&
      per_cu > 0) {
    E->num_cu = prop.multiProcessorCount;
    E->coop_blocks = prop.multiProcessorCount * std::min(per_cu, 4);
  }
  *out = E;
  return DMF_OK;
}

int dmf_engine_destroy(dmf_engine* E) {
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
  hipFree(D.order);
  hipFree(D.wl);
  hipFree(D.wl2);
  hipFree(D.part);
  hipFree(D.side);
  hipFree(D.pfinal);
  hipFree(D.ctrl);
  hipFree(D.hkey);
  hipFree(D.hval);
  hipFree(D.owner);
  hipFree(E->d_usrc);
  hipFree(E->d_udst);
  hipFree(E->d_uw);
  hipFree(E->d_uarc);
  if (E->h_ctrl) hipHostFree(E->h_ctrl);
  if (E->ev0) hipEventDestroy(E->ev0);
  if (E->ev1) hipEventDestroy(E->ev1);
  if (E->own_stream && E->stream) hipStreamDestroy(E->stream);
  delete E;
  return DMF_OK;
}

int dmf_engine_set_stream(dmf_engine* E, void* hip_stream) {
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
}

int dmf_engine_reset(dmf_engine* E) {
  if (!E) return DMF_EINVAL;
  Dev& D = E->D;
  HIP_OK(hipMemcpyAsync(D.cap, E->d_cap0, sizeof(int32_t) * E->m, hipMemcpyDeviceToDevice, E->stream));
  HIP_OK(hipMemsetAsync(D.res, 0, sizeof(int32_t) * E->m, E->stream));
  HIP_OK(hipMemsetAsync(D.rres, 0, sizeof(int32_t) * E->m, E->stream));
  HIP_OK(hipMemsetAsync(D.excess, 0, sizeof(long long) * E->n, E->stream));
  HIP_OK(hipMemsetAsync(D.side, 0, E->n, E->stream));
  HIP_OK(hipStreamSynchronize(E->stream));
  return DMF_OK;
}

int dmf_engine_static_solve(dmf_engine* E, const dmf_params* p, int64_t* flow) {
  if (!E) return DMF_EINVAL;
  E->D.k = 0;
  return run_solve(E, p, 0, flow);
}

int dmf_engine_apply_batch_device(dmf_engine* E, int32_t k, const int32_t* d_src,
                                  const int32_t* d_dst, const int32_t* d_w, const dmf_params* p,
                                  int64_t* flow) {
  if (!E || k < 0 || (k > 0 && (!d_src || !d_dst || !d_w))) return DMF_EINVAL;
  int rc = ensure_update_capacity(E, std::max(k, 1));
  if (rc) return rc;
  E->D.usrc = d_src;
  E->D.udst = d_dst;
  E->D.uw = d_w;
  E->D.uarc = E->d_uarc;
  E->D.k = k;
  rc = run_solve(E, p, 1, flow);
  E->D.k = 0;
  return rc;
}

int dmf_engine_apply_batch(dmf_engine* E, int32_t k, const int32_t* src, const int32_t* dst,
                           const int32_t* w, const dmf_params* p, int64_t* flow) {
  if (!E || k < 0 || (k > 0 && (!src || !dst || !w))) return DMF_EINVAL;
  int rc = ensure_update_capacity(E, std::max(k, 1));
  if (rc) return rc;
  if (k > 0) {
    HIP_OK(hipMemcpyAsync(E->d_usrc, src, sizeof(int32_t) * k, hipMemcpyHostToDevice, E->stream));
    HIP_OK(hipMemcpyAsync(E->d_udst, dst, sizeof(int32_t) * k, hipMemcpyHostToDevice, E->stream));
    HIP_OK(hipMemcpyAsync(E->d_uw, w, sizeof(int32_t) * k, hipMemcpyHostToDevice, E->stream));
  }
  return dmf_engine_apply_batch_device(E, k, E->d_usrc, E->d_udst, E->d_uw, p, flow);
}

int dmf_engine_flow(dmf_engine* E, int64_t* flow) {
  if (!E || !flow) return DMF_EINVAL;
  HIP_OK(hipMemsetAsync(&E->D.ctrl->flow, 0, sizeof(long long), E->stream));
  hipLaunchKernelGGL(k_flow, dim3(blocks_for(E->n)), dim3(kBlock), 0, E->stream, E->D);
  HIP_OK(hipGetLastError());
  int rc = sync_ctrl(E);
  if (rc) return rc;
  *flow = E->h_ctrl->flow;
  return DMF_OK;
}

int dmf_engine_min_cut(dmf_engine* E, uint8_t* side) {
  if (!E || !side) return DMF_EINVAL;
  Dev& D = E->D;
  int32_t* mark = D.wl2;  // scratch between solves
  int rc = DMF_OK;
  do {
    if ((rc = set_dev_int(E, &D.ctrl->qcnt[0], 0))) break;
    if ((rc = set_dev_int(E, &D.ctrl->qcnt[1], 0))) break;
    if ((rc = set_dev_int(E, &D.ctrl->qcnt[2], 0))) break;
    LAUNCH(k_cut_init, blocks_for(E->n), D, mark);
    int lo = 0, lvl = 0;
    for (;;) {
      int32_t cnt = 0;
      if ((rc = read_int(E, &D.ctrl->qcnt[lvl % 3], &cnt))) break;
      if (cnt == 0) break;
      if ((rc = set_dev_int(E, &D.ctrl->qcnt[(lvl + 1) % 3], 0))) break;
      LAUNCH(k_cut_level, blocks_for(cnt), D, mark, lo, cnt, (lvl + 1) % 3);
      lo += cnt;
      ++lvl;
    }
    if (rc) break;
    std::vector<int32_t> h(E->n);
    if (hipMemcpyAsync(h.data(), mark, sizeof(int32_t) * E->n, hipMemcpyDeviceToHost, E->stream) != hipSuccess ||
        hipStreamSynchronize(E->stream) != hipSuccess) {
      rc = DMF_EDEVICE;
      break;
    }
    for (int32_t v = 0; v < E->n; ++v) side[v] = h[v] ? 1 : 0;
  } while (false);
  return rc;
}

int dmf_engine_read_state(dmf_engine* E, int32_t* res, int32_t* cap, int64_t* excess,
                          int32_t* height) {
  if (!E) return DMF_EINVAL;
  if (res) HIP_OK(hipMemcpyAsync(res, E->D.res, sizeof(int32_t) * E->m, hipMemcpyDeviceToHost, E->stream));
  if (cap) HIP_OK(hipMemcpyAsync(cap, E->D.cap, sizeof(int32_t) * E->m, hipMemcpyDeviceToHost, E->stream));
  if (excess)
    HIP_OK(hipMemcpyAsync(excess, E->D.excess, sizeof(long long) * E->n, hipMemcpyDeviceToHost, E->stream));
  if (height)
    HIP_OK(hipMemcpyAsync(height, E->D.height, sizeof(int32_t) * E->n, hipMemcpyDeviceToHost, E->stream));
  HIP_OK(hipStreamSynchronize(E->stream));
  return DMF_OK;
}

int dmf_engine_counters(dmf_engine* E, dmf_counters* c) {
  if (!E || !c) return DMF_EINVAL;
  *c = E->last;
  return DMF_OK;
}

int64_t dmf_engine_device_bytes(const dmf_engine* E) { return E ? E->bytes : 0; }

int dmf_recalculate_max_flow(const dmf_csr* g, int32_t k, const int32_t* src, const int32_t* dst,
                             const int32_t* w, int32_t s, int32_t t, int32_t kc, int32_t variant,
                             int64_t* flow, double* ms_static, double* ms_batch) {
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
}

int dmf_static_recalculate(const dmf_csr* g, int32_t k, const int32_t* src, const int32_t* dst,
                           const int32_t* w, int32_t s, int32_t t, int32_t kc, int32_t variant,
                           int64_t* flow, double* ms) {
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
      D.k = k;
      hipLaunchKernelGGL(k_upd_lookup, dim3(blocks_for(k)), dim3(kBlock), 0, E->stream, D);
      hipLaunchKernelGGL(k_upd_apply, dim3(blocks_for(k)), dim3(kBlock), 0, E->stream, D, 1);
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
}

}  // extern "C"
