// Learned scheduler: a small FP32 MLP (8 -> 32 -> 32 -> 4) evaluated by one
// wave64 on counters that live on the device.  It answers two questions:
//   * at the start of a batch: which variant (data, topo, pp) takes it;
//   * at the end of every push/repair round: fire the global relabel now?
// No reference counterpart: the reference fires the global relabel after every
// round and fixes the variant per binary.
//
// Features (all scaled to roughly [0,1]):
//   f0 active vertices after the round      log2(1+a) / log2(1+n)
//      (at batch start, where f0-f3 have no round to describe: the graph's
//      size, log2(1+n) / 32)
//   f1 pushes in the round                  log2(1+p) / log2(1+m)
//   f2 local relabels in the round          log2(1+r) / log2(1+n)
//   f3 change of excess(t) in the round     sign(d) log2(1+|d|) / 32
//   f4 batch size                           log2(1+k) / log2(1+m)
//   f5 fraction of capacity increases in the batch
//   f6 rounds since the last global relabel min(s,16) / 16
//   f7 levels of the last global relabel    log2(1+L) / 8
// Outputs: out[0] > 0 means relabel now; argmax(out[1..3]) is the variant.
#pragma once

#include "phases.hpp"

namespace dmfdev {

constexpr int kFeat = 8;
constexpr int kHid = 32;
constexpr int kOut = 4;
constexpr int kSchedParams = kFeat * kHid + kHid + kHid * kHid + kHid + kHid * kOut + kOut;

// Row-major weights in this order: w1[kFeat][kHid] b1[kHid] w2[kHid][kHid]
// b2[kHid] w3[kHid][kOut] b3[kOut].
struct SchedWeights {
  float w1[kFeat][kHid];
  float b1[kHid];
  float w2[kHid][kHid];
  float b2[kHid];
  float w3[kHid][kOut];
  float b3[kOut];
};
static_assert(sizeof(SchedWeights) == sizeof(float) * kSchedParams, "weight layout");

struct SchedState {
  unsigned long long prev_push;
  unsigned long long prev_relabel;
  long long prev_sink;
  float feat[kFeat];
  float out[kOut];
  int32_t relabel;  // decision: fire the global relabel before the next round
  int32_t variant;  // decision at batch start
  int32_t valid;    // outputs finite
  int32_t tag;      // tag of the round whose decision this is (cascade launches)
};

__device__ __forceinline__ float log2p1(double x) { return static_cast<float>(log2(1.0 + x)); }

// phase 0 = batch start (counters of the batch so far are ignored), 1 = round end.
// active: vertices listed for the round just run (both lists and both queues),
// as the host read them before the round's pushes.  inc_frac < 0: the
// fraction of capacity increases is read from the control block (the batch's
// lookup counted them on the device).  Called by one whole wave
// (64 lanes).  HOIST (weights in global memory): every weight a lane needs is
// loaded before the first product, so the evaluation costs a few memory
// round trips, not one per hidden unit, at the price of about 115 registers.
// Without it (weights already in LDS) the products read them as they go,
// which keeps the caller's register count, and its occupancy, low.
template <bool HOIST>
__device__ void sched_eval(const Dev& D, const SchedWeights* W, SchedState* S, int phase, int k, float inc_frac,
                           int rounds_since, int last_levels, double active) {
  const int lane = lane_id();
  const int hl = lane < kHid ? lane : 0;  // lanes >= kHid compute a copy of unit 0
  const int ol = lane < kOut ? lane : 0;
  float w1[kFeat], w2[HOIST ? kHid : 1], w3[HOIST ? kHid : 1];
#pragma unroll
  for (int i = 0; i < kFeat; ++i) w1[i] = W->w1[i][hl];
  if (HOIST) {
#pragma unroll
    for (int i = 0; i < kHid; ++i) w2[i] = W->w2[i][hl];
#pragma unroll
    for (int i = 0; i < kHid; ++i) w3[i] = W->w3[i][ol];
  }
  const float b1 = W->b1[hl], b2 = W->b2[hl], b3 = W->b3[ol];
  // totals of the work counters (kSlots copies each); uncached, so a caller
  // inside a cooperative launch sees the other blocks' atomics
  unsigned long long pu = ldru64(&D.ctrs[C_PUSH * kSlots + lane]);
  unsigned long long re = ldru64(&D.ctrs[C_RELABEL * kSlots + lane]);
  pu = static_cast<unsigned long long>(wave_sum_i64(static_cast<long long>(pu)));
  re = static_cast<unsigned long long>(wave_sum_i64(static_cast<long long>(re)));
  const long long sink = ldr64(&D.excess[D.t]);
  float f[kFeat];
  const float ln = log2p1(D.n), lm = log2p1(D.m);
  if (phase == 0) {
    // batch start: no round counters yet; slot 0 carries the graph's size
    f[0] = log2p1(D.n) / 32.f;
    f[1] = f[2] = f[3] = 0.f;
    f[6] = 0.f;
    f[7] = log2p1(last_levels) / 8.f;
  } else {
    const long long ds = sink - S->prev_sink;
    f[0] = log2p1(active) / ln;
    f[1] = log2p1(static_cast<double>(pu - S->prev_push)) / lm;
    f[2] = log2p1(static_cast<double>(re - S->prev_relabel)) / ln;
    f[3] = (ds < 0 ? -1.f : 1.f) * log2p1(static_cast<double>(ds < 0 ? -ds : ds)) / 32.f;
    f[6] = static_cast<float>(rounds_since < 16 ? rounds_since : 16) / 16.f;
    f[7] = log2p1(last_levels) / 8.f;
  }
  f[4] = log2p1(k) / lm;
  f[5] = inc_frac >= 0.f ? inc_frac : (k > 0 ? static_cast<float>(ldr(&D.ctrl->inc)) / static_cast<float>(k) : 0.5f);
  // layer 1: lane j < kHid computes hidden unit j
  float h1 = b1;
#pragma unroll
  for (int i = 0; i < kFeat; ++i) h1 += f[i] * w1[i];
  h1 = h1 > 0.f ? h1 : 0.f;
  float h2 = b2;
#pragma unroll 8
  for (int i = 0; i < kHid; ++i) h2 += __shfl(h1, i) * (HOIST ? w2[HOIST ? i : 0] : W->w2[i][hl]);
  h2 = h2 > 0.f ? h2 : 0.f;
  float o = b3;
#pragma unroll 8
  for (int i = 0; i < kHid; ++i) o += __shfl(h2, i) * (HOIST ? w3[HOIST ? i : 0] : W->w3[i][ol]);
  float outs[kOut];
#pragma unroll
  for (int j = 0; j < kOut; ++j) outs[j] = __shfl(o, j);
  if (lane == 0) {
    bool ok = true;
    for (int j = 0; j < kOut; ++j) ok = ok && isfinite(outs[j]);
    for (int i = 0; i < kFeat; ++i) S->feat[i] = f[i];
    for (int j = 0; j < kOut; ++j) S->out[j] = outs[j];
    S->valid = ok ? 1 : 0;
    S->relabel = outs[0] > 0.f ? 1 : 0;
    int best = 0;
    for (int j = 1; j < 3; ++j)
      if (outs[1 + j] > outs[1 + best]) best = j;
    S->variant = best;
    S->prev_push = pu;
    S->prev_relabel = re;
    S->prev_sink = sink;
  }
}

__global__ void __launch_bounds__(kWave) k_sched(Dev D, const SchedWeights* W, SchedState* S, int phase,
                                                 int k, float inc_frac, int rounds_since, int last_levels,
                                                 double active) {
  sched_eval<true>(D, W, S, phase, k, inc_frac, rounds_since, last_levels, active);
}

// Scheduler inputs of a round-end decision taken inside a cascade launch.
struct SchedRound {
  int k;
  float inc_frac;
  int rounds_since;
  int last_levels;
  double active;
  int32_t tag;  // the round's tag: the decision is taken once per tag
  // the chain was queued behind the relabel, before the host read the round's
  // list: a round k_list_snap skipped takes no decision, and active and
  // last_levels come from the control block (r_cnt, r_wpk; the cursor's level)
  int32_t spec;
};

// One cascade sub-round: pushes from the list in slot `slot` (sizes read from
// the control block, written by the previous pass) and lists the vertices for
// the next sub-round in slot (slot+1)%3 with stamp `stamp`.  Block 0 clears
// slot (slot+2)%3, the input of the sub-round before, for the sub-round
// after.  An empty list returns at once, so a chain of launches stops doing
// work when the excess has settled.  W != nullptr: the first launch of the
// round that finds its list empty, i.e. the launch right after the chain's
// last working sub-round, also takes the scheduler's round-end decision
// (block 0 stages the weights in LDS, wave 0 evaluates; every earlier launch
// of the stream has completed, so the round's counters are final) and marks
// it with the round's tag.
__global__ void __launch_bounds__(kBlock) k_cascade(Dev D, int slot, int kc, int stamp, int relist,
                                                    const SchedWeights* W, SchedState* S, SchedRound a) {
  // a chain queued behind a relabel whose round k_list_snap skipped does
  // nothing: while the BFS goes on, the ring slots are the BFS's
  if (a.spec && D.ctrl->r_skip) return;  // uniform
  // block 0 loads the decision tag beside the list sizes (one round trip for
  // the empty launches that end every chain)
  const bool decider = W != nullptr && blockIdx.x == 0;
  const int32_t tag_now = decider ? ldr(&S->tag) : a.tag;
  const Ctrl* cc = D.ctrl;
  const int cnt = cc->qcnt[slot];
  const unsigned long long wpk = cc->wq[slot];
  if (is_block0_thread0()) {
    D.ctrl->qcnt[(slot + 2) % 3] = 0;
    D.ctrl->wq[(slot + 2) % 3] = 0;
  }
  if (cnt == 0 && wpk == 0) {
    // uniform over block 0: the tag was written by an earlier launch
    if (decider && tag_now != a.tag) {
      __shared__ SchedWeights wl;
      const float* src = reinterpret_cast<const float*>(W);
      float* dst = reinterpret_cast<float*>(&wl);
      for (int i = threadIdx.x; i < kSchedParams; i += blockDim.x) dst[i] = src[i];
      __syncthreads();
      if (threadIdx.x < kWave) {
        const double active = a.spec ? static_cast<double>(cc->r_cnt) + static_cast<double>(cc->r_wpk >> 32)
                                     : a.active;
        const int last_levels = a.spec ? cc->bcur[0][1] : a.last_levels;
        sched_eval<false>(D, &wl, S, 1, a.k, a.inc_frac, a.rounds_since, last_levels, active);
        if (threadIdx.x == 0) str(&S->tag, a.tag);
      }
    }
    return;
  }
  Tally t;
  ph_push<false>(D, make_range(), cascade_list(D, slot), cnt, kc, PM_ALL, nullptr, wlist(D, slot, wpk), t,
                 Cascade{(slot + 1) % 3, stamp, relist});
  t.flush(D.ctrs);
}

}  // namespace dmfdev
