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
  int32_t pad;
};

__device__ __forceinline__ float log2p1(double x) { return static_cast<float>(log2(1.0 + x)); }

// phase 0 = batch start (counters of the batch so far are ignored), 1 = round end.
// active: vertices listed for the round just run (both lists and both queues),
// as the host read them before the round's pushes.
__global__ void __launch_bounds__(kWave) k_sched(Dev D, const SchedWeights* W, SchedState* S, int phase,
                                                 int k, float inc_frac, int rounds_since, int last_levels,
                                                 double active) {
  const int lane = threadIdx.x;
  // totals of the work counters (kSlots copies each)
  unsigned long long pu = D.ctrs[C_PUSH * kSlots + lane];
  unsigned long long re = D.ctrs[C_RELABEL * kSlots + lane];
  pu = static_cast<unsigned long long>(wave_sum_i64(static_cast<long long>(pu)));
  re = static_cast<unsigned long long>(wave_sum_i64(static_cast<long long>(re)));
  const long long sink = D.excess[D.t];
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
  f[5] = inc_frac;
  // layer 1: lane j < kHid computes hidden unit j
  float h1 = 0.f;
  if (lane < kHid) {
    h1 = W->b1[lane];
#pragma unroll
    for (int i = 0; i < kFeat; ++i) h1 += f[i] * W->w1[i][lane];
    h1 = h1 > 0.f ? h1 : 0.f;
  }
  float h2 = lane < kHid ? W->b2[lane] : 0.f;
  for (int i = 0; i < kHid; ++i) {
    const float hi = __shfl(h1, i);
    if (lane < kHid) h2 += hi * W->w2[i][lane];
  }
  h2 = (lane < kHid && h2 > 0.f) ? h2 : 0.f;
  float o = lane < kOut ? W->b3[lane] : 0.f;
  for (int i = 0; i < kHid; ++i) {
    const float hi = __shfl(h2, i);
    if (lane < kOut) o += hi * W->w3[i][lane];
  }
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

}  // namespace dmfdev
