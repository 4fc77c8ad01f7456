// dmf_serve_bench: streaming update batches against resident graphs.
//
//   dmf_serve_bench [--scale=23] [--ef=16] [--seed=5] [--batches=60] [--warmup=3]
//                   [--kmin=1000] [--kmax=1000000] [--inc=0.4] [--load=0.7 | --rate=R]
//                   [--sched=learned|plain] [--verify-every=10] [--sessions=1] [--json]
//
// Workload: directed R-MAT (a/b/c/d = 0.57/0.19/0.19/0.05, capacities uniform
// [1,100]) resident in one server session.  Batch sizes are log-uniform in
// [kmin, kmax] arc updates, 40% increases / 60% decreases (new capacity
// uniform [1,100] on the chosen side), arriving as a Poisson process.  The
// first `warmup` batches are served closed-loop and set the arrival rate:
// rate = load / mean service time (or --rate batches/s).  Latency is measured
// from a batch's arrival (submit) to its result, on the server clock.
// Verification: after the timed run, every verify-every-th batch (and the
// last) is recomputed from scratch (dmf_static_recalculate on the original
// graph plus every arc changed so far) and must equal the served flow.
// With --sessions=K > 1, K graphs (seeds seed, seed+1, ...) are resident at
// once, each with its own Poisson stream at rate / K (the warm-up rate splits
// the load over the sessions), so solves of different sessions run on the
// device together.  Reported besides: how many batches' solves overlapped
// another session's, and the most solves in flight at once.  Each session's
// batches are then replayed, in order, on a solo engine with the server's
// parameters, and every served flow must equal its solo flow (the result is
// the same alone and in the mix), on top of the recompute checks.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "dmf.h"

namespace {

uint64_t splitmix(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double u01(uint64_t& x) { return static_cast<double>(splitmix(x) >> 11) * (1.0 / 9007199254740992.0); }

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

double pct_rank(std::vector<double> x, double q) {  // nearest rank
  if (x.empty()) return 0;
  std::sort(x.begin(), x.end());
  size_t r = static_cast<size_t>(std::ceil(q / 100.0 * x.size()));
  return x[std::min(x.size(), std::max<size_t>(r, 1)) - 1];
}

[[noreturn]] void die(const char* what, int rc) {
  std::fprintf(stderr, "dmf_serve_bench: %s: %s (%d)\n", what, dmf_strerror(rc), rc);
  std::exit(2);
}

struct Batch {
  std::vector<int32_t> src, dst, w;
};

// One resident graph and its stream of batches.
struct Session {
  uint64_t seed = 0;
  int64_t m_in = 0;
  int32_t *pu = nullptr, *pv = nullptr, *pw = nullptr;
  dmf_csr g{};
  int32_t s = 0, t = 0;
  std::vector<Batch> B;
  std::vector<int64_t> counts;
  int32_t sid = 0;
  int64_t f0 = 0;
  double svc = 0, rate = 0;
  std::vector<dmf_result> R;
};

}  // namespace

int main(int argc, char** argv) {
  int scale = 23, ef = 16, nbatch = 60, warmup = 3, verify_every = 10, nsess = 1;
  uint64_t seed = 5;
  int64_t kmin = 1000, kmax = 1000000;
  double inc = 0.4, load = 0.7, rate = 0;
  bool learned = true, json = false;
  for (int i = 1; i < argc; ++i) {
    const char* x = argv[i];
    auto val = [&](const char* key) -> const char* {
      const size_t l = std::strlen(key);
      return std::strncmp(x, key, l) == 0 ? x + l : nullptr;
    };
    const char* a;
    if ((a = val("--scale="))) scale = std::atoi(a);
    else if ((a = val("--ef="))) ef = std::atoi(a);
    else if ((a = val("--seed="))) seed = std::strtoull(a, nullptr, 10);
    else if ((a = val("--batches="))) nbatch = std::atoi(a);
    else if ((a = val("--warmup="))) warmup = std::atoi(a);
    else if ((a = val("--kmin="))) kmin = std::atoll(a);
    else if ((a = val("--kmax="))) kmax = std::atoll(a);
    else if ((a = val("--inc="))) inc = std::atof(a);
    else if ((a = val("--load="))) load = std::atof(a);
    else if ((a = val("--rate="))) rate = std::atof(a);
    else if ((a = val("--sched="))) learned = std::strcmp(a, "plain") != 0;
    else if ((a = val("--verify-every="))) verify_every = std::atoi(a);
    else if ((a = val("--sessions="))) nsess = std::atoi(a);
    else if (!std::strcmp(x, "--json")) json = true;
    else {
      std::fprintf(stderr, "dmf_serve_bench: unknown argument %s\n", x);
      return 2;
    }
  }
  if (scale < 2 || ef < 1 || nbatch < 1 || warmup < 1 || kmin < 1 || kmax < kmin || load <= 0 || rate < 0 ||
      nsess < 1 || nsess > 16) {
    std::fprintf(stderr, "dmf_serve_bench: invalid arguments\n");
    return 2;
  }
  if (dmf_device_count() < 1) {
    std::fprintf(stderr, "dmf_serve_bench: no GPU visible\n");
    return 2;
  }
  const int total = warmup + nbatch;
  std::vector<Session> SS(nsess);
  int rc = 0;
  int64_t kbig = 0;
  for (int i = 0; i < nsess; ++i) {
    Session& Q = SS[i];
    Q.seed = seed + static_cast<uint64_t>(i);
    // ---- graph
    double t = now_s();
    if ((rc = dmf_gen_rmat(scale, ef, 0.57, 0.19, 0.19, Q.seed, 1, 100, &Q.m_in, &Q.pu, &Q.pv, &Q.pw)))
      die("gen_rmat", rc);
    std::printf("[serve] rmat-%d ef%d seed %llu: %lld edges (%.1f s)\n", scale, ef,
                static_cast<unsigned long long>(Q.seed), static_cast<long long>(Q.m_in), now_s() - t);
    std::fflush(stdout);
    t = now_s();
    if ((rc = dmf_build_csr(Q.m_in, Q.pu, Q.pv, Q.pw, &Q.g))) die("build_csr", rc);
    if ((rc = dmf_pick_source_sink(Q.g.n, Q.m_in, Q.pu, Q.pv, &Q.s, &Q.t))) die("pick_source_sink", rc);
    std::printf("[serve] csr n=%d m=%d s=%d t=%d (%.1f s)\n", Q.g.n, Q.g.m, Q.s, Q.t, now_s() - t);
    std::fflush(stdout);
    // ---- batches (log-uniform sizes, successive)
    t = now_s();
    Q.B.resize(total);
    Q.counts.resize(total);
    uint64_t st = Q.seed * 7919 + 11;
    std::vector<int32_t> cur(Q.pw, Q.pw + Q.m_in);
    const double l0 = std::log(static_cast<double>(kmin)), l1 = std::log(static_cast<double>(kmax));
    for (int b = 0; b < total; ++b) {
      Q.counts[b] = std::min<int64_t>(Q.m_in, static_cast<int64_t>(std::exp(l0 + (l1 - l0) * u01(st))));
      int32_t k = 0;
      int32_t *a = nullptr, *d = nullptr, *c = nullptr;
      if ((rc = dmf_gen_updates(Q.m_in, Q.pu, Q.pv, cur.data(), Q.counts[b], inc, 1, 100, Q.seed * 1000003ull + b,
                                1, &k, &a, &d, &c)))
        die("gen_updates", rc);
      Q.B[b].src.assign(a, a + k);
      Q.B[b].dst.assign(d, d + k);
      Q.B[b].w.assign(c, c + k);
      dmf_free(a);
      dmf_free(d);
      dmf_free(c);
    }
    for (auto& b : Q.B) kbig = std::max<int64_t>(kbig, static_cast<int64_t>(b.src.size()));
    std::printf("[serve] %d batches generated (%.1f s)\n", total, now_s() - t);
    std::fflush(stdout);
    Q.R.resize(total);
  }
  // ---- server
  dmf_server_config cfg;
  dmf_server_config_default(&cfg);
  cfg.params.scheduler = learned ? DMF_SCHED_LEARNED : DMF_SCHED_FIXED;
  cfg.params.mode = 2;
  // room for 4 of the largest batches, and 2 more per further session
  cfg.staging_bytes = std::max<int64_t>(int64_t{64} << 20, kbig * 12 * (4 + 2 * (nsess - 1)));
  dmf_server* S = nullptr;
  if ((rc = dmf_server_create(&cfg, &S))) die("server_create", rc);
  for (Session& Q : SS) {
    const double t = now_s();
    if ((rc = dmf_server_open(S, &Q.g, Q.s, Q.t, static_cast<int32_t>(kbig), &Q.sid, &Q.f0))) die("server_open", rc);
    std::printf("[serve] session %d open, static flow %lld (%.1f s)\n", Q.sid, static_cast<long long>(Q.f0),
                now_s() - t);
    std::fflush(stdout);
  }
  // warm-up: closed loop, one session after another; sets the arrival rates
  double svc_all = 0;
  for (Session& Q : SS) {
    for (int b = 0; b < warmup; ++b) {
      uint64_t tk = 0;
      if ((rc = dmf_server_submit(S, Q.sid, static_cast<int32_t>(Q.B[b].src.size()), Q.B[b].src.data(),
                                  Q.B[b].dst.data(), Q.B[b].w.data(), &tk)))
        die("submit", rc);
      if ((rc = dmf_server_wait(S, tk, -1, &Q.R[b]))) die("wait", rc);
      if (Q.R[b].status) die("warm-up batch", Q.R[b].status);
      Q.svc += Q.R[b].done_ms - Q.R[b].start_ms;
    }
    Q.svc /= warmup;
    svc_all += Q.svc;
  }
  svc_all /= nsess;
  // total offered rate: load / mean service time (or --rate), split evenly
  if (rate <= 0) rate = load * 1000.0 / std::max(svc_all, 1e-3);
  for (Session& Q : SS) Q.rate = rate / nsess;
  std::printf("[serve] warm-up mean service %.2f ms -> Poisson arrivals at %.2f batches/s (%d session%s)\n", svc_all,
              rate, nsess, nsess > 1 ? "s" : "");
  std::fflush(stdout);
  // ---- timed: per session, a producer thread submits at Poisson arrival
  // times and a waiter thread collects the results in order.
  // tickets[i][b]: 0 until the producer has submitted batch b (~0 if refused)
  std::vector<std::vector<std::atomic<uint64_t>>> tickets(nsess);
  std::vector<std::vector<int>> sub_rc(nsess, std::vector<int>(total, 0));
  for (auto& v : tickets) {
    std::vector<std::atomic<uint64_t>> w(total);
    v.swap(w);
    for (auto& x : v) x.store(0);
  }
  std::atomic<int> done{0};
  std::atomic<int> fail{0};
  const double t_start = now_s();
  std::vector<std::thread> threads;
  for (int i = 0; i < nsess; ++i) {
    threads.emplace_back([&, i] {
      Session& Q = SS[i];
      uint64_t st = Q.seed * 104729 + 3;
      double at = 0;
      for (int b = warmup; b < total; ++b) {
        at += -std::log(1.0 - u01(st)) / Q.rate;
        const double wait = t_start + at - now_s();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        uint64_t tk = 0;
        sub_rc[i][b] = dmf_server_submit(S, Q.sid, static_cast<int32_t>(Q.B[b].src.size()), Q.B[b].src.data(),
                                         Q.B[b].dst.data(), Q.B[b].w.data(), &tk);
        tickets[i][b].store(sub_rc[i][b] ? ~uint64_t{0} : tk, std::memory_order_release);
      }
    });
    threads.emplace_back([&, i] {
      Session& Q = SS[i];
      for (int b = warmup; b < total; ++b) {
        uint64_t tk;
        while ((tk = tickets[i][b].load(std::memory_order_acquire)) == 0)
          std::this_thread::sleep_for(std::chrono::microseconds(50));
        int r = sub_rc[i][b];
        if (!r) r = dmf_server_wait(S, tk, -1, &Q.R[b]);
        if (!r) r = Q.R[b].status;
        if (r) {
          std::fprintf(stderr, "dmf_serve_bench: session %d batch %d: %s (%d)\n", i, b, dmf_strerror(r), r);
          fail.store(1);
          return;
        }
        const int d = ++done;
        if (d % 10 == 0) {
          std::printf("[serve] %d/%d served\n", d, nbatch * nsess);
          std::fflush(stdout);
        }
      }
    });
  }
  for (auto& th : threads) th.join();
  if (fail.load()) die("served batch", DMF_EDEVICE);
  dmf_server_stats stats;
  dmf_server_get_stats(S, &stats);
  dmf_server_destroy(S);
  // ---- report
  std::vector<double> lat, svcs, queue;
  double arcs = 0, dev = 0, kk = 0, first_submit = 1e300, last_done = 0;
  struct Span {
    double a, b;
    int sess;
  };
  std::vector<Span> spans;
  for (int i = 0; i < nsess; ++i)
    for (int b = warmup; b < total; ++b) {
      const dmf_result& r = SS[i].R[b];
      lat.push_back(r.done_ms - r.submit_ms);
      svcs.push_back(r.done_ms - r.start_ms);
      queue.push_back(r.start_ms - r.submit_ms);
      arcs += static_cast<double>(r.arc_visits);
      dev += r.device_ms;
      kk += r.k;
      first_submit = std::min(first_submit, r.submit_ms);
      last_done = std::max(last_done, r.done_ms);
      spans.push_back({r.start_ms, r.done_ms, i});
    }
  const int nserved = nbatch * nsess;
  const double bps = nserved * 1000.0 / (last_done - first_submit);
  // solves of different sessions that ran at the same time (server clock)
  int overlapped = 0, peak = 0;
  for (size_t x = 0; x < spans.size(); ++x) {
    int inflight = 1;
    bool other = false;
    for (size_t y = 0; y < spans.size(); ++y) {
      if (x == y) continue;
      const bool meet = spans[y].a < spans[x].b && spans[x].a < spans[y].b;
      if (meet && spans[y].sess != spans[x].sess) other = true;
      if (spans[y].a <= spans[x].a && spans[x].a < spans[y].b) ++inflight;
    }
    overlapped += other ? 1 : 0;
    peak = std::max(peak, inflight);
  }
  // ---- verification (after the timed run)
  int verified = 0, mismatches = 0, replayed = 0;
  for (int i = 0; i < nsess; ++i) {
    Session& Q = SS[i];
    if (verify_every > 0) {
      std::vector<int32_t> cur(Q.pw, Q.pw + Q.m_in);
      for (int b = 0; b < total; ++b) {
        int32_t k = 0;
        int32_t *a = nullptr, *d = nullptr, *c = nullptr;
        if ((rc = dmf_gen_updates(Q.m_in, Q.pu, Q.pv, cur.data(), Q.counts[b], inc, 1, 100, Q.seed * 1000003ull + b,
                                  1, &k, &a, &d, &c)))
          die("gen_updates", rc);
        dmf_free(a);
        dmf_free(d);
        dmf_free(c);
        const bool check = b >= warmup && ((b - warmup + 1) % verify_every == 0 || b == total - 1);
        if (!check) continue;
        std::vector<int32_t> cs, cd, cw;
        for (int64_t e = 0; e < Q.m_in; ++e)
          if (cur[e] != Q.pw[e]) {
            cs.push_back(Q.pu[e]);
            cd.push_back(Q.pv[e]);
            cw.push_back(cur[e]);
          }
        int64_t fs = 0;
        double ms = 0;
        if ((rc = dmf_static_recalculate(&Q.g, static_cast<int32_t>(cs.size()), cs.data(), cd.data(), cw.data(), Q.s,
                                         Q.t, 0, DMF_VARIANT_DATA, &fs, &ms)))
          die("static_recalculate", rc);
        ++verified;
        if (fs != Q.R[b].flow) {
          ++mismatches;
          std::printf("MISMATCH session %d batch %d served %lld recomputed %lld\n", i, b - warmup,
                      static_cast<long long>(Q.R[b].flow), static_cast<long long>(fs));
        }
        std::printf("[serve] verified session %d batch %d: flow %lld (%zu arcs changed so far)\n", i, b - warmup,
                    static_cast<long long>(fs), cs.size());
        std::fflush(stdout);
      }
    }
    if (nsess > 1) {
      // the same batches alone: a solo engine with the server's parameters
      dmf_engine* E = nullptr;
      if ((rc = dmf_engine_create(&Q.g, Q.s, Q.t, &E))) die("engine_create", rc);
      dmf_params p = cfg.params;
      dmf_params d;
      dmf_params_default(&Q.g, &d);
      p.kc = d.kc;
      dmf_params ps = p;
      ps.scheduler = DMF_SCHED_FIXED;
      ps.variant = DMF_VARIANT_PP;
      int64_t f = 0;
      if ((rc = dmf_engine_static_solve(E, &ps, &f))) die("solo static solve", rc);
      if (f != Q.f0) {
        ++mismatches;
        std::printf("MISMATCH session %d static served %lld alone %lld\n", i, static_cast<long long>(Q.f0),
                    static_cast<long long>(f));
      }
      for (int b = 0; b < total; ++b) {
        if ((rc = dmf_engine_apply_batch(E, static_cast<int32_t>(Q.B[b].src.size()), Q.B[b].src.data(),
                                         Q.B[b].dst.data(), Q.B[b].w.data(), &p, &f)))
          die("solo batch", rc);
        ++replayed;
        if (f != Q.R[b].flow) {
          ++mismatches;
          std::printf("MISMATCH session %d batch %d served %lld alone %lld\n", i, b, static_cast<long long>(Q.R[b].flow),
                      static_cast<long long>(f));
        }
      }
      dmf_engine_destroy(E);
      std::printf("[serve] session %d: %d batches replayed alone\n", i, total);
      std::fflush(stdout);
    }
  }
  std::printf("served %d batches (mean %.0f updates) at %.2f batches/s offered %.2f: latency p50 %.2f ms p99 %.2f "
              "ms; service p50 %.2f ms; queue p50 %.2f ms; GTEPS %.3f; peak staged %lld B; verified %d, "
              "mismatches %d\n",
              nserved, kk / nserved, bps, rate, pct_rank(lat, 50), pct_rank(lat, 99), pct_rank(svcs, 50),
              pct_rank(queue, 50), arcs / (dev * 1e6), static_cast<long long>(stats.peak_staged_bytes), verified,
              mismatches);
  if (nsess > 1)
    std::printf("sessions %d: %d of %d solves overlapped another session's, at most %d in flight; %d batches "
                "replayed alone, all equal: %s\n",
                nsess, overlapped, nserved, peak, replayed, mismatches ? "no" : "yes");
  if (json) {
    std::printf("{\"workload\":\"rmat-%d ef%d seed %llu, %d Poisson batches log-uniform %lld-%lld updates, %.0f%% "
                "increases\",\"n\":%d,\"m\":%d,\"sessions\":%d,\"scheduler\":\"%s\",\"offered_rate\":%.3f,"
                "\"batches_per_s\":%.3f,\"latency_p50_ms\":%.3f,\"latency_p99_ms\":%.3f,\"service_p50_ms\":%.3f,"
                "\"service_p99_ms\":%.3f,\"queue_p50_ms\":%.3f,\"mean_updates\":%.1f,\"gteps\":%.4f,\"hbm_held\":%lld,"
                "\"peak_staged_bytes\":%lld,\"overlapped\":%d,\"peak_in_flight\":%d,\"replayed_alone\":%d,"
                "\"verified\":%d,\"mismatches\":%d}\n",
                scale, ef, static_cast<unsigned long long>(seed), nbatch, static_cast<long long>(kmin),
                static_cast<long long>(kmax), inc * 100, SS[0].g.n, SS[0].g.m, nsess, learned ? "learned" : "plain",
                rate, bps, pct_rank(lat, 50), pct_rank(lat, 99), pct_rank(svcs, 50), pct_rank(svcs, 99),
                pct_rank(queue, 50), kk / nserved, arcs / (dev * 1e6), static_cast<long long>(stats.hbm_held),
                static_cast<long long>(stats.peak_staged_bytes), overlapped, peak, replayed, verified, mismatches);
  }
  for (Session& Q : SS) {
    dmf_csr_free(&Q.g);
    dmf_free(Q.pu);
    dmf_free(Q.pv);
    dmf_free(Q.pw);
  }
  return mismatches ? 1 : 0;
}
