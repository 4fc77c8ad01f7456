// dmf_serve_bench: streaming update batches against one resident graph.
//
//   dmf_serve_bench [--scale=23] [--ef=16] [--seed=5] [--batches=60] [--warmup=3]
//                   [--kmin=1000] [--kmax=1000000] [--inc=0.4] [--load=0.7 | --rate=R]
//                   [--sched=learned|plain] [--verify-every=10] [--json]
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

}  // namespace

int main(int argc, char** argv) {
  int scale = 23, ef = 16, nbatch = 60, warmup = 3, verify_every = 10;
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
    else if (!std::strcmp(x, "--json")) json = true;
    else {
      std::fprintf(stderr, "dmf_serve_bench: unknown argument %s\n", x);
      return 2;
    }
  }
  if (scale < 2 || ef < 1 || nbatch < 1 || warmup < 1 || kmin < 1 || kmax < kmin || load <= 0 || rate < 0) {
    std::fprintf(stderr, "dmf_serve_bench: invalid arguments\n");
    return 2;
  }
  if (dmf_device_count() < 1) {
    std::fprintf(stderr, "dmf_serve_bench: no GPU visible\n");
    return 2;
  }
  // ---- graph
  double t = now_s();
  int64_t m_in = 0;
  int32_t *pu = nullptr, *pv = nullptr, *pw = nullptr;
  int rc = dmf_gen_rmat(scale, ef, 0.57, 0.19, 0.19, seed, 1, 100, &m_in, &pu, &pv, &pw);
  if (rc) die("gen_rmat", rc);
  std::printf("[serve] rmat-%d ef%d seed %llu: %lld edges (%.1f s)\n", scale, ef,
              static_cast<unsigned long long>(seed), static_cast<long long>(m_in), now_s() - t);
  std::fflush(stdout);
  t = now_s();
  dmf_csr g{};
  if ((rc = dmf_build_csr(m_in, pu, pv, pw, &g))) die("build_csr", rc);
  int32_t s = 0, tt = 0;
  if ((rc = dmf_pick_source_sink(g.n, m_in, pu, pv, &s, &tt))) die("pick_source_sink", rc);
  std::printf("[serve] csr n=%d m=%d s=%d t=%d (%.1f s)\n", g.n, g.m, s, tt, now_s() - t);
  std::fflush(stdout);
  // ---- batches (log-uniform sizes, successive)
  t = now_s();
  const int total = warmup + nbatch;
  std::vector<Batch> B(total);
  std::vector<int64_t> counts(total);
  {
    uint64_t st = seed * 7919 + 11;
    std::vector<int32_t> cur(pw, pw + m_in);
    const double l0 = std::log(static_cast<double>(kmin)), l1 = std::log(static_cast<double>(kmax));
    for (int b = 0; b < total; ++b) {
      counts[b] = std::min<int64_t>(m_in, static_cast<int64_t>(std::exp(l0 + (l1 - l0) * u01(st))));
      int32_t k = 0;
      int32_t *a = nullptr, *d = nullptr, *c = nullptr;
      if ((rc = dmf_gen_updates(m_in, pu, pv, cur.data(), counts[b], inc, 1, 100, seed * 1000003ull + b, 1, &k,
                                &a, &d, &c)))
        die("gen_updates", rc);
      B[b].src.assign(a, a + k);
      B[b].dst.assign(d, d + k);
      B[b].w.assign(c, c + k);
      dmf_free(a);
      dmf_free(d);
      dmf_free(c);
    }
  }
  int64_t kbig = 0;
  for (auto& b : B) kbig = std::max<int64_t>(kbig, static_cast<int64_t>(b.src.size()));
  std::printf("[serve] %d batches generated (%.1f s)\n", total, now_s() - t);
  std::fflush(stdout);
  // ---- server
  dmf_server_config cfg;
  dmf_server_config_default(&cfg);
  cfg.params.scheduler = learned ? DMF_SCHED_LEARNED : DMF_SCHED_FIXED;
  cfg.params.mode = 2;
  cfg.staging_bytes = std::max<int64_t>(int64_t{64} << 20, kbig * 12 * 4);  // room for 4 of the largest
  dmf_server* S = nullptr;
  if ((rc = dmf_server_create(&cfg, &S))) die("server_create", rc);
  int32_t sid = 0;
  int64_t f0 = 0;
  t = now_s();
  if ((rc = dmf_server_open(S, &g, s, tt, static_cast<int32_t>(kbig), &sid, &f0))) die("server_open", rc);
  std::printf("[serve] session open, static flow %lld (%.1f s)\n", static_cast<long long>(f0), now_s() - t);
  std::fflush(stdout);
  std::vector<dmf_result> R(total);
  // warm-up: closed loop, sets the arrival rate
  double svc = 0;
  for (int b = 0; b < warmup; ++b) {
    uint64_t tk = 0;
    if ((rc = dmf_server_submit(S, sid, static_cast<int32_t>(B[b].src.size()), B[b].src.data(), B[b].dst.data(),
                                B[b].w.data(), &tk)))
      die("submit", rc);
    if ((rc = dmf_server_wait(S, tk, -1, &R[b]))) die("wait", rc);
    if (R[b].status) die("warm-up batch", R[b].status);
    svc += R[b].done_ms - R[b].start_ms;
  }
  svc /= warmup;
  if (rate <= 0) rate = load * 1000.0 / std::max(svc, 1e-3);
  std::printf("[serve] warm-up mean service %.2f ms -> Poisson arrivals at %.2f batches/s\n", svc, rate);
  std::fflush(stdout);
  // ---- timed: a producer thread submits at Poisson arrival times
  // tickets[b]: 0 until the producer has submitted batch b (~0 if refused)
  std::vector<std::atomic<uint64_t>> tickets(total);
  std::vector<int> sub_rc(total, 0);  // written before tickets[b] is published
  for (auto& x : tickets) x.store(0);
  const double t_start = now_s();
  std::thread producer([&] {
    uint64_t st = seed * 104729 + 3;
    double at = 0;
    for (int b = warmup; b < total; ++b) {
      at += -std::log(1.0 - u01(st)) / rate;
      const double wait = t_start + at - now_s();
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      uint64_t tk = 0;
      sub_rc[b] = dmf_server_submit(S, sid, static_cast<int32_t>(B[b].src.size()), B[b].src.data(),
                                    B[b].dst.data(), B[b].w.data(), &tk);
      tickets[b].store(sub_rc[b] ? ~uint64_t{0} : tk, std::memory_order_release);
    }
  });
  int done = 0;
  for (int b = warmup; b < total; ++b) {
    uint64_t tk;
    while ((tk = tickets[b].load(std::memory_order_acquire)) == 0)
      std::this_thread::sleep_for(std::chrono::microseconds(50));
    if (sub_rc[b]) die("submit", sub_rc[b]);
    if ((rc = dmf_server_wait(S, tk, -1, &R[b]))) die("wait", rc);
    if (R[b].status) die("served batch", R[b].status);
    if (++done % 10 == 0) {
      std::printf("[serve] %d/%d served\n", done, nbatch);
      std::fflush(stdout);
    }
  }
  producer.join();
  dmf_server_stats stats;
  dmf_server_get_stats(S, &stats);
  dmf_server_destroy(S);
  // ---- report
  std::vector<double> lat, svcs, queue;
  double arcs = 0, dev = 0, kk = 0;
  for (int b = warmup; b < total; ++b) {
    lat.push_back(R[b].done_ms - R[b].submit_ms);
    svcs.push_back(R[b].done_ms - R[b].start_ms);
    queue.push_back(R[b].start_ms - R[b].submit_ms);
    arcs += static_cast<double>(R[b].arc_visits);
    dev += R[b].device_ms;
    kk += R[b].k;
  }
  const double span_ms = R[total - 1].done_ms - R[warmup].submit_ms;
  const double bps = nbatch * 1000.0 / span_ms;
  // ---- verification (after the timed run)
  int verified = 0, mismatches = 0;
  if (verify_every > 0) {
    std::vector<int32_t> cur(pw, pw + m_in);
    for (int b = 0; b < total; ++b) {
      int32_t k = 0;
      int32_t *a = nullptr, *d = nullptr, *c = nullptr;
      if ((rc = dmf_gen_updates(m_in, pu, pv, cur.data(), counts[b], inc, 1, 100, seed * 1000003ull + b, 1, &k,
                                &a, &d, &c)))
        die("gen_updates", rc);
      dmf_free(a);
      dmf_free(d);
      dmf_free(c);
      const bool check = b >= warmup && ((b - warmup + 1) % verify_every == 0 || b == total - 1);
      if (!check) continue;
      std::vector<int32_t> cs, cd, cw;
      for (int64_t e = 0; e < m_in; ++e)
        if (cur[e] != pw[e]) {
          cs.push_back(pu[e]);
          cd.push_back(pv[e]);
          cw.push_back(cur[e]);
        }
      int64_t fs = 0;
      double ms = 0;
      if ((rc = dmf_static_recalculate(&g, static_cast<int32_t>(cs.size()), cs.data(), cd.data(), cw.data(), s, tt,
                                       0, DMF_VARIANT_DATA, &fs, &ms)))
        die("static_recalculate", rc);
      ++verified;
      if (fs != R[b].flow) {
        ++mismatches;
        std::printf("MISMATCH batch %d served %lld recomputed %lld\n", b - warmup, static_cast<long long>(R[b].flow),
                    static_cast<long long>(fs));
      }
      std::printf("[serve] verified batch %d: flow %lld (%zu arcs changed so far)\n", b - warmup,
                  static_cast<long long>(fs), cs.size());
      std::fflush(stdout);
    }
  }
  std::printf("served %d batches (mean %.0f updates) at %.2f batches/s offered %.2f: latency p50 %.2f ms p99 %.2f "
              "ms; service p50 %.2f ms; queue p50 %.2f ms; GTEPS %.3f; peak staged %lld B; verified %d, "
              "mismatches %d\n",
              nbatch, kk / nbatch, bps, rate, pct_rank(lat, 50), pct_rank(lat, 99), pct_rank(svcs, 50),
              pct_rank(queue, 50), arcs / (dev * 1e6), static_cast<long long>(stats.peak_staged_bytes), verified,
              mismatches);
  if (json) {
    std::printf("{\"workload\":\"rmat-%d ef%d seed %llu, %d Poisson batches log-uniform %lld-%lld updates, %.0f%% "
                "increases\",\"n\":%d,\"m\":%d,\"scheduler\":\"%s\",\"offered_rate\":%.3f,\"batches_per_s\":%.3f,"
                "\"latency_p50_ms\":%.3f,\"latency_p99_ms\":%.3f,\"service_p50_ms\":%.3f,\"service_p99_ms\":%.3f,"
                "\"queue_p50_ms\":%.3f,\"mean_updates\":%.1f,\"gteps\":%.4f,\"hbm_held\":%lld,"
                "\"peak_staged_bytes\":%lld,\"verified\":%d,\"mismatches\":%d}\n",
                scale, ef, static_cast<unsigned long long>(seed), nbatch, static_cast<long long>(kmin),
                static_cast<long long>(kmax), inc * 100, g.n, g.m, learned ? "learned" : "plain", rate, bps,
                pct_rank(lat, 50), pct_rank(lat, 99), pct_rank(svcs, 50), pct_rank(svcs, 99), pct_rank(queue, 50),
                kk / nbatch, arcs / (dev * 1e6), static_cast<long long>(stats.hbm_held),
                static_cast<long long>(stats.peak_staged_bytes), verified, mismatches);
  }
  dmf_csr_free(&g);
  dmf_free(pu);
  dmf_free(pv);
  dmf_free(pw);
  return mismatches ? 1 : 0;
}
