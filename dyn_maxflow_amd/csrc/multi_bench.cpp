// dmf_multi_bench: many uneven independent max-flow instances solved together.
//
//   dmf_multi_bench [--instances=256] [--log-lo=10] [--log-hi=16] [--ef-lo=8]
//                   [--ef-hi=32] [--seed=4] [--batches=5] [--pct=5]
//                   [--repeats=3] [--mode=persistent|stepped|graph|rounds|all, or a comma list]
//                   [--variant=data|topo|pp] [--sched=plain|learned]
//                   [--cascade=K] [--no-autotune] [--tune-interval] [--no-solo] [--json]
//
// Workload (all seeded): instance i draws x uniform in [log_lo, log_hi) and
// has n_i = floor(2^x) vertices, an edge factor uniform in [ef_lo, ef_hi] and
// n_i * ef_i R-MAT samples (a/b/c/d = 0.57/0.19/0.19/0.05, dmf_gen_rmat_n,
// seed seed*1000003+i), capacities uniform [1,100], s = max out-degree,
// t = max in-degree.  Each instance gets `batches` successive update batches
// of pct% of its input edges (half increases, half decreases).
//
// All instances live in one engine as a disjoint union (dmf_batch_create): a
// batch of updates for every instance is one dmf_batch_apply, i.e. one
// cooperative launch in the persistent driver.  Before timing, the degree-bucket
// thresholds are chosen by dmf_engine_autotune (with --tune-interval, for the fixed
// schedule, also the relabel interval: dmf_engine_autotune_schedule).  The solo baseline solves each
// instance in its own engine, one after another, and every instance's flow must
// equal its solo flow after every batch (a mismatch prints MISMATCH and the
// exit status is 1).  No reference counterpart: the reference solves one graph
// per process.
#include <algorithm>
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

struct Opts {
  int instances = 256;
  double log_lo = 10, log_hi = 16;
  int ef_lo = 8, ef_hi = 32;
  uint64_t seed = 4;
  int batches = 5;
  double pct = 5;
  int repeats = 3;
  std::string mode = "all";
  int variant = DMF_VARIANT_DATA;
  int sched = DMF_SCHED_FIXED;  // of the batched engine (the solo baseline keeps the fixed schedule)
  int cascade = 0;              // cascade sub-rounds per round of the batched engine
  bool autotune = true;
  bool tune_interval = false;  // fixed schedule: also autotune the relabel interval
  bool solo = true;
  bool json = false;
};

struct Instance {
  int32_t n_req = 0, ef = 0;
  uint64_t seed = 0;
  std::vector<int32_t> u, v, w0;
  dmf_csr g{};
  int32_t s = 0, t = 0;
  // batches[b] = (src, dst, w)
  std::vector<std::vector<int32_t>> bs, bd, bw;
};

uint64_t splitmix(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double u01(uint64_t& x) { return static_cast<double>(splitmix(x) >> 11) * (1.0 / 9007199254740992.0); }

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

[[noreturn]] void die(const char* what, int rc) {
  std::fprintf(stderr, "dmf_multi_bench: %s: %s (%d)\n", what, dmf_strerror(rc), rc);
  std::exit(2);
}

int make_instance(Instance& I, const Opts& o) {
  int64_t k = 0;
  int32_t *u = nullptr, *v = nullptr, *w = nullptr;
  int rc = dmf_gen_rmat_n(I.n_req, static_cast<int64_t>(I.n_req) * I.ef, 0.57, 0.19, 0.19, I.seed, 1, 100, &k,
                          &u, &v, &w);
  if (rc) return rc;
  I.u.assign(u, u + k);
  I.v.assign(v, v + k);
  I.w0.assign(w, w + k);
  dmf_free(u);
  dmf_free(v);
  dmf_free(w);
  if ((rc = dmf_build_csr(k, I.u.data(), I.v.data(), I.w0.data(), &I.g))) return rc;
  if ((rc = dmf_pick_source_sink(I.g.n, k, I.u.data(), I.v.data(), &I.s, &I.t))) return rc;
  std::vector<int32_t> cur = I.w0;
  const int64_t count = std::max<int64_t>(1, static_cast<int64_t>(std::floor(k * o.pct / 100.0)));
  for (int b = 0; b < o.batches; ++b) {
    int32_t kk = 0;
    int32_t *a = nullptr, *d = nullptr, *c = nullptr;
    if ((rc = dmf_gen_updates(k, I.u.data(), I.v.data(), cur.data(), count, 0.5, 1, 100,
                              I.seed * 31 + static_cast<uint64_t>(b) + 1, 1, &kk, &a, &d, &c)))
      return rc;
    I.bs.emplace_back(a, a + kk);
    I.bd.emplace_back(d, d + kk);
    I.bw.emplace_back(c, c + kk);
    dmf_free(a);
    dmf_free(d);
    dmf_free(c);
  }
  return DMF_OK;
}

double pct_rank(std::vector<double> x, double q) {  // nearest rank
  if (x.empty()) return 0;
  std::sort(x.begin(), x.end());
  size_t r = static_cast<size_t>(std::ceil(q / 100.0 * x.size()));
  return x[std::min(x.size(), std::max<size_t>(r, 1)) - 1];
}

struct ModeResult {
  std::string name;
  int32_t lane_max = 0, wide_t = 0, interval = 1;
  double static_ms = 0;
  std::vector<double> dev_ms, wall_ms;
  uint64_t arc_visits = 0;
  double arc_ms = 0;
  uint64_t launches = 0;
};

}  // namespace

int main(int argc, char** argv) {
  Opts o;
  for (int i = 1; i < argc; ++i) {
    const char* x = argv[i];
    auto val = [&](const char* key) -> const char* {
      const size_t l = std::strlen(key);
      return std::strncmp(x, key, l) == 0 ? x + l : nullptr;
    };
    const char* a;
    if ((a = val("--instances="))) o.instances = std::atoi(a);
    else if ((a = val("--log-lo="))) o.log_lo = std::atof(a);
    else if ((a = val("--log-hi="))) o.log_hi = std::atof(a);
    else if ((a = val("--ef-lo="))) o.ef_lo = std::atoi(a);
    else if ((a = val("--ef-hi="))) o.ef_hi = std::atoi(a);
    else if ((a = val("--seed="))) o.seed = std::strtoull(a, nullptr, 10);
    else if ((a = val("--batches="))) o.batches = std::atoi(a);
    else if ((a = val("--pct="))) o.pct = std::atof(a);
    else if ((a = val("--repeats="))) o.repeats = std::atoi(a);
    else if ((a = val("--mode="))) o.mode = a;
    else if ((a = val("--variant="))) {
      o.variant = !std::strcmp(a, "topo") ? DMF_VARIANT_TOPO : !std::strcmp(a, "pp") ? DMF_VARIANT_PP
                                                                                       : DMF_VARIANT_DATA;
    } else if ((a = val("--sched="))) {
      o.sched = !std::strcmp(a, "learned") ? DMF_SCHED_LEARNED : DMF_SCHED_FIXED;
    } else if ((a = val("--cascade="))) {
      o.cascade = std::atoi(a);
    } else if (!std::strcmp(x, "--no-autotune")) o.autotune = false;
    else if (!std::strcmp(x, "--tune-interval")) o.tune_interval = true;
    else if (!std::strcmp(x, "--no-solo")) o.solo = false;
    else if (!std::strcmp(x, "--json")) o.json = true;
    else {
      std::fprintf(stderr, "dmf_multi_bench: unknown argument %s\n", x);
      return 2;
    }
  }
  if (o.instances < 1 || o.log_lo < 1 || o.log_hi < o.log_lo || o.log_hi > 30 || o.ef_lo < 1 ||
      o.ef_hi < o.ef_lo || o.batches < 0 || o.pct <= 0 || o.repeats < 1) {
    std::fprintf(stderr, "dmf_multi_bench: invalid arguments\n");
    return 2;
  }
  if (dmf_device_count() < 1) {
    std::fprintf(stderr, "dmf_multi_bench: no GPU visible\n");
    return 2;
  }

  // ---- workload
  const double t_gen = now_ms();
  std::vector<Instance> inst(o.instances);
  uint64_t st = o.seed;
  for (int i = 0; i < o.instances; ++i) {
    const double x = o.log_lo + (o.log_hi - o.log_lo) * u01(st);
    inst[i].n_req = std::max(2, static_cast<int32_t>(std::floor(std::pow(2.0, x))));
    inst[i].ef = o.ef_lo + static_cast<int>(std::floor(u01(st) * (o.ef_hi - o.ef_lo + 1)));
    inst[i].seed = o.seed * 1000003ull + static_cast<uint64_t>(i);
  }
  {
    const int nt = std::max(1, std::min<int>(16, static_cast<int>(std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    std::vector<int> rcs(nt, 0);
    for (int w = 0; w < nt; ++w)
      pool.emplace_back([&, w] {
        for (int i = w; i < o.instances; i += nt) {
          const int rc = make_instance(inst[i], o);
          if (rc) rcs[w] = rc;
        }
      });
    for (auto& th : pool) th.join();
    for (int rc : rcs)
      if (rc) die("instance generation", rc);
  }
  int64_t n_tot = 0, m_tot = 0, upd_tot = 0;
  for (auto& I : inst) {
    n_tot += I.g.n;
    m_tot += I.g.m;
    for (auto& b : I.bs) upd_tot += static_cast<int64_t>(b.size());
  }
  std::printf("instances:%d vertices:%lld arcs:%lld updates:%lld gen_ms:%.1f\n", o.instances,
              static_cast<long long>(n_tot), static_cast<long long>(m_tot), static_cast<long long>(upd_tot),
              now_ms() - t_gen);
  std::fflush(stdout);

  std::vector<dmf_csr> graphs(o.instances);
  std::vector<int32_t> S(o.instances), T(o.instances);
  for (int i = 0; i < o.instances; ++i) {
    graphs[i] = inst[i].g;
    S[i] = inst[i].s;
    T[i] = inst[i].t;
  }
  // concatenated batches (instance order), one per batch index
  std::vector<std::vector<int32_t>> cnt(o.batches), cs(o.batches), cd(o.batches), cw(o.batches);
  for (int b = 0; b < o.batches; ++b)
    for (auto& I : inst) {
      cnt[b].push_back(static_cast<int32_t>(I.bs[b].size()));
      cs[b].insert(cs[b].end(), I.bs[b].begin(), I.bs[b].end());
      cd[b].insert(cd[b].end(), I.bd[b].begin(), I.bd[b].end());
      cw[b].insert(cw[b].end(), I.bw[b].begin(), I.bw[b].end());
    }

  // ---- solo baseline (and the expected flows)
  std::vector<std::vector<int64_t>> expect(o.batches + 1, std::vector<int64_t>(o.instances, -1));
  double solo_static_ms = 0;
  std::vector<double> solo_batch_ms(o.batches, 0.0);
  dmf_params p;
  dmf_params_default(&inst[0].g, &p);
  p.variant = o.variant;
  if (o.solo) {
    for (int i = 0; i < o.instances; ++i) {
      dmf_engine* e = nullptr;
      int rc = dmf_engine_create(&inst[i].g, inst[i].s, inst[i].t, &e);
      if (rc) die("solo engine", rc);
      dmf_params q;
      dmf_params_default(&inst[i].g, &q);
      q.variant = o.variant;
      q.kc = p.kc;
      q.mode = 1;
      dmf_counters c;
      int64_t f = 0;
      if ((rc = dmf_engine_static_solve(e, &q, &f))) die("solo static solve", rc);
      dmf_engine_counters(e, &c);
      solo_static_ms += c.device_ms;
      expect[0][i] = f;
      for (int b = 0; b < o.batches; ++b) {
        if ((rc = dmf_engine_apply_batch(e, static_cast<int32_t>(inst[i].bs[b].size()), inst[i].bs[b].data(),
                                         inst[i].bd[b].data(), inst[i].bw[b].data(), &q, &f)))
          die("solo batch", rc);
        dmf_engine_counters(e, &c);
        solo_batch_ms[b] += c.device_ms;
        expect[b + 1][i] = f;
      }
      dmf_engine_destroy(e);
    }
    std::printf("solo: static %.3f ms, batches", solo_static_ms);
    for (double x : solo_batch_ms) std::printf(" %.3f", x);
    std::printf(" ms (sum over instances, one engine each)\n");
    std::fflush(stdout);
  }

  // ---- batched engine
  dmf_engine* E = nullptr;
  int rc = dmf_batch_create(o.instances, graphs.data(), S.data(), T.data(), &E);
  if (rc) die("batch_create", rc);
  std::printf("batched engine: %lld device bytes\n", static_cast<long long>(dmf_engine_device_bytes(E)));
  std::vector<std::string> modes;
  if (o.mode == "all") {
    modes = {"persistent", "stepped", "graph", "rounds"};
  } else {
    for (size_t a = 0; a <= o.mode.size();) {  // comma-separated list
      const size_t b = std::min(o.mode.find(',', a), o.mode.size());
      if (b > a) modes.push_back(o.mode.substr(a, b - a));
      a = b + 1;
    }
  }
  for (const std::string& mname : modes)
    if (mname != "persistent" && mname != "stepped" && mname != "graph" && mname != "rounds") {
      std::fprintf(stderr, "dmf_multi_bench: unknown mode %s\n", mname.c_str());
      return 2;
    }
  std::vector<ModeResult> results;
  int mismatches = 0;
  std::vector<int64_t> flows(o.instances);
  for (const std::string& mname : modes) {
    ModeResult R;
    R.name = mname;
    dmf_params q = p;
    q.mode = mname == "persistent" ? 0 : mname == "stepped" ? 1 : mname == "graph" ? 2 : 3;
    q.scheduler = o.sched;
    q.cascade = o.cascade;
    if (o.autotune && o.tune_interval) {
      if ((rc = dmf_engine_autotune_schedule(E, &q, &R.lane_max, &R.wide_t))) die("autotune_schedule", rc);
    } else if (o.autotune) {
      if ((rc = dmf_engine_autotune(E, &q, &R.lane_max, &R.wide_t))) die("autotune", rc);
    }
    R.interval = q.relabel_interval;
    for (int rep = 0; rep < o.repeats + 1; ++rep) {  // repeat 0 warms up
      if ((rc = dmf_engine_reset(E))) die("reset", rc);
      if ((rc = dmf_batch_static_solve(E, &q, flows.data()))) die("batch static solve", rc);
      dmf_counters c;
      dmf_engine_counters(E, &c);
      if (rep == 1) R.static_ms = c.device_ms;
      for (int i = 0; o.solo && i < o.instances; ++i)
        if (flows[i] != expect[0][i]) {
          ++mismatches;
          std::printf("MISMATCH mode=%s static instance %d batched %lld solo %lld\n", mname.c_str(), i,
                      static_cast<long long>(flows[i]), static_cast<long long>(expect[0][i]));
        }
      for (int b = 0; b < o.batches; ++b) {
        const double w0 = now_ms();
        if ((rc = dmf_batch_apply(E, cnt[b].data(), cs[b].data(), cd[b].data(), cw[b].data(), &q, flows.data())))
          die("batch apply", rc);
        const double wall = now_ms() - w0;
        dmf_engine_counters(E, &c);
        if (rep > 0) {
          R.dev_ms.push_back(c.device_ms);
          R.wall_ms.push_back(wall);
          R.arc_visits += c.arc_visits;
          R.arc_ms += c.device_ms;
          R.launches += c.launches;
        }
        for (int i = 0; o.solo && i < o.instances; ++i)
          if (flows[i] != expect[b + 1][i]) {
            ++mismatches;
            std::printf("MISMATCH mode=%s batch %d instance %d batched %lld solo %lld\n", mname.c_str(), b, i,
                        static_cast<long long>(flows[i]), static_cast<long long>(expect[b + 1][i]));
          }
      }
    }
    std::printf("mode=%s lane_max=%d wide_t=%d interval=%d static %.3f ms; per batch (all %d instances) p50 %.3f ms "
                "p99 %.3f ms wall p50 %.3f ms; GTEPS %.3f; launches/batch %.1f\n",
                mname.c_str(), R.lane_max, R.wide_t, R.interval, R.static_ms, o.instances, pct_rank(R.dev_ms, 50),
                pct_rank(R.dev_ms, 99), pct_rank(R.wall_ms, 50), R.arc_visits / (R.arc_ms * 1e6),
                R.dev_ms.empty() ? 0.0 : static_cast<double>(R.launches) / R.dev_ms.size());
    std::fflush(stdout);
    results.push_back(R);
  }
  dmf_engine_destroy(E);
  for (auto& I : inst) dmf_csr_free(&I.g);

  double solo_p50 = 0;
  if (o.solo) solo_p50 = pct_rank(solo_batch_ms, 50);
  if (o.json) {
    std::printf("{\"workload\":\"%d RMAT instances, n log-uniform 2^%g-2^%g, EF %d-%d, seed %llu, %d batches of %g%%\","
                "\"scheduler\":\"%s\",\"cascade\":%d,\"vertices\":%lld,\"arcs\":%lld,\"updates\":%lld,\"solo_batch_p50_ms\":%.4f,\"verified\":%s,"
                "\"mismatches\":%d,\"modes\":[",
                o.instances, o.log_lo, o.log_hi, o.ef_lo, o.ef_hi, static_cast<unsigned long long>(o.seed),
                o.batches, o.pct, o.sched == DMF_SCHED_LEARNED ? "learned" : "plain", o.cascade,
                static_cast<long long>(n_tot), static_cast<long long>(m_tot),
                static_cast<long long>(upd_tot), solo_p50, o.solo ? "true" : "false", mismatches);
    for (size_t r = 0; r < results.size(); ++r) {
      const ModeResult& R = results[r];
      const double p50 = pct_rank(R.dev_ms, 50);
      std::printf("%s{\"mode\":\"%s\",\"lane_max\":%d,\"wide_t\":%d,\"relabel_interval\":%d,\"static_ms\":%.4f,"
                  "\"batch_p50_ms\":%.4f,\"batch_p99_ms\":%.4f,\"wall_p50_ms\":%.4f,\"gteps\":%.4f,\"speedup_vs_solo\":%.3f}",
                  r ? "," : "", R.name.c_str(), R.lane_max, R.wide_t, R.interval, R.static_ms, p50, pct_rank(R.dev_ms, 99),
                  pct_rank(R.wall_ms, 50), R.arc_visits / (R.arc_ms * 1e6), p50 > 0 ? solo_p50 / p50 : 0.0);
    }
    std::printf("]}\n");
  }
  if (mismatches) {
    std::printf("FAILED: %d flow mismatches between batched and solo solves\n", mismatches);
    return 1;
  }
  return 0;
}
