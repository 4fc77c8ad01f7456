// Bench drivers with the reference's command line and per-file output:
//   dyn_{data,topo,pp}_bench        <graph.bin> <update_prefix> <source> <sink> [flags]
//   dyn_{data,topo,pp}_bench_large  <graph.bin> <update_prefix> <source> <sink> <i> [flags]
// (main of code/dyn_data_bench.cu:495-567, code/dyn_topo_bench.cu:399-476,
//  code/dyn_pp_bench.cu:624-695; _large variants take one file index,
//  code/dyn_data_bench_large.cu:502).
//
// For each update file <prefix>_p<i>.bin and each kernel_cycles value in
// {m/n-1 (if m/n > 1), m/n (if m/n > 0), m/n+1}, the dynamic path (fresh
// state, static solve, the whole file as one batch) is compared with a static
// recompute on the updated graph, and a mismatch prints the PANIC line.
//
// Optional flags (after the positional arguments; not in the reference):
//   --mode=independent|successive  independent (default) applies every file to
//                                  the original graph, as the reference does;
//                                  successive keeps one resident state and
//                                  applies the files one after another.
//   --kc=N                         run only this kernel_cycles value
//   --relabel-interval=N           rounds between global relabels (default 1)
//   --driver=graph|rounds|stepped|persistent  engine driver (default graph: per-phase
//                                  launches with the BFS levels replayed from
//                                  captured hipGraphs); --stepped and
//                                  --persistent are short forms
//   --json                         one JSON line of counters per run
//   --no-static-check              skip the static recompute
// Compiled once per variant with -DDMF_BENCH_VARIANT="data|topo|pp" and
// -DDMF_BENCH_LARGE=0|1.  Uses only the C-ABI of include/dmf.h.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dmf.h"

#ifndef DMF_BENCH_VARIANT
#define DMF_BENCH_VARIANT "data"
#endif
#ifndef DMF_BENCH_LARGE
#define DMF_BENCH_LARGE 0
#endif

namespace {

struct Opts {
  bool successive = false;
  int kc_only = 0;
  int interval = 1;
  int driver = 2;  // dmf_params.mode
  bool json = false;
  bool static_check = true;
};

int variant_id() {
  if (std::strcmp(DMF_BENCH_VARIANT, "topo") == 0) return DMF_VARIANT_TOPO;
  if (std::strcmp(DMF_BENCH_VARIANT, "pp") == 0) return DMF_VARIANT_PP;
  return DMF_VARIANT_DATA;
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

void print_json(const char* what, int file, int kc, long long flow, const dmf_counters& c) {
  std::printf(
      "{\"run\":\"%s\",\"variant\":\"%s\",\"file\":%d,\"kc\":%d,\"flow\":%lld,\"device_ms\":%.6f,"
      "\"arc_visits\":%llu,\"vertex_visits\":%llu,\"pushes\":%llu,\"relabels\":%llu,"
      "\"repairs\":%llu,\"bfs_levels\":%llu,\"global_relabels\":%llu,\"rounds\":%llu,"
      "\"updates\":%llu,\"atomics\":%llu,\"launches\":%llu}\n",
      what, DMF_BENCH_VARIANT, file, kc, flow, c.device_ms, (unsigned long long)c.arc_visits,
      (unsigned long long)c.vertex_visits, (unsigned long long)c.pushes,
      (unsigned long long)c.relabels, (unsigned long long)c.repairs,
      (unsigned long long)c.bfs_levels, (unsigned long long)c.global_relabels,
      (unsigned long long)c.rounds, (unsigned long long)c.updates, (unsigned long long)c.atomics,
      (unsigned long long)c.launches);
}

void print_static_lines(int variant, int kc, long long flow, const dmf_counters& c) {
  if (variant == DMF_VARIANT_PP) {
    std::printf("static main push cnt:%d GPU Time: %.6f ms\n", (int)c.rounds, c.device_ms);
  } else {
    std::printf("static GPU Time: %.6f ms\n", c.device_ms);
    if (variant == DMF_VARIANT_TOPO) std::printf("flow:%d  kernel cycles:%d\n", (int)flow, kc);
  }
}

int fail(int rc, const char* what) {
  std::fprintf(stderr, "%s: %s (%d)\n", what, dmf_strerror(rc), rc);
  return rc;
}

// recalculate_max_flow (code/dyn_data_bench.cu:171-298): fresh state, static
// solve, the whole list as one batch, flow.
int dynamic_run(const dmf_csr& g, const std::vector<int32_t>& src, const std::vector<int32_t>& dst,
                const std::vector<int32_t>& w, int s, int t, int kc, const Opts& o, int file,
                long long* flow_out) {
  const int variant = variant_id();
  dmf_engine* E = nullptr;
  int rc = dmf_engine_create(&g, s, t, &E);
  if (rc) return fail(rc, "engine_create");
  dmf_params p;
  dmf_params_default(&g, &p);
  p.variant = variant;
  p.kc = kc;
  p.relabel_interval = o.interval;
  p.mode = o.driver;
  int64_t f = 0;
  rc = dmf_engine_static_solve(E, &p, &f);
  if (rc) {
    dmf_engine_destroy(E);
    return fail(rc, "static solve");
  }
  dmf_counters c;
  dmf_engine_counters(E, &c);
  print_static_lines(variant, kc, f, c);
  if (o.json) print_json("static", file, kc, f, c);
  const double t0 = now_ms();
  rc = dmf_engine_apply_batch(E, (int32_t)src.size(), src.data(), dst.data(), w.data(), &p, &f);
  const double total = now_ms() - t0;
  if (rc) {
    dmf_engine_destroy(E);
    return fail(rc, "apply batch");
  }
  dmf_engine_counters(E, &c);
  std::printf("total GPU Time: %.6f ms\n", total);
  std::printf("DYN TIME:%.6fms true_cnt=%d cnt=%d per batch time:%.6f\n", c.device_ms, 1, 1,
              c.device_ms);
  std::printf("Dynamic flow:%d\n", (int)f);
  if (o.json) print_json("dynamic", file, kc, f, c);
  dmf_engine_destroy(E);
  *flow_out = f;
  return 0;
}

// static_recalculate (code/dyn_data_bench.cu:325-492).
int static_run(const dmf_csr& g, const std::vector<int32_t>& src, const std::vector<int32_t>& dst,
               const std::vector<int32_t>& w, int s, int t, int kc, long long* flow_out) {
  int64_t f = 0;
  double ms = 0;
  int rc = dmf_static_recalculate(&g, (int32_t)src.size(), src.data(), dst.data(), w.data(), s, t,
                                  kc, variant_id(), &f, &ms);
  if (rc) return fail(rc, "static recalculate");
  if (variant_id() == DMF_VARIANT_PP) std::printf("static  UPDATE GPU Time: %.6f ms\n", ms);
  else std::printf("STATIC UPDATE GPU Time: %.6f ms\n", ms);
  *flow_out = f;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const int npos = DMF_BENCH_LARGE ? 6 : 5;
  if (argc < npos) {
    std::fprintf(stderr, "usage: %s <graph.bin> <update_prefix> <source> <sink>%s [flags]\n", argv[0],
                 DMF_BENCH_LARGE ? " <i>" : "");
    return 1;
  }
  Opts o;
  for (int a = npos; a < argc; ++a) {
    const char* x = argv[a];
    if (!std::strcmp(x, "--mode=successive")) o.successive = true;
    else if (!std::strcmp(x, "--mode=independent")) o.successive = false;
    else if (!std::strncmp(x, "--kc=", 5)) o.kc_only = std::atoi(x + 5);
    else if (!std::strncmp(x, "--relabel-interval=", 19)) o.interval = std::atoi(x + 19);
    else if (!std::strcmp(x, "--stepped") || !std::strcmp(x, "--driver=stepped")) o.driver = 1;
    else if (!std::strcmp(x, "--persistent") || !std::strcmp(x, "--driver=persistent")) o.driver = 0;
    else if (!std::strcmp(x, "--driver=graph")) o.driver = 2;
    else if (!std::strcmp(x, "--driver=rounds")) o.driver = 3;
    else if (!std::strcmp(x, "--json")) o.json = true;
    else if (!std::strcmp(x, "--no-static-check")) o.static_check = false;
    else {
      std::fprintf(stderr, "unknown flag %s\n", x);
      return 1;
    }
  }
  const char* graph_path = argv[1];
  const char* prefix = argv[2];
  const int source = std::atoi(argv[3]);
  const int sink = std::atoi(argv[4]);
  std::printf("source:%d sink:%d\n", source, sink);
  dmf_csr g;
  int rc = dmf_read_graph_bin(graph_path, &g);
  if (rc) return fail(rc, graph_path);
  if (variant_id() == DMF_VARIANT_TOPO) {
    std::printf("V:%d\n", g.n);
    std::printf("E:%d\n", g.m);
  }
  const int first = DMF_BENCH_LARGE ? std::atoi(argv[5]) : 1;
  const int last = DMF_BENCH_LARGE ? first : 10;
  dmf_engine* resident = nullptr;  // successive mode
  dmf_params rp;
  if (o.successive) {
    rc = dmf_engine_create(&g, source, sink, &resident);
    if (rc) return fail(rc, "engine_create");
    dmf_params_default(&g, &rp);
    rp.variant = variant_id();
    rp.relabel_interval = o.interval;
    rp.mode = o.driver;
    if (o.kc_only > 0) rp.kc = o.kc_only;
    int64_t f0 = 0;
    rc = dmf_engine_static_solve(resident, &rp, &f0);
    if (rc) return fail(rc, "static solve");
    dmf_counters c;
    dmf_engine_counters(resident, &c);
    print_static_lines(variant_id(), rp.kc, f0, c);
  }
  int panics = 0;
  std::vector<int32_t> all_src, all_dst, all_w;  // successive mode: every record applied so far
  for (int i = first; i <= last; ++i) {
    char name[1024];
    std::snprintf(name, sizeof(name), "%s_p%d.bin", prefix, i);
    std::printf("Processing file: %s\n", name);
    int32_t k = 0;
    int32_t *a = nullptr, *b = nullptr, *c = nullptr;
    std::vector<int32_t> src, dst, w;
    rc = dmf_read_updates_bin(name, &k, &a, &b, &c);
    if (rc) {
      std::fprintf(stderr, "Error: cannot open %s for reading\n", name);  // the reference continues
      k = 0;
    } else {
      src.assign(a, a + k);
      dst.assign(b, b + k);
      w.assign(c, c + k);
      dmf_free(a);
      dmf_free(b);
      dmf_free(c);
    }
    int selfloops = 0;
    for (int j = 0; j < k; ++j) selfloops += src[j] == dst[j];
    // code/update.hpp:72-83: k placeholder records precede the real ones
    std::printf("updates size:%d\n", 2 * k - selfloops);
    if (o.successive) {
      int64_t f = 0;
      rc = dmf_engine_apply_batch(resident, k, src.data(), dst.data(), w.data(), &rp, &f);
      if (rc) return fail(rc, "apply batch");
      dmf_counters cc;
      dmf_engine_counters(resident, &cc);
      std::printf("DYN TIME:%.6fms true_cnt=%d cnt=%d per batch time:%.6f\n", cc.device_ms, 1, 1,
                  cc.device_ms);
      std::printf("Dynamic flow:%d\n", (int)f);
      if (o.json) print_json("successive", i, rp.kc, f, cc);
      all_src.insert(all_src.end(), src.begin(), src.end());
      all_dst.insert(all_dst.end(), dst.begin(), dst.end());
      all_w.insert(all_w.end(), w.begin(), w.end());
      if (o.static_check) {
        long long stat = 0;
        if (static_run(g, all_src, all_dst, all_w, source, sink, rp.kc, &stat)) return 1;
        if (stat != f) {
          std::printf("<<<<<<<<<<<<<<<<<<<<<<<<PANIC NOT MATCHING dynflow:%d statflow:%d>>>>>>>>>>>>>>>>>>>>>>>>>>>\n",
                      (int)f, (int)stat);
          ++panics;
        }
      }
      continue;
    }
    const int avg = g.m / g.n;
    std::vector<int> kcs;
    if (o.kc_only > 0) {
      kcs.push_back(o.kc_only);
    } else {
      if (avg > 1) kcs.push_back(avg - 1);
      if (avg > 0) kcs.push_back(avg);
      kcs.push_back(avg + 1);
    }
    for (int kc : kcs) {
      std::printf("=======kernel cycles %d =======\n", kc);
      long long dyn = 0, stat = 0;
      if (dynamic_run(g, src, dst, w, source, sink, kc, o, i, &dyn)) return 1;
      if (!o.static_check) continue;
      if (static_run(g, src, dst, w, source, sink, kc, &stat)) return 1;
      if (dyn != stat) {
        std::printf("<<<<<<<<<<<<<<<<<<<<<<<<PANIC NOT MATCHING dynflow:%d statflow:%d>>>>>>>>>>>>>>>>>>>>>>>>>>>\n",
                    (int)dyn, (int)stat);
        ++panics;
      }
    }
  }
  if (resident) dmf_engine_destroy(resident);
  dmf_csr_free(&g);
  return panics ? 3 : 0;
}
