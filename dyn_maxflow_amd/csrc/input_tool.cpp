// dmf_input — seeded replacements for the reference's input pipeline
// (input_create/: weight.cpp, create_bin.cpp, update_gen.cpp + cut_p.cpp +
// create_txt_updates.sh, create_update_bin.cpp), writing the same file layouts.
//
//   dmf_input weight  <in.mtx|edges> <out.txt> <seed> [lo hi]   u v -> u v w, w ~ U[lo,hi] (default 1..100);
//                                                           '%' lines skipped (weight.cpp:35)
//   dmf_input csr     <edges.txt> <out.bin>                 create_bin.cpp: bidirectional CSR
//   dmf_input upd2bin <upd.txt> <out.bin>                   create_update_bin.cpp: first line k, then u v w
//   dmf_input rmat    <scale> <ef> <seed> <out_prefix> [lo hi]  <out_prefix>.txt + .bin, prints "s t"
//   dmf_input grid    <rows> <cols> <seed> <out_prefix> [lo hi] same, s = 0, t = n-1
//   dmf_input updates <edges.txt> <out_prefix> <pct_per_file> <inc_pct> <seed> [lo hi]
//        <out_prefix>_p1..p10 .txt/.bin, file i = first i*B records of one shuffled
//        update list, B = pct_per_file% of the input lines (cut_p.cpp:32-38 nesting)
//   dmf_input tree <edges.txt> <updates_dir> <prefix> <pct_per_file> <seed> [lo hi]
//        the reference's update tree (input_create/create_txt_updates.sh:17-49,
//        input_create/README.md:14-21): <updates_dir>/{dec,inc,mixed}/<prefix>_p1..p10
//        .txt/.bin as `updates` writes them, with 0%, 100% and 50% increases
//   dmf_input successive <edges.txt> <out_prefix> <inc_pct> <seed> [lo hi]
//        <out_prefix>_b1..b10 .bin, batch i = i% of the input lines, each drawn from
//        the capacities left by batches 1..i-1
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dmf.h"

namespace {

int die(const char* msg, int rc = 1) {
  std::fprintf(stderr, "dmf_input: %s\n", msg);
  return rc;
}

bool read_edges(const char* path, std::vector<int32_t>& u, std::vector<int32_t>& v,
                std::vector<int32_t>* w) {
  FILE* f = std::fopen(path, "r");
  if (!f) return false;
  char line[512];
  while (std::fgets(line, sizeof(line), f)) {
    if (line[0] == '%' || line[0] == '#' || line[0] == '\n') continue;
    long a, b, c = 0;
    int got = std::sscanf(line, "%ld %ld %ld", &a, &b, &c);
    if (got < 2 || (w && got < 3)) continue;
    u.push_back((int32_t)a);
    v.push_back((int32_t)b);
    if (w) w->push_back((int32_t)c);
  }
  std::fclose(f);
  return true;
}

bool write_edges(const std::string& path, const int32_t* u, const int32_t* v, const int32_t* w,
                 int64_t k, bool header) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return false;
  if (header) std::fprintf(f, "%lld\n", (long long)k);
  for (int64_t i = 0; i < k; ++i) std::fprintf(f, "%d %d %d\n", u[i], v[i], w[i]);
  return std::fclose(f) == 0;
}

int emit_graph(const std::string& prefix, int64_t k, int32_t* u, int32_t* v, int32_t* w,
               bool grid) {
  if (!write_edges(prefix + ".txt", u, v, w, k, false)) return die("cannot write edge list");
  dmf_csr g;
  int rc = dmf_build_csr(k, u, v, w, &g);
  if (rc) return die(dmf_strerror(rc));
  rc = dmf_write_graph_bin((prefix + ".bin").c_str(), &g);
  if (rc) return die(dmf_strerror(rc));
  int32_t s = 0, t = g.n - 1;
  if (!grid) dmf_pick_source_sink(g.n, k, u, v, &s, &t);
  std::printf("%d %d\n", s, t);
  dmf_csr_free(&g);
  return 0;
}

// Ten nested update files <prefix>_p1..p10 (.txt and .bin): file i holds the
// first i*B records of one update list of 10*B distinct input edges, B =
// pct_per_file% of the input lines (cut_p.cpp:32-38 nesting).
int write_nested(const std::string& prefix, const std::vector<int32_t>& u, const std::vector<int32_t>& v,
                 const std::vector<int32_t>& w, int pct_per_file, double inc, uint64_t seed, int lo, int hi) {
  const int64_t k = (int64_t)u.size();
  std::vector<int32_t> cur = w;
  const int64_t B = k * pct_per_file / 100;
  int32_t cnt = 0;
  int32_t *a = nullptr, *b = nullptr, *c = nullptr;
  int rc = dmf_gen_updates(k, u.data(), v.data(), cur.data(), 10 * B, inc, lo, hi, seed, 0, &cnt, &a, &b, &c);
  if (rc) return die(dmf_strerror(rc));
  for (int i = 1; i <= 10 && !rc; ++i) {
    const int64_t take = std::min<int64_t>(cnt, (int64_t)i * B);
    const std::string base = prefix + "_p" + std::to_string(i);
    if (!write_edges(base + ".txt", a, b, c, take, true)) {
      rc = die("cannot write updates");
      break;
    }
    const int wr = dmf_write_updates_bin((base + ".bin").c_str(), (int32_t)take, a, b, c);
    if (wr) rc = die(dmf_strerror(wr));
  }
  dmf_free(a);
  dmf_free(b);
  dmf_free(c);
  return rc;
}

bool make_dir(const std::string& path) { return mkdir(path.c_str(), 0755) == 0 || errno == EEXIST; }

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return die("usage: see the header of input_tool.cpp");
  const std::string cmd = argv[1];
  if (cmd == "weight" && argc >= 5) {
    std::vector<int32_t> u, v;
    if (!read_edges(argv[2], u, v, nullptr)) return die("cannot read input");
    const int lo = argc > 6 ? std::atoi(argv[5]) : 1, hi = argc > 6 ? std::atoi(argv[6]) : 100;
    // capacities from the seeded generator's stream: draw a 1-row grid-free
    // batch by generating an R-MAT-sized capacity vector is not needed; use
    // dmf_gen_grid's generator through a 1 x (k+1) grid's capacities.
    int64_t gk = 0;
    int32_t *gu = nullptr, *gv = nullptr, *gw = nullptr;
    int rc = dmf_gen_grid(1, (int32_t)(u.size() / 2 + 2), std::strtoull(argv[4], nullptr, 10), lo, hi,
                          &gk, &gu, &gv, &gw);
    if (rc || gk < (int64_t)u.size()) return die("capacity generation failed");
    if (!write_edges(argv[3], u.data(), v.data(), gw, (int64_t)u.size(), false))
      return die("cannot write output");
    dmf_free(gu);
    dmf_free(gv);
    dmf_free(gw);
    return 0;
  }
  if (cmd == "csr" && argc == 4) {
    std::vector<int32_t> u, v, w;
    if (!read_edges(argv[2], u, v, &w)) return die("cannot read input");
    dmf_csr g;
    int rc = dmf_build_csr((int64_t)u.size(), u.data(), v.data(), w.data(), &g);
    if (rc) return die(dmf_strerror(rc));
    rc = dmf_write_graph_bin(argv[3], &g);
    std::printf("CSR written to %s (%d nodes, %d edges)\n", argv[3], g.n, g.m);
    dmf_csr_free(&g);
    return rc ? die(dmf_strerror(rc)) : 0;
  }
  if (cmd == "upd2bin" && argc == 4) {
    FILE* f = std::fopen(argv[2], "r");
    if (!f) return die("cannot read input");
    long k = 0;
    if (std::fscanf(f, "%ld", &k) != 1 || k < 0) {
      std::fclose(f);
      return die("bad header");
    }
    std::vector<int32_t> a(k), b(k), c(k);
    for (long i = 0; i < k; ++i) {
      long x, y, z;
      if (std::fscanf(f, "%ld %ld %ld", &x, &y, &z) != 3) {
        std::fclose(f);
        return die("short update file");
      }
      a[i] = (int32_t)x;
      b[i] = (int32_t)y;
      c[i] = (int32_t)z;
    }
    std::fclose(f);
    int rc = dmf_write_updates_bin(argv[3], (int32_t)k, a.data(), b.data(), c.data());
    return rc ? die(dmf_strerror(rc)) : 0;
  }
  if ((cmd == "rmat" || cmd == "grid") && argc >= 6) {
    const int lo = argc > 7 ? std::atoi(argv[6]) : 1;
    const int hi = argc > 7 ? std::atoi(argv[7]) : (cmd == "grid" ? 1000 : 100);
    int64_t k = 0;
    int32_t *u = nullptr, *v = nullptr, *w = nullptr;
    int rc;
    if (cmd == "rmat")
      rc = dmf_gen_rmat(std::atoi(argv[2]), std::atoi(argv[3]), 0.57, 0.19, 0.19,
                        std::strtoull(argv[4], nullptr, 10), lo, hi, &k, &u, &v, &w);
    else
      rc = dmf_gen_grid(std::atoi(argv[2]), std::atoi(argv[3]), std::strtoull(argv[4], nullptr, 10), lo,
                        hi, &k, &u, &v, &w);
    if (rc) return die(dmf_strerror(rc));
    rc = emit_graph(argv[5], k, u, v, w, cmd == "grid");
    dmf_free(u);
    dmf_free(v);
    dmf_free(w);
    return rc;
  }
  if (cmd == "tree" && argc >= 7) {
    std::vector<int32_t> u, v, w;
    if (!read_edges(argv[2], u, v, &w)) return die("cannot read input");
    const std::string dir = argv[3], prefix = argv[4];
    const int pct = std::atoi(argv[5]);
    const uint64_t seed = std::strtoull(argv[6], nullptr, 10);
    const int lo = argc > 8 ? std::atoi(argv[7]) : 1, hi = argc > 8 ? std::atoi(argv[8]) : 100;
    if (!make_dir(dir)) return die("cannot create the updates directory");
    // the three update types of create_txt_updates.sh:17-24
    const struct { const char* name; double inc; } types[] = {{"dec", 0.0}, {"inc", 1.0}, {"mixed", 0.5}};
    for (int i = 0; i < 3; ++i) {
      const std::string sub = dir + "/" + types[i].name;
      if (!make_dir(sub)) return die("cannot create an update type directory");
      const int rc = write_nested(sub + "/" + prefix, u, v, w, pct, types[i].inc, seed + (uint64_t)i, lo, hi);
      if (rc) return rc;
    }
    return 0;
  }
  if ((cmd == "updates" && argc >= 7) || (cmd == "successive" && argc >= 6)) {
    std::vector<int32_t> u, v, w;
    if (!read_edges(argv[2], u, v, &w)) return die("cannot read input");
    const std::string prefix = argv[3];
    const bool succ = cmd == "successive";
    const double inc = std::atof(argv[succ ? 4 : 5]) / 100.0;
    const uint64_t seed = std::strtoull(argv[succ ? 5 : 6], nullptr, 10);
    const int ai = succ ? 6 : 7;
    const int lo = argc > ai + 1 ? std::atoi(argv[ai]) : 1;
    const int hi = argc > ai + 1 ? std::atoi(argv[ai + 1]) : 100;
    const int64_t k = (int64_t)u.size();
    std::vector<int32_t> cur = w;
    if (!succ) return write_nested(prefix, u, v, w, std::atoi(argv[4]), inc, seed, lo, hi);
    for (int i = 1; i <= 10; ++i) {
      int32_t cnt = 0;
      int32_t *a = nullptr, *b = nullptr, *c = nullptr;
      int rc = dmf_gen_updates(k, u.data(), v.data(), cur.data(), k * i / 100, inc, lo, hi,
                               seed * 1000 + i, 1, &cnt, &a, &b, &c);
      if (rc) return die(dmf_strerror(rc));
      rc = dmf_write_updates_bin((prefix + "_b" + std::to_string(i) + ".bin").c_str(), cnt, a, b, c);
      dmf_free(a);
      dmf_free(b);
      dmf_free(c);
      if (rc) return die(dmf_strerror(rc));
    }
    return 0;
  }
  return die("unknown command or wrong argument count; see input_tool.cpp");
}
