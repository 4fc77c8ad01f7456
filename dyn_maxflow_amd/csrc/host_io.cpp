// Host-side formats, CSR construction and seeded input generators.
//
// Formats follow the reference byte for byte:
//   graph .bin  : code/graph.hpp:202-223 (reader), input_create/create_bin.cpp:55-60 (writer)
//   update .bin : code/update.hpp:52-89 (reader), input_create/create_update_bin.cpp:46-49 (writer)
// The CSR builder reproduces create_bin's arc order (input_create/create_bin.cpp:95-119)
// with a hash index instead of a per-edge linear search.
#include "../../include/dmf.h"
#include "abi_guard.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

template <class T>
T* alloc_array(int64_t count) {
  if (count <= 0) count = 1;
  return static_cast<T*>(std::malloc(sizeof(T) * static_cast<size_t>(count)));
}

// ---------------------------------------------------------------- PRNG
struct SplitMix64 {
  uint64_t x;
  uint64_t next() {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
};

struct Xoshiro256ss {
  uint64_t s[4];
  explicit Xoshiro256ss(uint64_t seed) {
    SplitMix64 sm{seed};
    for (auto& w : s) w = sm.next();
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // Uniform double in [0,1) from the top 53 bits.
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
  // Uniform integer in [0, range) without bias (Lemire's multiply-shift with rejection).
  uint64_t below(uint64_t range) {
    uint64_t x = next();
    __uint128_t mprod = static_cast<__uint128_t>(x) * range;
    uint64_t low = static_cast<uint64_t>(mprod);
    if (low < range) {
      uint64_t thresh = (0 - range) % range;
      while (low < thresh) {
        x = next();
        mprod = static_cast<__uint128_t>(x) * range;
        low = static_cast<uint64_t>(mprod);
      }
    }
    return static_cast<uint64_t>(mprod >> 64);
  }
  int32_t between(int32_t lo, int32_t hi) {  // inclusive
    return lo + static_cast<int32_t>(below(static_cast<uint64_t>(hi - lo) + 1));
  }
};

// ----------------------------------------------------- open-addressing map
// Keys are (u << 32 | v) pairs; value is an int32 payload.  Empty slot = ~0.
struct PairMap {
  std::vector<uint64_t> keys;
  std::vector<int32_t> vals;
  uint64_t mask = 0;
  explicit PairMap(int64_t expected) {
    uint64_t cap = 16;
    while (cap < static_cast<uint64_t>(expected) * 2 + 16) cap <<= 1;
    keys.assign(cap, ~0ull);
    vals.assign(cap, -1);
    mask = cap - 1;
  }
  static uint64_t hash(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
  }
  // Returns pointer to the value slot; *inserted tells whether the key is new.
  int32_t* find_or_insert(uint64_t key, bool* inserted) {
    uint64_t h = hash(key) & mask;
    while (true) {
      if (keys[h] == key) {
        *inserted = false;
        return &vals[h];
      }
      if (keys[h] == ~0ull) {
        keys[h] = key;
        *inserted = true;
        return &vals[h];
      }
      h = (h + 1) & mask;
    }
  }
};

inline uint64_t pair_key(int32_t u, int32_t v) {
  return (static_cast<uint64_t>(static_cast<uint32_t>(u)) << 32) | static_cast<uint32_t>(v);
}

bool read_exact(FILE* f, void* dst, size_t bytes) {
  return std::fread(dst, 1, bytes, f) == bytes;
}

// Arc part of dmf_csr_check over tails [v0, v1); offsets already checked.
int check_arcs(const dmf_csr& g, int32_t v0, int32_t v1) {
  const int32_t n = g.n, m = g.m;
  const int32_t *row = g.row, *col = g.col, *cap = g.cap, *rev = g.rev;
  for (int32_t v = v0; v < v1; ++v) {
    for (int32_t e = row[v]; e < row[v + 1]; ++e) {
      const int32_t h = col[e], r = rev[e];
      if (h < 0 || h >= n || cap[e] < 0 || r < 0 || r >= m) return DMF_EFORMAT;
      if (rev[r] != e || col[r] != v) return DMF_EFORMAT;
    }
  }
  return DMF_OK;
}

}  // namespace

extern "C" {

const char* dmf_strerror(int status) {
  switch (status) {
    case DMF_OK: return "ok";
    case DMF_EINVAL: return "invalid argument";
    case DMF_EIO: return "i/o error";
    case DMF_ENOARC: return "update names an arc that is not in the graph";
    case DMF_ENOMEM: return "out of memory";
    case DMF_EDEVICE: return "HIP device error";
    case DMF_ENOCONV: return "round cap reached before convergence";
    case DMF_EFORMAT: return "malformed file or CSR";
    case DMF_EADMIT: return "request exceeds the admission budget";
    case DMF_EAGAIN: return "result not ready";
    case DMF_ESIZE: return "input exceeds the fixed workspace";
    default: return "unknown status";
  }
}

void dmf_free(void* p) { std::free(p); }

void dmf_csr_free(dmf_csr* g) {
  if (!g) return;
  std::free(g->row);
  std::free(g->col);
  std::free(g->cap);
  std::free(g->rev);
  g->row = g->col = g->cap = g->rev = nullptr;
  g->n = g->m = 0;
}

int dmf_csr_check(const dmf_csr* g) {
  return guarded([&]() -> int {
    if (!g || g->n < 0 || g->m < 0 || !g->row || (g->m > 0 && (!g->col || !g->cap || !g->rev)))
      return DMF_EINVAL;
    const int32_t n = g->n, m = g->m;
    const int32_t* row = g->row;
    if (row[0] != 0 || row[n] != m) return DMF_EFORMAT;
    for (int32_t v = 0; v < n; ++v)
      if (row[v] > row[v + 1]) return DMF_EFORMAT;
    // The arc pass reads rev[rev[e]] and col[rev[e]] at random, ~25 ns an arc on one
    // core, so large graphs split it over threads by arc-balanced vertex ranges.
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int32_t parts = m < (1 << 22) ? 1 : static_cast<int32_t>(std::min(hw, 16u));
    std::vector<int32_t> cut(parts + 1, n);
    cut[0] = 0;
    for (int32_t i = 1; i < parts; ++i) {
      const int64_t target = static_cast<int64_t>(m) * i / parts;
      cut[i] = static_cast<int32_t>(std::lower_bound(row, row + n + 1, target) - row);
      cut[i] = std::max(cut[i], cut[i - 1]);
    }
    std::vector<int> rc(parts, DMF_OK);
    std::vector<std::thread> pool;
    for (int32_t i = 1; i < parts; ++i) {
      try {
        pool.emplace_back([&, i] { rc[i] = check_arcs(*g, cut[i], cut[i + 1]); });
      } catch (...) {  // no thread available: check the range here
        rc[i] = check_arcs(*g, cut[i], cut[i + 1]);
      }
    }
    rc[0] = check_arcs(*g, cut[0], cut[1]);
    for (auto& th : pool) th.join();
    for (int r : rc)
      if (r != DMF_OK) return r;
    return DMF_OK;
  });
}

int dmf_read_graph_bin(const char* path, dmf_csr* out) {
  return guarded([&]() -> int {
    if (!path || !out) return DMF_EINVAL;
    std::memset(out, 0, sizeof(*out));
    FILE* f = std::fopen(path, "rb");
    if (!f) return DMF_EIO;
    int32_t hdr[2];
    if (!read_exact(f, hdr, sizeof(hdr))) {
      std::fclose(f);
      return DMF_EIO;
    }
    const int32_t n = hdr[0], m = hdr[1];
    if (n < 0 || m < 0) {
      std::fclose(f);
      return DMF_EFORMAT;
    }
    dmf_csr g{};
    g.n = n;
    g.m = m;
    g.row = alloc_array<int32_t>(static_cast<int64_t>(n) + 1);
    g.col = alloc_array<int32_t>(m);
    g.cap = alloc_array<int32_t>(m);
    g.rev = alloc_array<int32_t>(m);
    if (!g.row || !g.col || !g.cap || !g.rev) {
      dmf_csr_free(&g);
      std::fclose(f);
      return DMF_ENOMEM;
    }
    const bool ok = read_exact(f, g.row, sizeof(int32_t) * (static_cast<size_t>(n) + 1)) &&
                    read_exact(f, g.col, sizeof(int32_t) * static_cast<size_t>(m)) &&
                    read_exact(f, g.cap, sizeof(int32_t) * static_cast<size_t>(m)) &&
                    read_exact(f, g.rev, sizeof(int32_t) * static_cast<size_t>(m));
    std::fclose(f);
    if (!ok) {
      dmf_csr_free(&g);
      return DMF_EIO;
    }
    // structural validation the reference does not do
    if (dmf_csr_check(&g) != DMF_OK) {
      dmf_csr_free(&g);
      return DMF_EFORMAT;
    }
    *out = g;
    return DMF_OK;
  });
}

int dmf_write_graph_bin(const char* path, const dmf_csr* g) {
  return guarded([&]() -> int {
    if (!path || !g || g->n < 0 || g->m < 0) return DMF_EINVAL;
    FILE* f = std::fopen(path, "wb");
    if (!f) return DMF_EIO;
    const int32_t hdr[2] = {g->n, g->m};
    bool ok = std::fwrite(hdr, sizeof(int32_t), 2, f) == 2;
    ok = ok && std::fwrite(g->row, sizeof(int32_t), static_cast<size_t>(g->n) + 1, f) ==
                   static_cast<size_t>(g->n) + 1;
    ok = ok && std::fwrite(g->col, sizeof(int32_t), g->m, f) == static_cast<size_t>(g->m);
    ok = ok && std::fwrite(g->cap, sizeof(int32_t), g->m, f) == static_cast<size_t>(g->m);
    ok = ok && std::fwrite(g->rev, sizeof(int32_t), g->m, f) == static_cast<size_t>(g->m);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? DMF_OK : DMF_EIO;
  });
}

int dmf_build_csr(int64_t k, const int32_t* u, const int32_t* v, const int32_t* w,
                  dmf_csr* out) {
  return guarded([&]() -> int {
    if (!out || k < 0 || (k > 0 && (!u || !v || !w))) return DMF_EINVAL;
    std::memset(out, 0, sizeof(*out));
    int64_t max_id = -1;
    for (int64_t i = 0; i < k; ++i) {
      if (u[i] < 0 || v[i] < 0) return DMF_EINVAL;
      if (u[i] == v[i]) continue;  // create_bin.cpp:85
      max_id = std::max<int64_t>(max_id, std::max(u[i], v[i]));
    }
    const int64_t n = max_id + 1;  // create_bin.cpp:92
    if (n > INT32_MAX - 1) return DMF_EINVAL;

    // Pass 1: discover arcs in the order create_bin appends them to adj[] and give
    // each arc its position within its tail's list.  An arc is keyed (tail, head).
    std::vector<int32_t> deg(static_cast<size_t>(n), 0);
    struct Rec {
      int32_t tail, head, pos;
    };
    std::vector<Rec> order;  // arcs in creation order
    order.reserve(static_cast<size_t>(2 * k));
    std::vector<int64_t> capsum;  // capacity accumulated per created arc
    capsum.reserve(static_cast<size_t>(2 * k));
    PairMap arc_id(2 * k);  // (tail,head) -> creation index
    for (int64_t i = 0; i < k; ++i) {
      const int32_t a = u[i], b = v[i];
      if (a == b) continue;
      // arc ids are int32: stop before one would wrap (m <= INT32_MAX below)
      if (order.size() + 2 > static_cast<size_t>(INT32_MAX)) return DMF_EINVAL;
      bool ins = false;
      int32_t* id = arc_id.find_or_insert(pair_key(a, b), &ins);
      if (ins) {
        *id = static_cast<int32_t>(order.size());
        order.push_back({a, b, deg[a]++});
        capsum.push_back(w[i]);
      } else {
        capsum[*id] += w[i];  // duplicate weights summed (create_bin.cpp:107)
      }
      int32_t* rid = arc_id.find_or_insert(pair_key(b, a), &ins);
      if (ins) {  // reverse arc with capacity 0 (create_bin.cpp:116)
        *rid = static_cast<int32_t>(order.size());
        order.push_back({b, a, deg[b]++});
        capsum.push_back(0);
      }
    }
    const int64_t m = static_cast<int64_t>(order.size());

    dmf_csr g{};
    g.n = static_cast<int32_t>(n);
    g.m = static_cast<int32_t>(m);
    g.row = alloc_array<int32_t>(n + 1);
    g.col = alloc_array<int32_t>(m);
    g.cap = alloc_array<int32_t>(m);
    g.rev = alloc_array<int32_t>(m);
    if (!g.row || !g.col || !g.cap || !g.rev) {
      dmf_csr_free(&g);
      return DMF_ENOMEM;
    }
    g.row[0] = 0;
    for (int64_t x = 0; x < n; ++x) g.row[x + 1] = g.row[x] + deg[x];
    std::vector<int32_t> csr_index(static_cast<size_t>(m));
    for (int64_t id = 0; id < m; ++id) {
      const Rec& r = order[id];
      const int32_t e = g.row[r.tail] + r.pos;
      csr_index[id] = e;
      g.col[e] = r.head;
      if (capsum[id] > INT32_MAX || capsum[id] < INT32_MIN) {
        dmf_csr_free(&g);
        return DMF_EINVAL;
      }
      g.cap[e] = static_cast<int32_t>(capsum[id]);
    }
    for (int64_t id = 0; id < m; ++id) {
      const Rec& r = order[id];
      bool ins = false;
      const int32_t rid = *arc_id.find_or_insert(pair_key(r.head, r.tail), &ins);
      g.rev[csr_index[id]] = csr_index[rid];
    }
    *out = g;
    return DMF_OK;
  });
}

int dmf_read_updates_bin(const char* path, int32_t* k, int32_t** src, int32_t** dst,
                         int32_t** w) {
  return guarded([&]() -> int {
    if (!path || !k || !src || !dst || !w) return DMF_EINVAL;
    *k = 0;
    *src = *dst = *w = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return DMF_EIO;
    int32_t cnt = 0;
    if (!read_exact(f, &cnt, sizeof(cnt))) {
      std::fclose(f);
      return DMF_EIO;
    }
    if (cnt < 0) {
      std::fclose(f);
      return DMF_EFORMAT;
    }
    int32_t* a = alloc_array<int32_t>(cnt);
    int32_t* b = alloc_array<int32_t>(cnt);
    int32_t* c = alloc_array<int32_t>(cnt);
    if (!a || !b || !c) {
      std::free(a);
      std::free(b);
      std::free(c);
      std::fclose(f);
      return DMF_ENOMEM;
    }
    const size_t bytes = sizeof(int32_t) * static_cast<size_t>(cnt);
    const bool ok = read_exact(f, a, bytes) && read_exact(f, b, bytes) && read_exact(f, c, bytes);
    std::fclose(f);
    if (!ok) {
      std::free(a);
      std::free(b);
      std::free(c);
      return DMF_EIO;
    }
    *k = cnt;
    *src = a;
    *dst = b;
    *w = c;
    return DMF_OK;
  });
}

int dmf_write_updates_bin(const char* path, int32_t k, const int32_t* src, const int32_t* dst,
                          const int32_t* w) {
  return guarded([&]() -> int {
    if (!path || k < 0 || (k > 0 && (!src || !dst || !w))) return DMF_EINVAL;
    FILE* f = std::fopen(path, "wb");
    if (!f) return DMF_EIO;
    bool ok = std::fwrite(&k, sizeof(int32_t), 1, f) == 1;
    ok = ok && std::fwrite(src, sizeof(int32_t), k, f) == static_cast<size_t>(k);
    ok = ok && std::fwrite(dst, sizeof(int32_t), k, f) == static_cast<size_t>(k);
    ok = ok && std::fwrite(w, sizeof(int32_t), k, f) == static_cast<size_t>(k);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? DMF_OK : DMF_EIO;
  });
}

int dmf_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                 uint64_t seed, int32_t cap_lo, int32_t cap_hi, int64_t* k, int32_t** u,
                 int32_t** v, int32_t** w) {
  return guarded([&]() -> int {
    if (scale < 1 || scale > 30 || edge_factor < 1) return DMF_EINVAL;
    return dmf_gen_rmat_n(int32_t{1} << scale, (int64_t{1} << scale) * edge_factor, a, b, c, seed, cap_lo,
                          cap_hi, k, u, v, w);
  });
}

// R-MAT over [0, n): samples are drawn on the enclosing power of two and those
// with an endpoint >= n are dropped like self-loops and duplicates, so for a
// power-of-two n the output equals dmf_gen_rmat's.
int dmf_gen_rmat_n(int32_t n, int64_t samples, double a, double b, double c, uint64_t seed,
                   int32_t cap_lo, int32_t cap_hi, int64_t* k, int32_t** u, int32_t** v, int32_t** w) {
  return guarded([&]() -> int {
    if (n < 2 || n > (int32_t{1} << 30) || samples < 1 || !k || !u || !v || !w || cap_lo > cap_hi ||
        a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
      return DMF_EINVAL;
    int32_t scale = 1;
    while ((int64_t{1} << scale) < n) ++scale;
    Xoshiro256ss rng(seed);
    PairMap seen(samples);
    int32_t* U = alloc_array<int32_t>(samples);
    int32_t* V = alloc_array<int32_t>(samples);
    if (!U || !V) {
      std::free(U);
      std::free(V);
      return DMF_ENOMEM;
    }
    const double ab = a + b, abc = a + b + c;
    int64_t kept = 0;
    for (int64_t i = 0; i < samples; ++i) {
      uint32_t x = 0, y = 0;
      for (int32_t bit = 0; bit < scale; ++bit) {
        const double r = rng.uniform();
        uint32_t bx, by;
        if (r < a) {
          bx = 0; by = 0;
        } else if (r < ab) {
          bx = 0; by = 1;
        } else if (r < abc) {
          bx = 1; by = 0;
        } else {
          bx = 1; by = 1;
        }
        x = (x << 1) | bx;
        y = (y << 1) | by;
      }
      if (x == y || x >= static_cast<uint32_t>(n) || y >= static_cast<uint32_t>(n)) continue;
      bool ins = false;
      seen.find_or_insert(pair_key(static_cast<int32_t>(x), static_cast<int32_t>(y)), &ins);
      if (!ins) continue;
      U[kept] = static_cast<int32_t>(x);
      V[kept] = static_cast<int32_t>(y);
      ++kept;
    }
    int32_t* W = alloc_array<int32_t>(kept);
    if (!W) {
      std::free(U);
      std::free(V);
      return DMF_ENOMEM;
    }
    Xoshiro256ss crng(seed ^ 0xC0FFEE123456789ull);
    for (int64_t i = 0; i < kept; ++i) W[i] = crng.between(cap_lo, cap_hi);
    *k = kept;
    *u = U;
    *v = V;
    *w = W;
    return DMF_OK;
  });
}

int dmf_gen_grid(int32_t rows, int32_t cols, uint64_t seed, int32_t cap_lo, int32_t cap_hi,
                 int64_t* k, int32_t** u, int32_t** v, int32_t** w) {
  return guarded([&]() -> int {
    if (rows < 1 || cols < 1 || !k || !u || !v || !w || cap_lo > cap_hi) return DMF_EINVAL;
    if (static_cast<int64_t>(rows) * cols > INT32_MAX) return DMF_EINVAL;
    const int64_t total = 2 * (static_cast<int64_t>(rows) * (cols - 1) +
                               static_cast<int64_t>(cols) * (rows - 1));
    int32_t* U = alloc_array<int32_t>(total);
    int32_t* V = alloc_array<int32_t>(total);
    int32_t* W = alloc_array<int32_t>(total);
    if (!U || !V || !W) {
      std::free(U);
      std::free(V);
      std::free(W);
      return DMF_ENOMEM;
    }
    Xoshiro256ss rng(seed);
    int64_t i = 0;
    // Each vertex lists its out-edges in the order up, left, right, down.
    for (int32_t r = 0; r < rows; ++r) {
      for (int32_t c = 0; c < cols; ++c) {
        const int32_t id = r * cols + c;
        const int32_t nb[4] = {r > 0 ? id - cols : -1, c > 0 ? id - 1 : -1,
                               c + 1 < cols ? id + 1 : -1, r + 1 < rows ? id + cols : -1};
        for (int32_t j = 0; j < 4; ++j) {
          if (nb[j] < 0) continue;
          U[i] = id;
          V[i] = nb[j];
          W[i] = rng.between(cap_lo, cap_hi);
          ++i;
        }
      }
    }
    *k = i;
    *u = U;
    *v = V;
    *w = W;
    return DMF_OK;
  });
}

int dmf_pick_source_sink(int32_t n, int64_t k, const int32_t* u, const int32_t* v, int32_t* s,
                         int32_t* t) {
  return guarded([&]() -> int {
    if (n < 2 || k < 0 || !s || !t || (k > 0 && (!u || !v))) return DMF_EINVAL;
    std::vector<int64_t> outd(n, 0), ind(n, 0);
    for (int64_t i = 0; i < k; ++i) {
      if (u[i] < 0 || u[i] >= n || v[i] < 0 || v[i] >= n) return DMF_EINVAL;
      if (u[i] == v[i]) continue;
      ++outd[u[i]];
      ++ind[v[i]];
    }
    int32_t bs = 0;
    for (int32_t x = 1; x < n; ++x)
      if (outd[x] > outd[bs]) bs = x;
    int32_t bt = (bs == 0) ? 1 : 0;
    for (int32_t x = 0; x < n; ++x) {
      if (x == bs) continue;
      if (ind[x] > ind[bt]) bt = x;
    }
    *s = bs;
    *t = bt;
    return DMF_OK;
  });
}

int dmf_gen_updates(int64_t k, const int32_t* u, const int32_t* v, int32_t* cur_w, int64_t count,
                    double inc_frac, int32_t cap_lo, int32_t cap_hi, uint64_t seed, int apply,
                    int32_t* out_k, int32_t** src, int32_t** dst, int32_t** w) {
  return guarded([&]() -> int {
    if (k < 0 || count < 0 || !out_k || !src || !dst || !w || cap_lo >= cap_hi ||
        (k > 0 && (!u || !v || !cur_w)))
      return DMF_EINVAL;
    if (count > k) count = k;
    if (count > INT32_MAX) return DMF_EINVAL;
    int32_t* S = alloc_array<int32_t>(count);
    int32_t* D = alloc_array<int32_t>(count);
    int32_t* W = alloc_array<int32_t>(count);
    if (!S || !D || !W) {
      std::free(S);
      std::free(D);
      std::free(W);
      return DMF_ENOMEM;
    }
    Xoshiro256ss rng(seed);
    // Partial Fisher-Yates over edge indices: distinct edges, no replacement.
    // The permutation is kept sparse (positions displaced so far; every other
    // position still holds its own index), so a batch costs O(count), not O(k).
    std::unordered_map<int64_t, int64_t> moved;
    moved.reserve(static_cast<size_t>(std::min<int64_t>(k, 2 * count + 16)));
    auto at = [&moved](int64_t x) {
      const auto it = moved.find(x);
      return it == moved.end() ? x : it->second;
    };
    int64_t got = 0;
    for (int64_t i = 0; i < k && got < count; ++i) {
      const int64_t j = i + static_cast<int64_t>(rng.below(static_cast<uint64_t>(k - i)));
      const int64_t e = at(j);
      if (j != i) moved[j] = at(i);  // position i is never read again
      moved.erase(i);
      if (u[e] == v[e]) continue;
      const bool inc = rng.uniform() < inc_frac;
      const int32_t cur = cur_w[e];
      int32_t nw;
      if (inc) {
        if (cur >= cap_hi) continue;  // at the bound: draw another edge
        nw = rng.between(std::max(cur + 1, cap_lo), cap_hi);
      } else {
        if (cur <= cap_lo) continue;
        nw = rng.between(cap_lo, std::min(cur - 1, cap_hi));
      }
      S[got] = u[e];
      D[got] = v[e];
      W[got] = nw;
      if (apply) cur_w[e] = nw;
      ++got;
    }
    *out_k = static_cast<int32_t>(got);
    *src = S;
    *dst = D;
    *w = W;
    return DMF_OK;
  });
}

}  // extern "C"
