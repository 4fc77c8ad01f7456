/*
 * oracle.c — CPU restatement of the reference's dynamic push-relabel, plus an
 * int64 Dinic max-flow, used ONLY to compute expected results in tests.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only to check or to time a CPU
 * baseline.  The engine under dyn_maxflow_amd/ never calls it.
 *
 * Contents
 *   orc_dinic          int64 Dinic on the bidirectional CSR; S_min = BFS from s
 *                      over arcs with residual > 0 after the last phase.
 *   orc_pr_*           sequential restatement of the three reference variants:
 *                        data : code/dyn_data.h, code/dyn_data_bench.cu:8-170
 *                        topo : code/dyn_topo.h, code/dyn_topo_bench.cu:16-119
 *                        pp   : code/dyn_pp.h,   code/dyn_pp_bench.cu:6-258
 *                      Each GPU kernel is executed as one sequential pass in
 *                      ascending thread order, which is one legal schedule of
 *                      the reference's racy kernels.
 *   orc_pr_*_omp       the data variant with the push and repair passes run by
 *                      OpenMP threads using atomic read-modify-writes (the CPU
 *                      baseline the GPU numbers are reported beside).
 *
 * Pinning (see tests/test_oracle.py): Dinic against scipy's maximum_flow and
 * networkx on random graphs, against brute-force cut enumeration for n <= 12
 * and against closed forms (paths, parallel paths, grids, bipartite matching);
 * the push-relabel restatement against Dinic (flow value and S_min) after every
 * batch; formats against the reference's own create_bin/create_update_bin
 * compiled into oracle/_ref.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ===================================================================== Dinic */

/* Max-flow value into *flow and S_min (1 = source side) into side (may be NULL).
 * Arc residuals start at cap[e]; pushing f on e moves f from r[e] to r[rev[e]],
 * the standard paired-arc residual graph of the bidirectional CSR. */
int orc_dinic(int32_t n, int32_t m, const int32_t* row, const int32_t* col,
              const int32_t* cap, const int32_t* rev, int32_t s, int32_t t,
              int64_t* flow, uint8_t* side) {
  if (n <= 0 || s < 0 || s >= n || t < 0 || t >= n || s == t) return -1;
  int64_t* r = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
  int32_t* level = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* queue = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* it = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* stack_v = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* stack_e = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  if (!r || !level || !queue || !it || !stack_v || !stack_e) {
    free(r); free(level); free(queue); free(it); free(stack_v); free(stack_e);
    return -2;
  }
  for (int32_t e = 0; e < m; ++e) r[e] = cap[e];
  int64_t total = 0;
  for (;;) {
    for (int32_t v = 0; v < n; ++v) level[v] = -1;
    int32_t qh = 0, qt = 0;
    level[s] = 0;
    queue[qt++] = s;
    while (qh < qt) {
      int32_t v = queue[qh++];
      for (int32_t e = row[v]; e < row[v + 1]; ++e) {
        int32_t w = col[e];
        if (r[e] > 0 && level[w] < 0) {
          level[w] = level[v] + 1;
          queue[qt++] = w;
        }
      }
    }
    if (level[t] < 0) break;
    for (int32_t v = 0; v < n; ++v) it[v] = row[v];
    /* Blocking flow: iterative DFS along level-increasing arcs. */
    for (;;) {
      int32_t depth = 0;
      stack_v[0] = s;
      int found = 0;
      while (depth >= 0) {
        int32_t v = stack_v[depth];
        if (v == t) { found = 1; break; }
        int advanced = 0;
        for (; it[v] < row[v + 1]; ++it[v]) {
          int32_t e = it[v];
          int32_t w = col[e];
          if (r[e] > 0 && level[w] == level[v] + 1) {
            stack_e[depth] = e;
            stack_v[++depth] = w;
            advanced = 1;
            break;
          }
        }
        if (!advanced) {
          level[v] = -1; /* dead end: prune */
          --depth;
          if (depth >= 0) ++it[stack_v[depth]];
        }
      }
      if (!found) break;
      int64_t f = INT64_MAX;
      for (int32_t i = 0; i < depth; ++i)
        if (r[stack_e[i]] < f) f = r[stack_e[i]];
      for (int32_t i = 0; i < depth; ++i) {
        r[stack_e[i]] -= f;
        r[rev[stack_e[i]]] += f;
      }
      total += f;
    }
  }
  if (side) {
    memset(side, 0, (size_t)n);
    int32_t qh = 0, qt = 0;
    side[s] = 1;
    queue[qt++] = s;
    while (qh < qt) {
      int32_t v = queue[qh++];
      for (int32_t e = row[v]; e < row[v + 1]; ++e) {
        int32_t w = col[e];
        if (r[e] > 0 && !side[w]) {
          side[w] = 1;
          queue[qt++] = w;
        }
      }
    }
  }
  *flow = total;
  free(r); free(level); free(queue); free(it); free(stack_v); free(stack_e);
  return 0;
}

/* ========================================================= push-relabel state */

typedef struct orc_pr {
  int32_t n, m, s, t;
  int32_t *row, *col, *rev; /* owned copies */
  int32_t *cap0;            /* original capacities (reset) */
  int32_t *cap;             /* d_weight: current capacities */
  int32_t *res;             /* d_residual_capacity */
  int32_t *rres;            /* d_rev_residual_capacity: mirror of res[rev[e]] */
  int64_t *excess;          /* d_excess (int64 here; int32 in the reference) */
  int32_t *height, *level;
  int32_t *wl, *wl2, *part;
  uint8_t *partB, *partfinal;
  int64_t rounds, pushes, relabels, repairs, bfs_levels, arc_visits;
  int32_t max_rounds;
  int32_t threads; /* 0/1 = sequential, >1 = OpenMP data variant */
} orc_pr;

static void* xmalloc(size_t bytes) { return malloc(bytes ? bytes : 1); }

orc_pr* orc_pr_create(int32_t n, int32_t m, const int32_t* row, const int32_t* col,
                      const int32_t* cap, const int32_t* rev, int32_t s, int32_t t) {
  if (n <= 0 || m < 0 || s < 0 || s >= n || t < 0 || t >= n || s == t) return NULL;
  orc_pr* g = (orc_pr*)calloc(1, sizeof(orc_pr));
  if (!g) return NULL;
  g->n = n; g->m = m; g->s = s; g->t = t;
  size_t nb = (size_t)n, mb = (size_t)m;
  g->row = (int32_t*)xmalloc(sizeof(int32_t) * (nb + 1));
  g->col = (int32_t*)xmalloc(sizeof(int32_t) * mb);
  g->rev = (int32_t*)xmalloc(sizeof(int32_t) * mb);
  g->cap0 = (int32_t*)xmalloc(sizeof(int32_t) * mb);
  g->cap = (int32_t*)xmalloc(sizeof(int32_t) * mb);
  g->res = (int32_t*)xmalloc(sizeof(int32_t) * mb);
  g->rres = (int32_t*)xmalloc(sizeof(int32_t) * mb);
  g->excess = (int64_t*)xmalloc(sizeof(int64_t) * nb);
  g->height = (int32_t*)xmalloc(sizeof(int32_t) * nb);
  g->level = (int32_t*)xmalloc(sizeof(int32_t) * nb);
  g->wl = (int32_t*)xmalloc(sizeof(int32_t) * nb);
  g->wl2 = (int32_t*)xmalloc(sizeof(int32_t) * nb);
  g->part = (int32_t*)xmalloc(sizeof(int32_t) * nb);
  g->partB = (uint8_t*)calloc(nb, 1);
  g->partfinal = (uint8_t*)calloc(nb, 1);
  memcpy(g->row, row, sizeof(int32_t) * (nb + 1));
  memcpy(g->col, col, sizeof(int32_t) * mb);
  memcpy(g->rev, rev, sizeof(int32_t) * mb);
  memcpy(g->cap0, cap, sizeof(int32_t) * mb);
  memcpy(g->cap, cap, sizeof(int32_t) * mb);
  memset(g->res, 0, sizeof(int32_t) * mb);
  memset(g->rres, 0, sizeof(int32_t) * mb);
  memset(g->excess, 0, sizeof(int64_t) * nb);
  g->max_rounds = 1 << 30;
  g->threads = 1;
  return g;
}

void orc_pr_destroy(orc_pr* g) {
  if (!g) return;
  free(g->row); free(g->col); free(g->rev); free(g->cap0); free(g->cap);
  free(g->res); free(g->rres); free(g->excess); free(g->height); free(g->level);
  free(g->wl); free(g->wl2); free(g->part); free(g->partB); free(g->partfinal);
  free(g);
}

void orc_pr_set_threads(orc_pr* g, int32_t threads) { g->threads = threads; }
void orc_pr_set_max_rounds(orc_pr* g, int32_t r) { g->max_rounds = r; }

void orc_pr_reset(orc_pr* g) {
  memcpy(g->cap, g->cap0, sizeof(int32_t) * (size_t)g->m);
  memset(g->partB, 0, (size_t)g->n);
}

/* K1: staticMaxFlow_kernel_1 (code/dyn_data.h:25-35) */
static void k_init(orc_pr* g) {
  for (int32_t e = 0; e < g->m; ++e) {
    g->res[e] = g->cap[e];
    g->rres[e] = g->cap[g->rev[e]];
  }
  for (int32_t v = 0; v < g->n; ++v) g->excess[v] = 0;
}

/* K2: staticMaxFlow_kernel_7 (code/dyn_data.h:39-56): saturate every residual
 * arc out of s. */
static void k_saturate_source(orc_pr* g) {
  const int32_t v = g->s;
  for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
    int32_t w = g->col[e];
    int32_t d = g->res[e];
    if (d > 0) {
      g->excess[w] += d;
      g->excess[v] -= d;
      g->res[e] -= d;
      g->rres[e] += d;
      int32_t p = g->rev[e];
      g->res[p] += d;
      g->rres[p] -= d;
    }
  }
}

/* K2p: pp staticMaxFlow_kernel_7 (code/dyn_pp.h:39-72): K2 plus pulling every
 * residual arc into t. */
static void k_saturate_source_sink_pp(orc_pr* g) {
  k_saturate_source(g);
  const int32_t v = g->t;
  for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
    int32_t w = g->col[e];
    int32_t d = g->rres[e];
    if (d > 0) {
      g->excess[v] += d;
      g->excess[w] -= d;
      g->res[e] += d;
      g->rres[e] -= d;
      int32_t p = g->rev[e];
      g->res[p] -= d;
      g->rres[p] += d;
    }
  }
}

/* K3 + K4: my_init_1 (code/dyn_data.h:60-67) and the level loop of
 * staticMaxFlow_kernel_11 (code/dyn_data.h:72-87) driven as in
 * code/dyn_data_bench.cu:33-47.  Roots: t and every v != s with excess < 0.
 * The topo roots (staticMaxFlow_kernel_10, code/dyn_topo.h:60-71) are the same
 * set. */
static void global_relabel(orc_pr* g) {
  const int32_t n = g->n;
  for (int32_t v = 0; v < n; ++v) {
    g->height[v] = n;
    g->level[v] = -1;
    if (v == g->t) g->level[v] = 0;
    if (v != g->s && g->excess[v] < 0) g->level[v] = 0;
  }
  int32_t hops = 0;
  int flag;
  do {
    flag = 0;
    for (int32_t v = 0; v < n; ++v) {
      if (g->level[v] != hops) continue;
      g->height[v] = hops;
      for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
        int32_t w = g->col[e];
        g->arc_visits++;
        if (g->rres[e] > 0 && g->level[w] == -1) {
          g->level[w] = hops + 1;
          flag = 1;
        }
      }
    }
    ++hops;
    g->bfs_levels++;
  } while (flag);
}

/* K5: create_wl (code/dyn_data.h:217-238) in ascending vertex order. */
static int32_t build_worklist(orc_pr* g) {
  int32_t cnt = 0;
  for (int32_t v = 0; v < g->n; ++v)
    if (g->excess[v] > 0 && g->height[v] < g->n && v != g->s && v != g->t) g->wl[cnt++] = v;
  return cnt;
}

/* One push-or-relabel step of staticMaxFlow_kernel_14 (code/dyn_data.h:94-129).
 * Returns 0 when the vertex is no longer active (the kernel's `break`). */
static int push_relabel_step(orc_pr* g, int32_t v) {
  int64_t ex1 = g->excess[v];
  if (!(ex1 > 0 && g->height[v] < g->n)) return 0;
  int32_t hh = INT_MAX, v0 = -1, fe = -1;
  for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
    int32_t vv = g->col[e];
    int32_t h1 = g->height[vv];
    g->arc_visits++;
    if (h1 < hh && g->res[e] > 0) {
      v0 = vv;
      hh = h1;
      fe = e;
    }
  }
  if (g->height[v] > hh && v0 != -1) {
    int32_t fec = g->res[fe];
    int32_t p = g->rev[fe];
    int32_t d = (ex1 < fec) ? (int32_t)ex1 : fec;
    g->excess[v] -= d;
    g->excess[v0] += d;
    g->res[fe] -= d;
    g->rres[fe] += d;
    g->res[p] += d;
    g->rres[p] -= d;
    g->pushes++;
  } else if (v0 != -1) {
    g->height[v] = hh + 1;
    g->relabels++;
  }
  return 1;
}

/* K6 (data): staticMaxFlow_kernel_14 (code/dyn_data.h:89-131), kc cycles. */
static void push_pass(orc_pr* g, const int32_t* list, int32_t cnt, int32_t kc) {
  for (int32_t i = 0; i < cnt; ++i) {
    int32_t v = list[i];
    int32_t cycle = kc;
    while (cycle--)
      if (!push_relabel_step(g, v)) break;
  }
}

/* K7: staticMaxFlow_kernel_17 (code/dyn_data.h:132-153) for one vertex.
 * only_mask restricts to neighbours with only_mask[vv] == want (pp variants). */
static void repair_vertex(orc_pr* g, int32_t v, const uint8_t* only_mask, uint8_t want) {
  int32_t h = g->height[v];
  for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
    int32_t vv = g->col[e];
    if (only_mask && only_mask[vv] != want) continue;
    g->arc_visits++;
    if (g->rres[e] > 0 && g->height[vv] > h + 1) {
      int32_t d = g->rres[e];
      g->rres[e] -= d;
      g->res[e] += d;
      int32_t p = g->rev[e];
      g->rres[p] += d;
      g->res[p] -= d;
      g->excess[vv] -= d;
      g->excess[v] += d;
      g->repairs++;
    }
  }
}

/* K8 + K9: incremental_kernel_20 / incremental_kernel_22 (code/dyn_data.h:157-183). */
static void clamp_and_recompute(orc_pr* g) {
  for (int32_t v = 0; v < g->n; ++v) {
    for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
      if (g->res[e] < 0) {
        g->rres[e] += g->res[e];
        g->res[e] = 0;
      }
      if (g->rres[e] < 0) {
        g->res[e] += g->rres[e];
        g->rres[e] = 0;
      }
    }
  }
  for (int32_t v = 0; v < g->n; ++v) {
    int64_t e1 = 0;
    for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) e1 += (int64_t)g->res[e] - g->cap[e];
    g->excess[v] = e1;
  }
}

/* K10: OnAdd_kernel (code/dyn_data.h:186-206), applied in record order (so a
 * later record for the same arc wins).  Returns -3 if an arc is missing (the
 * reference writes index -1 there); the state is then left unchanged. */
static int apply_updates(orc_pr* g, int32_t k, const int32_t* src, const int32_t* dst,
                         const int32_t* w) {
  int32_t* arc = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(k > 0 ? k : 1));
  for (int32_t i = 0; i < k; ++i) {
    arc[i] = -1;
    int32_t a = src[i], b = dst[i];
    if (a < 0 || a >= g->n || b < 0 || b >= g->n) { free(arc); return -1; }
    if (a == b) continue;
    for (int32_t e = g->row[a]; e < g->row[a + 1]; ++e)
      if (g->col[e] == b) arc[i] = e; /* last match, as the reference */
    if (arc[i] < 0) { free(arc); return -3; }
  }
  for (int32_t i = 0; i < k; ++i) {
    int32_t e = arc[i];
    if (e < 0) continue;
    int32_t delta = w[i] - g->cap[e];
    g->res[e] += delta;
    g->rres[g->rev[e]] += delta;
    g->cap[e] = w[i];
  }
  free(arc);
  return 0;
}

/* update_kernel (code/dyn_data_bench.cu:300-322): capacities only. */
static int apply_caps_only(orc_pr* g, int32_t k, const int32_t* src, const int32_t* dst,
                           const int32_t* w) {
  for (int32_t i = 0; i < k; ++i) {
    int32_t a = src[i], b = dst[i];
    if (a < 0 || a >= g->n || b < 0 || b >= g->n) return -1;
    if (a == b) continue;
    int32_t arc = -1;
    for (int32_t e = g->row[a]; e < g->row[a + 1]; ++e)
      if (g->col[e] == b) arc = e;
    if (arc < 0) return -3;
    g->cap[arc] = w[i];
  }
  return 0;
}

/* ------------------------------------------------ OpenMP data-variant passes */
#define AADD32(p, v) __atomic_fetch_add((p), (v), __ATOMIC_RELAXED)
#define AADD64(p, v) __atomic_fetch_add((p), (v), __ATOMIC_RELAXED)
#define ALD32(p) __atomic_load_n((p), __ATOMIC_RELAXED)
#define ALD64(p) __atomic_load_n((p), __ATOMIC_RELAXED)

static void push_pass_omp(orc_pr* g, int32_t cnt, int32_t kc) {
  const int32_t n = g->n;
  int64_t pushes = 0, relabels = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(g->threads) reduction(+ : pushes, relabels)
  for (int32_t i = 0; i < cnt; ++i) {
    int32_t v = g->wl[i];
    int32_t cycle = kc;
    while (cycle--) {
      int64_t ex1 = ALD64(&g->excess[v]);
      int32_t hv = ALD32(&g->height[v]);
      if (!(ex1 > 0 && hv < n)) break;
      int32_t hh = INT_MAX, v0 = -1, fe = -1;
      for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
        int32_t vv = g->col[e];
        int32_t h1 = ALD32(&g->height[vv]);
        if (h1 < hh && ALD32(&g->res[e]) > 0) { v0 = vv; hh = h1; fe = e; }
      }
      if (hv > hh && v0 != -1) {
        int32_t fec = ALD32(&g->res[fe]);
        int32_t p = g->rev[fe];
        int32_t d = (ex1 < fec) ? (int32_t)ex1 : fec;
        AADD64(&g->excess[v], -(int64_t)d);
        AADD64(&g->excess[v0], (int64_t)d);
        AADD32(&g->res[fe], -d);
        AADD32(&g->rres[fe], d);
        AADD32(&g->res[p], d);
        AADD32(&g->rres[p], -d);
        pushes++;
      } else if (v0 != -1) {
        __atomic_store_n(&g->height[v], hh + 1, __ATOMIC_RELAXED);
        relabels++;
      }
    }
  }
  g->pushes += pushes;
  g->relabels += relabels;
}

static void repair_pass_omp(orc_pr* g, int32_t cnt) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(g->threads)
  for (int32_t i = 0; i < cnt; ++i) {
    int32_t v = g->wl[i];
    int32_t h = g->height[v];
    for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
      int32_t vv = g->col[e];
      int32_t d = ALD32(&g->rres[e]);
      if (d > 0 && g->height[vv] > h + 1) {
        int32_t p = g->rev[e];
        AADD32(&g->rres[e], -d);
        AADD32(&g->res[e], d);
        AADD32(&g->rres[p], d);
        AADD32(&g->res[p], -d);
        AADD64(&g->excess[vv], -(int64_t)d);
        AADD64(&g->excess[v], (int64_t)d);
      }
    }
  }
}

static void global_relabel_omp(orc_pr* g) {
  const int32_t n = g->n;
#pragma omp parallel for schedule(static) num_threads(g->threads)
  for (int32_t v = 0; v < n; ++v) {
    g->height[v] = n;
    g->level[v] = (v == g->t || (v != g->s && g->excess[v] < 0)) ? 0 : -1;
  }
  int32_t hops = 0;
  int flag;
  do {
    flag = 0;
#pragma omp parallel for schedule(dynamic, 1024) num_threads(g->threads) reduction(| : flag)
    for (int32_t v = 0; v < n; ++v) {
      if (g->level[v] != hops) continue;
      g->height[v] = hops;
      for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
        int32_t w = g->col[e];
        if (g->rres[e] > 0 && ALD32(&g->level[w]) == -1) {
          __atomic_store_n(&g->level[w], hops + 1, __ATOMIC_RELAXED);
          flag = 1;
        }
      }
    }
    ++hops;
    g->bfs_levels++;
  } while (flag);
}

static int32_t build_worklist_omp(orc_pr* g) {
  /* order-free compaction; ascending order is not needed for correctness */
  int32_t cnt = 0;
  for (int32_t v = 0; v < g->n; ++v)
    if (g->excess[v] > 0 && g->height[v] < g->n && v != g->s && v != g->t) g->wl[cnt++] = v;
  return cnt;
}

/* ============================================================ data variant */

/* The solve loop of staticMaxFlow / Incremental (code/dyn_data_bench.cu:33-87,
 * :113-167): relabel, worklist, then {push; repair; relabel; worklist} until the
 * worklist is empty.  Returns 0, or -6 when max_rounds is reached. */
static int data_loop(orc_pr* g, int32_t kc) {
  int omp = g->threads > 1;
  if (omp) global_relabel_omp(g); else global_relabel(g);
  int32_t active = omp ? build_worklist_omp(g) : build_worklist(g);
  while (active > 0) {
    if (g->rounds >= g->max_rounds) return -6;
    if (omp) {
      push_pass_omp(g, active, kc);
      repair_pass_omp(g, active);
      global_relabel_omp(g);
      active = build_worklist_omp(g);
    } else {
      push_pass(g, g->wl, active, kc);
      for (int32_t i = 0; i < active; ++i) repair_vertex(g, g->wl[i], NULL, 0);
      global_relabel(g);
      active = build_worklist(g);
    }
    g->rounds++;
  }
  return 0;
}

/* ============================================================ topo variant */

/* staticMaxFlow_kernel_14 (topo, code/dyn_topo.h:90-136) over all vertices.
 * Returns the flag1 value (some vertex was active). */
static int topo_push_pass(orc_pr* g, int32_t kc) {
  int flag1 = 0;
  for (int32_t v = 0; v < g->n; ++v) {
    if (g->excess[v] > 0 && v != g->s && v != g->t && g->height[v] < g->n) {
      flag1 = 1;
      int32_t cycle = kc;
      do {
        if (!push_relabel_step(g, v)) break;
        cycle = cycle - 1;
      } while (cycle > 0);
    }
  }
  return flag1;
}

/* code/dyn_topo_bench.cu:41-61 / :94-117: relabel, then push and repair over all
 * vertices, until the push pass saw no active vertex. */
static int topo_loop(orc_pr* g, int32_t kc) {
  int flag1;
  do {
    if (g->rounds >= g->max_rounds) return -6;
    global_relabel(g);
    flag1 = topo_push_pass(g, kc);
    for (int32_t v = 0; v < g->n; ++v) repair_vertex(g, v, NULL, 0); /* code/dyn_topo.h:137-157 */
    g->rounds++;
  } while (flag1);
  return 0;
}

/* ============================================================== pp variant */

/* pp_bfs_init + pp_bfs (code/dyn_pp.h:76-83, :99-120), level loop as
 * code/dyn_pp_bench.cu:22-31. */
static void pp_relabel(orc_pr* g) {
  const int32_t n = g->n;
  for (int32_t v = 0; v < n; ++v) {
    g->level[v] = -1;
    g->height[v] = n;
    if (v == g->s || v == g->t) g->level[v] = 0;
    if ((g->partB[v] && g->excess[v] < 0) || (!g->partB[v] && g->excess[v] > 0)) g->level[v] = 0;
  }
  int32_t hops = 0;
  int flag;
  do {
    flag = 0;
    for (int32_t v = 0; v < n; ++v) {
      if (g->level[v] != hops) continue;
      int push = g->partB[v];
      g->height[v] = hops;
      for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
        int32_t w = g->col[e];
        if (w == g->s || w == g->t) continue;
        g->arc_visits++;
        if (push && g->rres[e] > 0 && g->level[w] == -1) { g->level[w] = hops + 1; flag = 1; }
        if (!push && g->res[e] > 0 && g->level[w] == -1) { g->level[w] = hops + 1; flag = 1; }
      }
    }
    ++hops;
    g->bfs_levels++;
  } while (flag);
}

/* create_pp_wl (code/dyn_pp.h:499-545): push list (B, excess > 0) into wl,
 * pull list (A, excess < 0) into wl2. */
static void pp_worklists(orc_pr* g, int32_t* np, int32_t* nq) {
  int32_t a = 0, b = 0;
  for (int32_t v = 0; v < g->n; ++v) {
    if (v == g->s || v == g->t) continue;
    if (g->partB[v]) {
      if (g->excess[v] > 0 && g->height[v] < g->n) g->wl[a++] = v;
    } else {
      if (g->excess[v] < 0 && g->height[v] < g->n) g->wl2[b++] = v;
    }
  }
  *np = a;
  *nq = b;
}

/* pp_push_kernel (code/dyn_pp.h:191-232): K6 restricted to B neighbours. */
static void pp_push_pass(orc_pr* g, int32_t cnt, int32_t kc) {
  for (int32_t i = 0; i < cnt; ++i) {
    int32_t v = g->wl[i];
    int32_t cycle = kc;
    while (cycle--) {
      int64_t ex1 = g->excess[v];
      if (!(ex1 > 0 && g->height[v] < g->n)) break;
      int32_t hh = INT_MAX, v0 = -1, fe = -1;
      for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
        int32_t vv = g->col[e];
        if (!g->partB[vv]) continue;
        g->arc_visits++;
        int32_t h1 = g->height[vv];
        if (h1 < hh && g->res[e] > 0) { v0 = vv; hh = h1; fe = e; }
      }
      if (g->height[v] > hh && v0 != -1) {
        int32_t fec = g->res[fe];
        int32_t p = g->rev[fe];
        int32_t d = (ex1 < fec) ? (int32_t)ex1 : fec;
        g->excess[v] -= d; g->excess[v0] += d;
        g->res[fe] -= d; g->rres[fe] += d;
        g->res[p] += d; g->rres[p] -= d;
        g->pushes++;
      } else if (v0 != -1) {
        g->height[v] = hh + 1;
        g->relabels++;
      }
    }
  }
}

/* pull_kernel (code/dyn_pp.h:237-280): a deficit vertex in A pulls from its
 * lowest A neighbour that can send to it, or relabels. */
static void pp_pull_pass(orc_pr* g, int32_t cnt, int32_t kc) {
  for (int32_t i = 0; i < cnt; ++i) {
    int32_t v = g->wl2[i];
    int32_t cycle = kc;
    while (cycle--) {
      int64_t ex1 = -g->excess[v];
      if (!(ex1 > 0)) break;
      int32_t hh = INT_MAX, v0 = -1, fe = -1;
      for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
        int32_t vv = g->col[e];
        if (g->partB[vv]) continue;
        g->arc_visits++;
        int32_t h1 = g->height[vv];
        if (h1 < hh && g->rres[e] > 0) { v0 = vv; hh = h1; fe = e; }
      }
      if (g->height[v] > hh && v0 != -1) {
        int32_t rec = g->rres[fe];
        int32_t p = g->rev[fe];
        int32_t d = (ex1 < rec) ? (int32_t)ex1 : rec;
        g->excess[v0] -= d; g->excess[v] += d;
        g->rres[fe] -= d; g->res[fe] += d;
        g->rres[p] += d; g->res[p] -= d;
        g->pushes++;
      } else if (v0 != -1) {
        g->height[v] = hh + 1;
        g->relabels++;
      }
    }
  }
}

/* remove_pull (code/dyn_pp.h:381-403). */
static void pp_remove_pull_vertex(orc_pr* g, int32_t v) {
  int32_t h = g->height[v];
  for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
    int32_t vv = g->col[e];
    if (g->partB[vv]) continue;
    g->arc_visits++;
    if (g->res[e] > 0 && g->height[vv] > h + 1) {
      int32_t d = g->res[e];
      g->rres[e] += d; g->res[e] -= d;
      int32_t p = g->rev[e];
      g->rres[p] -= d; g->res[p] += d;
      g->excess[v] -= d; g->excess[vv] += d;
      g->repairs++;
    }
  }
}

/* part_pp (code/dyn_pp_bench.cu:6-88). */
static int pp_part(orc_pr* g, int32_t kc) {
  int32_t np, nq;
  pp_relabel(g);
  pp_worklists(g, &np, &nq);
  while (np + nq > 0) {
    if (g->rounds >= g->max_rounds) return -6;
    if (np > 0) {
      pp_push_pass(g, np, kc);
      for (int32_t i = 0; i < np; ++i) repair_vertex(g, g->wl[i], g->partB, 1);
    }
    if (nq > 0) {
      pp_pull_pass(g, nq, kc);
      for (int32_t i = 0; i < nq; ++i) pp_remove_pull_vertex(g, g->wl2[i]);
    }
    pp_relabel(g);
    pp_worklists(g, &np, &nq);
    g->rounds++;
  }
  return 0;
}

/* final_bfs_init + final_bfs (code/dyn_pp.h:86-93, :123-140) over the part list. */
static void pp_final_relabel(orc_pr* g, int32_t nodes) {
  for (int32_t i = 0; i < nodes; ++i) {
    int32_t v = g->part[i];
    g->level[v] = -1;
    g->height[v] = g->n;
    if (g->excess[v] < 0) g->level[v] = 0;
  }
  int32_t hops = 0;
  int flag;
  do {
    flag = 0;
    for (int32_t i = 0; i < nodes; ++i) {
      int32_t v = g->part[i];
      if (g->level[v] != hops) continue;
      g->height[v] = hops;
      for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
        int32_t w = g->col[e];
        if (g->partB[w]) continue;
        g->arc_visits++;
        if (g->rres[e] > 0 && g->level[w] == -1) { g->level[w] = hops + 1; flag = 1; }
      }
    }
    ++hops;
    g->bfs_levels++;
  } while (flag);
}

/* part_final (code/dyn_pp_bench.cu:91-152). */
static int pp_final(orc_pr* g, int32_t kc) {
  const int32_t n = g->n;
  int32_t nodes = 0;
  /* find_part (code/dyn_pp.h:330-356) */
  for (int32_t v = 0; v < n; ++v) {
    g->partfinal[v] = 0;
    if (g->level[v] == -1) {
      g->part[nodes++] = v;
      g->partB[v] = 0;
      g->partfinal[v] = 1;
    }
    g->level[v] = -1;
    g->height[v] = n;
  }
  if (nodes == 0) return 0;
  pp_final_relabel(g, nodes);
  int32_t active = 0;
  for (int32_t i = 0; i < nodes; ++i) { /* create_final_wl (code/dyn_pp.h:619-642) */
    int32_t v = g->part[i];
    if (g->excess[v] > 0 && g->height[v] < n) g->wl[active++] = v;
  }
  while (active > 0) {
    if (g->rounds >= g->max_rounds) return -6;
    push_pass(g, g->wl, active, kc); /* push_kernel (code/dyn_pp.h:145-185) */
    for (int32_t i = 0; i < active; ++i) repair_vertex(g, g->wl[i], g->partfinal, 1);
    pp_final_relabel(g, nodes);
    active = 0;
    for (int32_t i = 0; i < nodes; ++i) {
      int32_t v = g->part[i];
      if (g->excess[v] > 0 && g->height[v] < n) g->wl[active++] = v;
    }
    g->rounds++;
  }
  /* find_final_cut (code/dyn_pp.h:452-460) */
  for (int32_t v = 0; v < n; ++v)
    if (!g->partB[v] && g->level[v] != -1) g->partB[v] = 1;
  return 0;
}

/* pp incremental_kernel_20 (code/dyn_pp.h:405-436): clamp, saturate the A->B
 * direction of every arc crossing the stored cut, recompute excess. */
static void pp_clamp_cut_recompute(orc_pr* g) {
  for (int32_t v = 0; v < g->n; ++v) {
    uint8_t self = g->partB[v];
    int64_t e1 = 0;
    for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
      int32_t nbr = g->col[e];
      if (g->res[e] < 0) { g->rres[e] += g->res[e]; g->res[e] = 0; }
      if (g->rres[e] < 0) { g->res[e] += g->rres[e]; g->rres[e] = 0; }
      if (self != g->partB[nbr]) {
        if (self) { int32_t d = g->rres[e]; g->rres[e] -= d; g->res[e] += d; }
        else { int32_t d = g->res[e]; g->res[e] -= d; g->rres[e] += d; }
      }
      e1 += (int64_t)g->res[e] - g->cap[e];
    }
    g->excess[v] = e1;
  }
}

/* ============================================================= public API */

static void reset_counters(orc_pr* g) {
  g->rounds = g->pushes = g->relabels = g->repairs = g->bfs_levels = g->arc_visits = 0;
}

/* Static solve from zero flow with the given variant (0 data, 1 topo, 2 pp). */
int orc_pr_static(orc_pr* g, int32_t variant, int32_t kc) {
  reset_counters(g);
  k_init(g);
  if (variant == 2) {
    /* code/dyn_pp_bench.cu:176-232 */
    k_saturate_source_sink_pp(g);
    int rc = data_loop(g, kc);
    if (rc) return rc;
    for (int32_t v = 0; v < g->n; ++v) g->partB[v] = (g->level[v] != -1); /* find_cut */
    return 0;
  }
  k_saturate_source(g);
  return variant == 1 ? topo_loop(g, kc) : data_loop(g, kc);
}

/* One batch: OnAdd, then the variant's Incremental. */
int orc_pr_apply(orc_pr* g, int32_t k, const int32_t* src, const int32_t* dst,
                 const int32_t* w, int32_t variant, int32_t kc) {
  reset_counters(g);
  int rc = apply_updates(g, k, src, dst, w);
  if (rc) return rc;
  if (variant == 2) {
    /* code/dyn_pp_bench.cu:242-258 */
    pp_clamp_cut_recompute(g);
    rc = pp_part(g, kc);
    if (rc) return rc;
    return pp_final(g, kc);
  }
  clamp_and_recompute(g);
  k_saturate_source(g);
  return variant == 1 ? topo_loop(g, kc) : data_loop(g, kc);
}

/* Capacities only (update_kernel), for static_recalculate. */
int orc_pr_set_caps(orc_pr* g, int32_t k, const int32_t* src, const int32_t* dst,
                    const int32_t* w) {
  return apply_caps_only(g, k, src, dst, w);
}

/* recalculate_max_flow_kernel_39 (code/dyn_data.h:208-215), without mutating. */
int64_t orc_pr_flow(const orc_pr* g) {
  int64_t f = g->excess[g->t];
  for (int32_t v = 0; v < g->n; ++v)
    if (g->excess[v] < 0 && v != g->s) f += g->excess[v];
  return f;
}

/* Bare excess[t], what the reference's static data/topo paths return
 * (code/dyn_data_bench.cu:470-471). */
int64_t orc_pr_excess_t(const orc_pr* g) { return g->excess[g->t]; }

/* S_min from a converged state: vertices reachable over arcs with res > 0 from
 * s or from any v != t holding positive excess.  At convergence every such
 * excess sits on the source side of the cut and can be returned to s, which
 * makes those vertices reachable from s in the completed flow. */
void orc_pr_min_cut(const orc_pr* g, uint8_t* side) {
  const int32_t n = g->n;
  int32_t* q = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)n);
  int32_t qt = 0, qh = 0;
  memset(side, 0, (size_t)n);
  for (int32_t v = 0; v < n; ++v)
    if (v == g->s || (v != g->t && g->excess[v] > 0)) { side[v] = 1; q[qt++] = v; }
  while (qh < qt) {
    int32_t v = q[qh++];
    for (int32_t e = g->row[v]; e < g->row[v + 1]; ++e) {
      int32_t w = g->col[e];
      if (g->res[e] > 0 && !side[w]) { side[w] = 1; q[qt++] = w; }
    }
  }
  free(q);
}

void orc_pr_read_state(const orc_pr* g, int32_t* res, int32_t* rres, int32_t* cap,
                       int64_t* excess, int32_t* height) {
  if (res) memcpy(res, g->res, sizeof(int32_t) * (size_t)g->m);
  if (rres) memcpy(rres, g->rres, sizeof(int32_t) * (size_t)g->m);
  if (cap) memcpy(cap, g->cap, sizeof(int32_t) * (size_t)g->m);
  if (excess) memcpy(excess, g->excess, sizeof(int64_t) * (size_t)g->n);
  if (height) memcpy(height, g->height, sizeof(int32_t) * (size_t)g->n);
}

void orc_pr_counters(const orc_pr* g, int64_t* out6) {
  out6[0] = g->rounds; out6[1] = g->pushes; out6[2] = g->relabels;
  out6[3] = g->repairs; out6[4] = g->bfs_levels; out6[5] = g->arc_visits;
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
