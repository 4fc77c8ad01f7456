"""Pins the oracle before anything is checked against it.

Dinic: against scipy's maximum_flow and networkx (independent implementations),
brute-force enumeration of all s-t cuts on tiny graphs, and closed forms
(path = min capacity, disjoint paths = sum, disconnected = 0, grid corner to
corner with uniform capacity c = 2c, unit bipartite = matching size).
S_min: it is a minimum cut and it is the smallest one (contained in every
minimum cut found by enumeration).
Push-relabel restatement (all three variants, static and after every update
batch): equal to Dinic in value and S_min, and its residual state satisfies
capacity and net-inflow invariants.  A planted mistake (one capacity changed,
a broken mirror) is caught by these checks."""
import itertools
import os

import numpy as np
import pytest

import dyn_maxflow_amd as dmf
import oracle
from util import caps_after, rmat_case, small_random_graph

VARIANTS = ["data", "topo", "pp"]


def scipy_value(g, s, t, cap=None):
    import scipy.sparse as sp
    from scipy.sparse.csgraph import maximum_flow
    cap = g.cap if cap is None else cap
    M = sp.csr_matrix((np.asarray(cap, np.int32), (g.tails(), g.col)), shape=(g.n, g.n))
    M.sum_duplicates()
    return int(maximum_flow(M, s, t, method="dinic").flow_value)


def nx_value(g, s, t):
    import networkx as nx
    G = nx.DiGraph()
    G.add_nodes_from(range(g.n))
    for a, b, c in zip(g.tails().tolist(), g.col.tolist(), g.cap.tolist()):
        if G.has_edge(a, b):
            G[a][b]["capacity"] += c
        else:
            G.add_edge(a, b, capacity=c)
    return int(nx.maximum_flow_value(G, s, t))


def cut_capacity(g, side):
    tails = g.tails()
    mask = (side[tails] == 1) & (side[g.col] == 0)
    return int(g.cap[mask].sum())


def test_dinic_vs_scipy_and_networkx():
    rng = np.random.default_rng(1)
    for trial in range(25):
        n = int(rng.integers(5, 60))
        g, _ = small_random_graph(rng, n, int(rng.integers(n, 6 * n)))
        s, t = 0, n - 1
        f, side = oracle.dinic(g, s, t)
        assert f == scipy_value(g, s, t) == nx_value(g, s, t)
        assert cut_capacity(g, side) == f
        assert side[s] == 1 and side[t] == 0


def test_dinic_vs_bruteforce_cut_enumeration_and_smin_minimality():
    rng = np.random.default_rng(2)
    for trial in range(20):
        n = int(rng.integers(4, 11))
        g, _ = small_random_graph(rng, n, int(rng.integers(n, 4 * n)), cap_hi=6)
        s, t = 0, n - 1
        f, side = oracle.dinic(g, s, t)
        others = [v for v in range(n) if v not in (s, t)]
        best, min_sets = None, []
        for bits in itertools.product([0, 1], repeat=len(others)):
            S = np.zeros(n, np.uint8)
            S[s] = 1
            for v, b in zip(others, bits):
                S[v] = b
            c = cut_capacity(g, S)
            if best is None or c < best:
                best, min_sets = c, [S]
            elif c == best:
                min_sets.append(S)
        assert f == best
        inter = np.bitwise_and.reduce(np.array(min_sets), axis=0)
        assert np.array_equal(side, inter), "S_min must be the intersection of all min cuts"


def test_dinic_closed_forms():
    # path 0->1->2->3 with capacities 7, 3, 9: value = min
    g = dmf.build_csr([0, 1, 2], [1, 2, 3], [7, 3, 9])
    assert oracle.dinic(g, 0, 3)[0] == 3
    # three disjoint paths: value = sum of their bottlenecks
    g = dmf.build_csr([0, 1, 0, 2, 0, 3], [1, 4, 2, 4, 3, 4], [5, 2, 4, 8, 1, 1])
    assert oracle.dinic(g, 0, 4)[0] == 2 + 4 + 1
    # t unreachable: 0, and S_min is the component of s
    g = dmf.build_csr([0, 1, 3], [1, 2, 2], [4, 4, 4])
    f, side = oracle.dinic(g, 0, 3)
    assert f == 0 and side.tolist() == [1, 1, 1, 0]
    # grid corner to corner, every capacity c: two arcs leave the corner
    for c in (1, 5):
        u, v, w = dmf.gen_grid(6, 6, 1, c, c)
        g = dmf.build_csr(u, v, w)
        assert oracle.dinic(g, 0, 35)[0] == 2 * c
    # unit bipartite graph: value = maximum matching size
    import networkx as nx
    rng = np.random.default_rng(3)
    L, R = 12, 10
    edges = {(int(a), int(b)) for a, b in zip(rng.integers(0, L, 40), rng.integers(0, R, 40))}
    B = nx.Graph()
    B.add_nodes_from([("l", i) for i in range(L)] + [("r", j) for j in range(R)])
    B.add_edges_from((("l", a), ("r", b)) for a, b in edges)
    msize = len(nx.bipartite.maximum_matching(B, top_nodes=[("l", i) for i in range(L)])) // 2
    s, t = L + R, L + R + 1
    us = [s] * L + [a for a, _ in edges] + [L + j for j in range(R)]
    vs = list(range(L)) + [L + b for _, b in edges] + [t] * R
    g = dmf.build_csr(us, vs, [1] * len(us))
    assert oracle.dinic(g, s, t)[0] == msize


def test_planted_mistakes_are_caught():
    g, s, t, _ = rmat_case(8, 8, 5)
    f, side = oracle.dinic(g, s, t)
    # a dropped capacity term on one arc of the min cut changes the value
    tails = g.tails()
    cut_arcs = np.flatnonzero((side[tails] == 1) & (side[g.col] == 0) & (g.cap > 0))
    cap = g.cap.copy()
    cap[cut_arcs[0]] -= 1
    assert oracle.dinic(g, s, t, cap=cap)[0] == f - 1
    # a broken residual mirror or excess is reported by check_state
    pr = oracle.PushRelabel(g, s, t)
    pr.static("data", 4)
    res, rres, capn, ex, h = pr.state()
    assert oracle.check_state(g, res, capn, ex) == []
    res2 = res.copy()
    res2[0] += 1
    assert oracle.check_state(g, res2, capn, ex)
    ex2 = ex.copy()
    ex2[t] += 1
    assert oracle.check_state(g, res, capn, ex2)


@pytest.mark.parametrize("variant", VARIANTS)
def test_push_relabel_static_matches_dinic(variant):
    for scale, seed in ((6, 1), (8, 2), (10, 3), (11, 4)):
        g, s, t, _ = rmat_case(scale, 16, seed)
        f, side = oracle.dinic(g, s, t)
        pr = oracle.PushRelabel(g, s, t)
        assert pr.static(variant, max(1, g.m // g.n)) == f
        assert np.array_equal(pr.min_cut(), side)
        res, _, cap, ex, _ = pr.state()
        assert oracle.check_state(g, res, cap, ex) == []


@pytest.mark.parametrize("variant", VARIANTS)
def test_push_relabel_batches_match_dinic(variant):
    for scale, seed in ((8, 1), (10, 2), (12, 3)):
        g, s, t, (u, v, w) = rmat_case(scale, 16, seed)
        kc = max(1, g.m // g.n)
        pr = oracle.PushRelabel(g, s, t)
        pr.static(variant, kc)
        cur = w.copy()
        applied = []
        for b in range(1, 6):
            src, dst, nw = dmf.gen_updates(u, v, cur, len(u) * b // 100 + 1, seed * 100 + b)
            applied.append((src, dst, nw))
            f = pr.apply(src, dst, nw, variant, kc)
            cap = caps_after(g, applied)
            fd, side = oracle.dinic(g, s, t, cap=cap)
            assert f == fd, (scale, seed, b)
            assert np.array_equal(pr.min_cut(), side)
            res, _, capn, ex, _ = pr.state()
            assert np.array_equal(capn, cap)
            assert oracle.check_state(g, res, capn, ex) == []


def test_push_relabel_golden_fixtures(golden_dir, expected):
    g = dmf.read_graph_bin(os.path.join(golden_dir, "johnson.bin"))
    J = expected["johnson"]
    for variant in VARIANTS:
        pr = oracle.PushRelabel(g, J["s"], J["t"])
        assert pr.static(variant, max(1, g.m // g.n)) == J["static_flow"]
        assert np.flatnonzero(pr.min_cut()).tolist() == J["static_smin"]
        for b in J["batches"]:  # independent files: each applied to the original graph
            src, dst, w = dmf.read_updates_bin(os.path.join(golden_dir, b["file"]))
            pr2 = oracle.PushRelabel(g, J["s"], J["t"])
            pr2.static(variant, max(1, g.m // g.n))
            assert pr2.apply(src, dst, w, variant, max(1, g.m // g.n)) == b["flow"]
            assert np.flatnonzero(pr2.min_cut()).tolist() == b["smin"]
    rg = dmf.read_graph_bin(os.path.join(golden_dir, "rmat8.bin"))
    R = expected["rmat8"]
    assert oracle.dinic(rg, R["s"], R["t"])[0] == R["flow"]


def test_openmp_baseline_matches():
    g, s, t, (u, v, w) = rmat_case(12, 16, 7)
    kc = max(1, g.m // g.n)
    pr = oracle.PushRelabel(g, s, t, threads=4)
    f = pr.static("data", kc)
    assert f == oracle.dinic(g, s, t)[0]
    cur = w.copy()
    src, dst, nw = dmf.gen_updates(u, v, cur, len(u) // 20, 9)
    assert pr.apply(src, dst, nw, "data", kc) == oracle.dinic(g, s, t, cap=caps_after(g, [(src, dst, nw)]))[0]


def test_edge_cases_of_the_restatement():
    # s and t disconnected: value 0 for every variant
    g = dmf.build_csr([0, 1, 2, 3], [1, 0, 3, 2], [5, 5, 5, 5])
    for variant in VARIANTS:
        pr = oracle.PushRelabel(g, 0, 3)
        assert pr.static(variant, 2) == 0
    # empty batch leaves the value unchanged; self-loop updates are skipped
    g, s, t, (u, v, w) = rmat_case(8, 8, 3)
    pr = oracle.PushRelabel(g, s, t)
    f0 = pr.static("data", 4)
    assert pr.apply([], [], [], "data", 4) == f0
    assert pr.apply([s], [s], [50], "data", 4) == f0
    # duplicate records for one arc: the later one wins
    a, b = int(u[0]), int(v[0])
    pr.apply([a, a], [b, b], [1, 77], "data", 4)
    cap = pr.state()[2]
    e = [i for i in range(g.row[a], g.row[a + 1]) if g.col[i] == b][-1]
    assert cap[e] == 77


def _path_flow_state(g, flows):
    """Residuals of a flow given per (tail, head) input edge: res[e] = cap[e] - f(e)
    on the edge's arc and cap[rev] + f(e) on its reverse."""
    res = g.cap.astype(np.int64).copy()
    tails = g.tails()
    for (a, b), f in flows.items():
        e = int(np.flatnonzero((tails == a) & (g.col == b))[0])
        res[e] -= f
        res[g.rev[e]] += f
    ex = np.zeros(g.n, np.int64)
    for (a, b), f in flows.items():
        ex[b] += f
        ex[a] -= f
    return res.astype(np.int32), ex


def test_check_flow_and_cut_capacity_pinned():
    """check_flow accepts a hand-built maximum flow and catches a planted
    conservation error; cut_capacity equals hand-computed cut sums and, for
    S_min, Dinic's value (max-flow = min-cut) on random graphs."""
    # s=0 -> 1 -> 3=t and 0 -> 2 -> 3, capacities 5/3 and 4/6: max flow 3 + 4 = 7
    g = dmf.build_csr([0, 1, 0, 2], [1, 3, 2, 3], [5, 3, 4, 6])
    res, ex = _path_flow_state(g, {(0, 1): 3, (1, 3): 3, (0, 2): 4, (2, 3): 4})
    assert oracle.check_flow(g, res, g.cap, ex, [0], [3], 7) == []
    assert oracle.dinic(g, 0, 3)[0] == 7
    # planted: one unit more into vertex 1 than leaves it
    res2, ex2 = _path_flow_state(g, {(0, 1): 4, (1, 3): 3, (0, 2): 4, (2, 3): 4})
    bad = oracle.check_flow(g, res2, g.cap, ex2, [0], [3], 7)
    assert any("conservation" in b for b in bad) and any("source excess" in b for b in bad)
    # planted: the right state reported with a wrong value
    assert any("sink excess" in b for b in oracle.check_flow(g, res, g.cap, ex, [0], [3], 8))
    # cut sums by hand: {0} -> 5 + 4, {0,1} -> 3 + 4, {0,1,2} -> 3 + 6
    for S, c in (({0}, 9), ({0, 1}, 7), ({0, 1, 2}, 9)):
        side = np.zeros(g.n, np.uint8)
        side[list(S)] = 1
        assert oracle.cut_capacity(g, g.cap, side) == c == cut_capacity(g, side)
    rng = np.random.default_rng(12)
    for _ in range(20):
        n = int(rng.integers(5, 50))
        gr, _ = small_random_graph(rng, n, int(rng.integers(n, 5 * n)))
        f, side = oracle.dinic(gr, 0, n - 1)
        assert oracle.cut_capacity(gr, gr.cap, side) == f


def test_small_config_on_the_cpu_path():
    """The small plain configuration run on the CPU push-relabel path: RMAT-14,
    EF16, seed 1; static solve, then 10 batches of 1% of the input edges, flow
    value and S_min equal to Dinic after every batch (data variant, kc = m/n)."""
    g, s, t, (u, v, w) = rmat_case(14, 16, 1)
    kc = max(1, g.m // g.n)
    pr = oracle.PushRelabel(g, s, t, threads=1)
    try:
        f = pr.static("data", kc)
        fd, side = oracle.dinic(g, s, t)
        assert f == fd and np.array_equal(pr.min_cut(), side)
        cur = w.copy()
        applied = []
        for b in range(10):
            src, dst, nw = dmf.gen_updates(u, v, cur, len(u) // 100, 100 + b)
            applied.append((src, dst, nw))
            f = pr.apply(src, dst, nw, "data", kc)
            fd, side = oracle.dinic(g, s, t, cap=caps_after(g, applied))
            assert f == fd, (b, f, fd)
            assert np.array_equal(pr.min_cut(), side), b
    finally:
        pr.close()
