"""GPU parity: the HIP engine (through the C-ABI) against the oracle.

Bar: bit-exact max-flow value and S_min (vertices reachable from s in the
residual graph of a maximum flow) after the static solve and after every update
batch, plus the residual-state invariants (0 <= res, pair consistency,
excess = net inflow).  Runs every variant (data, topo, pp) under the three drivers
(persistent one-launch solve, stepped per-phase launches, stepped with the BFS
levels replayed from captured graphs)."""
import os
import sys

import numpy as np
import pytest

import oracle
from util import caps_after, rmat_case

pytestmark = pytest.mark.gpu

VARIANTS = ["data", "topo", "pp"]
MODES = [0, 1, 2, 3]  # 0 persistent, 1 stepped, 2 stepped with graph-replayed BFS levels, 3 round graphs


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_engine(eng, g, cap, s, t, flow, where):
    fd, side = oracle.dinic(g, s, t, cap=cap)
    assert flow == fd, f"{where}: flow {flow} != dinic {fd}"
    smin = eng.min_cut()
    assert np.array_equal(smin, side), f"{where}: S_min differs on {(smin != side).sum()} vertices"
    res, capd, ex, _ = eng.state()
    assert np.array_equal(capd, cap), f"{where}: capacities differ"
    bad = oracle.check_state(g, res, capd, ex)
    assert not bad, f"{where}: {bad}"


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("variant", VARIANTS)
def test_golden_johnson(gpu, golden_dir, expected, variant, mode):
    dmf = gpu
    g = dmf.read_graph_bin(os.path.join(golden_dir, "johnson.bin"))
    J = expected["johnson"]
    s, t = J["s"], J["t"]
    p = dmf.default_params(g, variant=variant, mode=mode)
    say(f"[johnson] variant={variant} mode={mode} static")
    with dmf.Engine(g, s, t) as eng:
        assert eng.static_solve(p) == J["static_flow"]
        assert np.flatnonzero(eng.min_cut()).tolist() == J["static_smin"]
    for b in J["batches"]:  # independent files on the original graph (code/README.md:24)
        say(f"[johnson] variant={variant} mode={mode} batch {b['file']}")
        src, dst, w = dmf.read_updates_bin(os.path.join(golden_dir, b["file"]))
        with dmf.Engine(g, s, t) as eng:
            eng.static_solve(p)
            assert eng.apply_batch(src, dst, w, p) == b["flow"]
            assert np.flatnonzero(eng.min_cut()).tolist() == b["smin"]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("variant", VARIANTS)
def test_rmat_successive_batches(gpu, variant, mode):
    dmf = gpu
    for scale, seed in ((9, 1), (12, 2)):
        g, s, t, (u, v, w) = rmat_case(scale, 16, seed)
        p = dmf.default_params(g, variant=variant, mode=mode)
        say(f"[rmat] scale={scale} seed={seed} variant={variant} mode={mode} static")
        with dmf.Engine(g, s, t) as eng:
            f = eng.static_solve(p)
            check_engine(eng, g, g.cap, s, t, f, "static")
            cur = w.copy()
            applied = []
            for b in range(1, 6):
                src, dst, nw = dmf.gen_updates(u, v, cur, len(u) * b // 100 + 1, seed * 1000 + b)
                applied.append((src, dst, nw))
                say(f"[rmat] scale={scale} variant={variant} mode={mode} batch {b} k={len(src)}")
                f = eng.apply_batch(src, dst, nw, p)
                check_engine(eng, g, caps_after(g, applied), s, t, f, f"batch {b}")


@pytest.mark.parametrize("variant", VARIANTS)
def test_rmat14_config_ten_one_percent_batches(gpu, variant):
    """Small plain configuration: RMAT-14, EF16, seed 1, 10 batches of 1%."""
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(14, 16, 1)
    p = dmf.default_params(g, variant=variant)
    say(f"[rmat14] variant={variant}")
    with dmf.Engine(g, s, t) as eng:
        f = eng.static_solve(p)
        check_engine(eng, g, g.cap, s, t, f, "static")
        cur = w.copy()
        applied = []
        for b in range(10):
            src, dst, nw = dmf.gen_updates(u, v, cur, len(u) // 100, 100 + b)
            applied.append((src, dst, nw))
            f = eng.apply_batch(src, dst, nw, p)
            check_engine(eng, g, caps_after(g, applied), s, t, f, f"batch {b}")


@pytest.mark.parametrize("variant", VARIANTS)
def test_reference_entry_points_agree(gpu, golden_dir, expected, variant):
    """dynamic == static recompute (the reference's PANIC check,
    code/dyn_data_bench.cu:535-539) and both equal Dinic."""
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(11, 16, 4)
    kc = max(1, g.m // g.n)
    cur = w.copy()
    src, dst, nw = dmf.gen_updates(u, v, cur, len(u) // 20, 77, apply=False)
    say(f"[entry points] variant={variant}")
    dyn, ms_s, ms_b = dmf.recalculate_max_flow(g, src, dst, nw, s, t, kc, variant)
    stat, ms = dmf.static_recalculate(g, src, dst, nw, s, t, kc, variant)
    fd, _ = oracle.dinic(g, s, t, cap=caps_after(g, [(src, dst, nw)]))
    assert dyn == stat == fd
    assert ms_s > 0 and ms_b > 0 and ms > 0


@pytest.mark.parametrize("scatter", ["0", "1"])
@pytest.mark.parametrize("mode", MODES)
def test_batch_edge_cases(gpu, mode, scatter, monkeypatch):
    """Update records through the arc index (keyed by the caller's ids) with
    the engine numbering scattered and not."""
    dmf = gpu
    monkeypatch.setenv("DMF_VERTEX_SCATTER", scatter)
    g, s, t, (u, v, w) = rmat_case(10, 8, 3)
    p = dmf.default_params(g, mode=mode)
    say(f"[edge cases] mode={mode} scatter={scatter}")
    with dmf.Engine(g, s, t) as eng:
        f0 = eng.static_solve(p)
        # empty batch
        assert eng.apply_batch([], [], [], p) == f0
        # self-loop records are skipped
        assert eng.apply_batch([s, 3], [s, 3], [7, 9], p) == f0
        # an absent arc fails the whole batch and changes nothing
        tails = g.tails()
        present = set(zip(tails.tolist(), g.col.tolist()))
        a = next(x for x in range(g.n) if (s, x) not in present and x != s)
        rc = eng.apply_batch_status([int(u[0]), s], [int(v[0]), a], [1, 5], p)
        assert rc == dmf.ENOARC
        res, cap, ex, _ = eng.state()
        assert np.array_equal(cap, g.cap)
        # out-of-range vertex id
        rc = eng.apply_batch_status([g.n + 5], [0], [1], p)
        assert rc == dmf.EINVAL
        assert eng.flow() == f0
        # a negative capacity is refused like a malformed CSR, state untouched
        rc = eng.apply_batch_status([int(u[0]), int(u[1])], [int(v[0]), int(v[1])], [5, -3], p)
        assert rc == dmf.EINVAL
        assert eng.flow() == f0
        _, capn, _, _ = eng.state()
        assert np.array_equal(capn, g.cap)
        # out-of-range parameters are refused before any device work
        for field, val in (("mode", 7), ("mode", -1), ("variant", 9), ("scheduler", 4)):
            bad = dmf.default_params(g, mode=mode)
            setattr(bad, field, val)
            assert eng.apply_batch_status([int(u[0])], [int(v[0])], [1], bad) == dmf.EINVAL, field
        assert eng.flow() == f0
        res2, cap2, _, _ = eng.state()
        assert np.array_equal(cap2, g.cap) and np.array_equal(res2, res)
        # duplicate records for one arc: the last wins
        a, b = int(u[1]), int(v[1])
        f = eng.apply_batch([a, a, a], [b, b, b], [100, 1, 42], p)
        cap_exp = caps_after(g, [([a], [b], [42])])
        check_engine(eng, g, cap_exp, s, t, f, "duplicates")
        # the same arc again with an increase on every arc out of s
        so = g.col[g.row[s]:g.row[s + 1]]
        f = eng.apply_batch(np.full(len(so), s), so, np.full(len(so), 100), p)
        cap_exp = caps_after(g, [([a], [b], [42]), (np.full(len(so), s), so, np.full(len(so), 100))])
        check_engine(eng, g, cap_exp, s, t, f, "source increases")
        # every arc into t decreased to 1
        ti = np.flatnonzero(g.col == t)
        f = eng.apply_batch(tails[ti], np.full(len(ti), t), np.ones(len(ti), np.int32), p)
        cap_exp2 = cap_exp.copy()
        cap_exp2[ti] = 1
        check_engine(eng, g, cap_exp2, s, t, f, "sink decreases")


def test_disconnected_source_and_sink(gpu):
    dmf = gpu
    g = dmf.build_csr([0, 1, 2, 3, 4], [1, 0, 3, 2, 5], [5, 5, 5, 5, 1])
    for variant in VARIANTS:
        for mode in MODES:
            with dmf.Engine(g, 0, 3) as eng:
                p = dmf.default_params(g, variant=variant, mode=mode)
                assert eng.static_solve(p) == 0
                assert eng.min_cut().tolist() == [1, 1, 0, 0, 0, 0]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("interval", [2, 5])
def test_relabel_interval_keeps_results(gpu, interval, variant, mode):
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(12, 16, 6)
    p = dmf.default_params(g, variant=variant, relabel_interval=interval, mode=mode)
    say(f"[interval] {interval} variant={variant} mode={mode}")
    with dmf.Engine(g, s, t) as eng:
        f = eng.static_solve(p)
        check_engine(eng, g, g.cap, s, t, f, "static")
        cur = w.copy()
        applied = []
        for b in range(3):
            src, dst, nw = dmf.gen_updates(u, v, cur, len(u) // 30, 500 + b)
            applied.append((src, dst, nw))
            f = eng.apply_batch(src, dst, nw, p)
            check_engine(eng, g, caps_after(g, applied), s, t, f, f"batch {b}")


def test_counters_are_populated(gpu):
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(12, 16, 8)
    p = dmf.default_params(g)
    with dmf.Engine(g, s, t) as eng:
        eng.static_solve(p)
        c = eng.counters()
        assert c["arc_visits"] > g.m // 2
        assert c["global_relabels"] >= 1 and c["bfs_levels"] >= 1
        assert c["device_ms"] > 0
        cur = w.copy()
        src, dst, nw = dmf.gen_updates(u, v, cur, len(u) // 50, 3)
        eng.apply_batch(src, dst, nw, p)
        c = eng.counters()
        assert c["updates"] == len(src)
        assert eng.device_bytes() > 20 * g.m


def test_learned_scheduler_keeps_results(gpu, tmp_path):
    """Whatever the scheduler decides (variant per batch, relabel timing), the
    flow value and S_min must stay exact; its trace must be written."""
    dmf = gpu
    import json
    g, s, t, (u, v, w) = rmat_case(12, 16, 9)
    trace = tmp_path / "trace.jsonl"
    os.environ["DMF_TRACE"] = str(trace)
    try:
        eng = dmf.Engine(g, s, t)
    finally:
        os.environ.pop("DMF_TRACE", None)
    with eng:
        p = dmf.default_params(g, scheduler=dmf.SCHED_LEARNED)
        f = eng.static_solve(p)
        check_engine(eng, g, g.cap, s, t, f, "static")
        cur = w.copy()
        applied = []
        for b in range(4):
            src, dst, nw = dmf.gen_updates(u, v, cur, len(u) * (b + 1) // 50, 900 + b)
            applied.append((src, dst, nw))
            say(f"[learned] batch {b}")
            f = eng.apply_batch(src, dst, nw, p)
            check_engine(eng, g, caps_after(g, applied), s, t, f, f"learned batch {b}")
    recs = [json.loads(l) for l in open(trace)]
    assert any(r["what"] == "round" for r in recs)
    assert sum(r["what"] == "summary" for r in recs) == 5
    assert all(len(r["feat"]) == 8 for r in recs if r["what"] != "summary")


@pytest.mark.parametrize("mode", MODES)
def test_batched_instances_match_solo_solves(gpu, mode):
    """Independent instances solved together (one disjoint union, one launch per
    phase for all of them): every instance's flow and S_min equal its solo
    engine's and Dinic's after every batch."""
    dmf = gpu
    rng = np.random.default_rng(4)
    cases = []
    for i in range(12):
        scale = int(rng.integers(7, 12))
        g, s, t, (u, v, w) = rmat_case(scale, int(rng.integers(8, 33)), 40 + i)
        cases.append((g, s, t, u, v, w))
    p = dmf.default_params(cases[0][0], mode=mode)
    say(f"[batched] mode={mode} instances={len(cases)}")
    solos = [dmf.Engine(c[0], c[1], c[2]) for c in cases]
    try:
        with dmf.BatchEngine([c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases]) as be:
            assert be.instances() == len(cases)
            flows = be.static_solve_all(p)
            for i, (g, s, t, u, v, w) in enumerate(cases):
                assert flows[i] == oracle.dinic(g, s, t)[0] == solos[i].static_solve(p), i
            curs = [c[5].copy() for c in cases]
            applied = [[] for _ in cases]
            for b in range(3):
                ups = []
                for i, (g, s, t, u, v, w) in enumerate(cases):
                    k = max(1, len(u) * (b + 1) // 50)
                    ups.append(dmf.gen_updates(u, v, curs[i], k, 600 + 10 * i + b))
                    applied[i].append(ups[-1])
                say(f"[batched] mode={mode} batch {b}")
                flows = be.apply_all(ups, p)
                assert np.array_equal(flows, be.flows())
                sides = be.min_cut_all()
                for i, (g, s, t, u, v, w) in enumerate(cases):
                    fd, side = oracle.dinic(g, s, t, cap=caps_after(g, applied[i]))
                    f_solo = solos[i].apply_batch(*ups[i], p)
                    assert flows[i] == fd == f_solo, (b, i)
                    assert np.array_equal(sides[i], side), (b, i)
    finally:
        for e in solos:
            e.close()


def test_autotune_keeps_results(gpu):
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(12, 16, 10)
    with dmf.Engine(g, s, t) as eng:
        p = dmf.default_params(g)
        lm, wt = eng.autotune(p)
        assert lm > 0 and wt > 0
        # rebuilding the wide lists per candidate releases the old ones
        assert 0 < eng.device_bytes() <= dmf.engine_footprint(g.n, g.m)
        f = eng.static_solve(p)
        check_engine(eng, g, g.cap, s, t, f, "autotuned static")
        cur = w.copy()
        src, dst, nw = dmf.gen_updates(u, v, cur, len(u) // 25, 77)
        f = eng.apply_batch(src, dst, nw, p)
        check_engine(eng, g, caps_after(g, [(src, dst, nw)]), s, t, f, "autotuned batch")


@pytest.mark.parametrize("mode", MODES)
def test_variant_switching_between_batches(gpu, mode):
    """Push-pull batches continue from a push-pull static state; from any other
    static state a push-pull request runs data-driven.  Every flow and S_min
    equals Dinic's through mixed sequences of variants."""
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(10, 16, 1)
    cur = w.copy()
    bl = [dmf.gen_updates(u, v, cur, len(u) * p // 100, 50 + i) for i, p in enumerate((2, 5, 3, 4, 2))]
    for static_variant, seq in (("data", ["data", "pp", "topo", "pp", "data"]),
                                ("topo", ["pp", "pp", "data", "topo", "pp"]),
                                ("pp", ["data", "pp", "topo", "pp", "data"]),
                                ("pp", ["pp", "topo", "topo", "pp", "pp"])):
        say(f"[switch] mode={mode} static={static_variant} seq={seq}")
        with dmf.Engine(g, s, t) as eng:
            eng.static_solve(dmf.default_params(g, variant=static_variant, mode=mode))
            for i, var in enumerate(seq):
                f = eng.apply_batch(*bl[i], dmf.default_params(g, variant=var, mode=mode))
                ran = eng.counters()["variant"]
                want = dmf.VARIANTS[var]
                if var == "pp" and static_variant != "pp":
                    want = dmf.VARIANT_DATA
                assert ran == want, (static_variant, seq, i, ran)
                check_engine(eng, g, caps_after(g, bl[:i + 1]), s, t, f, f"{static_variant} {seq[:i + 1]}")
        # the learned scheduler may pick any variant from a push-pull state
        with dmf.Engine(g, s, t) as eng:
            eng.static_solve(dmf.default_params(g, variant="pp", mode=mode))
            for i in range(len(bl)):
                f = eng.apply_batch(*bl[i], dmf.default_params(g, mode=1, scheduler=dmf.SCHED_LEARNED))
                check_engine(eng, g, caps_after(g, bl[:i + 1]), s, t, f, f"learned batch {i}")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("interval", [1, 3])
def test_lazy_start_keeps_results(gpu, mode, interval):
    """Data-driven batches that start on the previous solve's labels (no
    relabel before the first round) still end bit-exact."""
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(12, 16, 9)
    p = dmf.default_params(g, relabel_interval=interval, mode=mode, lazy_start=1)
    say(f"[lazy] mode={mode} interval={interval}")
    with dmf.Engine(g, s, t) as eng:
        f = eng.static_solve(dmf.default_params(g, mode=mode))
        check_engine(eng, g, g.cap, s, t, f, "static")
        cur = w.copy()
        applied = []
        for b, pct in enumerate((1, 4, 2, 6)):
            applied.append(dmf.gen_updates(u, v, cur, len(u) * pct // 100, 800 + b))
            f = eng.apply_batch(*applied[-1], p)
            check_engine(eng, g, caps_after(g, applied), s, t, f, f"lazy batch {b}")


def degree_boundary_graph(dmf, seed=5):
    """s -> hubs whose adjacency degree (CSR row, reverse arcs included) is
    exactly 4/5, 16/17, 64/65, 256/257 and 512/513: the default lane threshold
    of push and repair, the topology-driven lane threshold (the static wave
    list), the BFS wide threshold, one 256-arc chunk vs two, two chunks vs
    three.  Hubs feed a pool of leaves that drain into t and into each other,
    so every hub is active in the push phases and on the BFS frontier."""
    rng = np.random.default_rng(seed)
    degs = [4, 5, 16, 17, 64, 65, 256, 257, 512, 513]
    s, t = 0, 1
    hubs = list(range(2, 2 + len(degs)))
    nleaf = 700
    leaf0 = 2 + len(degs)
    u, v, w = [], [], []
    for h, d in zip(hubs, degs):
        u.append(s); v.append(h); w.append(int(rng.integers(2000, 6000)))
        for x in rng.choice(nleaf, size=d - 1, replace=False):  # d-1 out-arcs + the arc from s
            u.append(h); v.append(leaf0 + int(x)); w.append(int(rng.integers(1, 40)))
    for x in range(nleaf):
        u.append(leaf0 + x); v.append(t); w.append(int(rng.integers(1, 60)))
    pairs = set()
    while len(pairs) < 1500:
        a, b = (int(z) for z in rng.integers(0, nleaf, size=2))
        if a != b and (b, a) not in pairs:
            pairs.add((a, b))
    for a, b in sorted(pairs):
        u.append(leaf0 + a); v.append(leaf0 + b); w.append(int(rng.integers(1, 30)))
    u, v, w = (np.array(x, np.int32) for x in (u, v, w))
    g = dmf.build_csr(u, v, w)
    deg = np.diff(g.row)
    assert [int(deg[h]) for h in hubs] == degs
    return g, s, t, (u, v, w)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("variant", VARIANTS)
def test_degree_boundaries(gpu, variant, mode):
    dmf = gpu
    g, s, t, (u, v, w) = degree_boundary_graph(dmf)
    p = dmf.default_params(g, variant=variant, mode=mode)
    say(f"[degrees] variant={variant} mode={mode} static")
    with dmf.Engine(g, s, t) as eng:
        f = eng.static_solve(p)
        check_engine(eng, g, g.cap, s, t, f, "static")
        cur = w.copy()
        applied = []
        for b in range(3):
            src, dst, nw = dmf.gen_updates(u, v, cur, len(u) // 10, 500 + b)
            applied.append((src, dst, nw))
            say(f"[degrees] variant={variant} mode={mode} batch {b} k={len(src)}")
            f = eng.apply_batch(src, dst, nw, p)
            check_engine(eng, g, caps_after(g, applied), s, t, f, f"batch {b}")


@pytest.mark.parametrize("mode", MODES)
def test_vertex_scatter_is_invisible(gpu, mode, monkeypatch):
    """The engine may renumber the vertices of a graph with hubs internally.
    Flows, S_min, capacities and the residual state read back must be in the
    caller's numbering and equal, scattered or not, through batches of every
    variant (including arcs into and out of s and t)."""
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(11, 16, 4)
    cur = w.copy()
    bl = [dmf.gen_updates(u, v, cur, len(u) * pc // 100, 90 + i) for i, pc in enumerate((3, 6, 2))]
    so = g.col[g.row[s]:g.row[s + 1]]
    bl.append((np.full(len(so), s), so, np.full(len(so), 77)))
    seqs = {}
    for scatter in ("0", "1"):
        monkeypatch.setenv("DMF_VERTEX_SCATTER", scatter)
        say(f"[scatter] mode={mode} scatter={scatter}")
        flows, cuts = [], []
        with dmf.Engine(g, s, t) as eng:
            f = eng.static_solve(dmf.default_params(g, variant="pp", mode=mode))
            check_engine(eng, g, g.cap, s, t, f, f"scatter={scatter} static")
            flows.append(f)
            cuts.append(eng.min_cut())
            done = []
            for i, var in enumerate(("data", "pp", "topo", "data")):
                f = eng.apply_batch(*bl[i], dmf.default_params(g, variant=var, mode=mode))
                done.append(bl[i])
                check_engine(eng, g, caps_after(g, done), s, t, f, f"scatter={scatter} batch {i} {var}")
                flows.append(f)
                cuts.append(eng.min_cut())
        seqs[scatter] = (flows, cuts)
    assert seqs["0"][0] == seqs["1"][0]
    for a, b in zip(seqs["0"][1], seqs["1"][1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("variant", VARIANTS)
def test_completion_after_every_batch(gpu, variant, mode):
    """The maximum preflow of every solve is completed into a maximum flow on
    the device: same value, conservation at every non-terminal vertex, same
    S_min; the next batch continues from the completed state and is exact."""
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(12, 16, 3)
    p = dmf.default_params(g, variant=variant, mode=mode)
    say(f"[complete] variant={variant} mode={mode}")
    with dmf.Engine(g, s, t) as eng:
        f = eng.static_solve(p)
        applied = []
        cur = w.copy()
        for b in range(5):
            cap = caps_after(g, applied)
            check_engine(eng, g, cap, s, t, f, f"before completion {b}")
            side = eng.min_cut()
            assert eng.complete_flow() == f
            res, capd, ex, _ = eng.state()
            bad = oracle.check_flow(g, res, capd, ex, [s], [t], f)
            assert not bad, f"completed {b}: {bad}"
            assert np.array_equal(eng.min_cut(), side)
            assert eng.complete_flow() == f  # idempotent
            applied.append(dmf.gen_updates(u, v, cur, len(u) * (b + 2) // 100, 300 + b))
            f = eng.apply_batch(*applied[-1], p)
        check_engine(eng, g, caps_after(g, applied), s, t, f, "last batch")


def test_lazy_and_learned_batches_without_min_cut(gpu):
    """Lazily started and learned-scheduler batches, with no min_cut (or any
    other read) between them, end exact, as do the same batches with S_min
    read between them.  (Round and relabel counts are not compared: the
    lock-free pushes make them vary from run to run; the flows do not.)"""
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(12, 16, 11)
    cur = w.copy()
    bl = [dmf.gen_updates(u, v, cur, len(u) * pc // 100, 700 + i) for i, pc in enumerate((2, 5, 1, 7, 3))]
    runs = {}
    for sched in ("lazy", "learned"):
        for between in (False, True):
            p = (dmf.default_params(g, mode=1, lazy_start=1) if sched == "lazy"
                 else dmf.default_params(g, mode=1, scheduler=dmf.SCHED_LEARNED))
            say(f"[no-min-cut] {sched} min_cut between batches: {between}")
            ctr = []
            with dmf.Engine(g, s, t) as eng:
                eng.static_solve(dmf.default_params(g, variant="pp", mode=1))
                for b in bl:
                    f = eng.apply_batch(*b, p)
                    ctr.append(f)
                    if between:
                        eng.min_cut()
                check_engine(eng, g, caps_after(g, bl), s, t, f, f"{sched} last batch")
            runs[(sched, between)] = ctr
        assert runs[(sched, False)] == runs[(sched, True)], sched


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("interval", [1, 3])
def test_early_exit_relabel_keeps_results(gpu, mode, interval, monkeypatch):
    """Relabels that stop once every active vertex is labelled (forced on for
    the fixed schedule here; the learned scheduler uses them by default) end
    exact: convergence is only accepted after a full relabel."""
    dmf = gpu
    monkeypatch.setenv("DMF_EARLY_RELABEL", "2")
    g, s, t, (u, v, w) = rmat_case(12, 16, 13)
    p = dmf.default_params(g, relabel_interval=interval, mode=mode)
    say(f"[early] mode={mode} interval={interval}")
    with dmf.Engine(g, s, t) as eng:
        f = eng.static_solve(dmf.default_params(g, variant="pp", mode=mode))
        check_engine(eng, g, g.cap, s, t, f, "static")
        cur = w.copy()
        applied = []
        for b, pct in enumerate((1, 4, 2, 8)):
            applied.append(dmf.gen_updates(u, v, cur, len(u) * pct // 100, 1300 + b))
            f = eng.apply_batch(*applied[-1], p)
            check_engine(eng, g, caps_after(g, applied), s, t, f, f"early batch {b}")


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("variant", ["data", "pp"])
def test_cascade_subrounds_keep_results(gpu, mode, variant, monkeypatch):
    """Push sub-rounds chained on the device after each data-driven round
    (forced on here for every schedule) end exact, together with the learned
    scheduler and early-exit relabels."""
    dmf = gpu
    monkeypatch.setenv("DMF_CASCADE", "8")
    g, s, t, (u, v, w) = rmat_case(12, 16, 14)
    say(f"[cascade] variant={variant} mode={mode}")
    with dmf.Engine(g, s, t) as eng:
        f = eng.static_solve(dmf.default_params(g, variant=variant, mode=mode))
        check_engine(eng, g, g.cap, s, t, f, "static")
        cur = w.copy()
        applied = []
        for b, pct in enumerate((1, 4, 2, 8, 3)):
            applied.append(dmf.gen_updates(u, v, cur, len(u) * pct // 100, 1400 + b))
            p = (dmf.default_params(g, variant=variant, mode=mode) if b % 2 == 0
                 else dmf.default_params(g, mode=mode, scheduler=dmf.SCHED_LEARNED))
            f = eng.apply_batch(*applied[-1], p)
            check_engine(eng, g, caps_after(g, applied), s, t, f, f"cascade batch {b}")


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("env", [("DMF_BFS_RESTRICT", "1"), ("DMF_BFS_RESTRICT", "64"),
                                 ("DMF_BFS_BOTTOM_UP", "100000")])
def test_bottom_up_relabels_keep_results(gpu, mode, env, monkeypatch):
    """Bottom-up BFS levels.  DMF_BFS_RESTRICT: the relabels from t inside a
    solve try only the last complete labelled set as candidates, at nearly every
    level (1) or only at the largest frontiers (64), under the cascade schedule
    and the learned scheduler.  DMF_BFS_BOTTOM_UP: every relabel from t takes
    its levels bottom-up over all vertices.  Exact after every batch on an
    R-MAT graph and on a grid, including capacity increases that let more
    vertices reach t."""
    dmf = gpu
    if env[0] == "DMF_BFS_BOTTOM_UP":
        monkeypatch.setenv("DMF_BFS_RESTRICT", "0")
    monkeypatch.setenv(*env)
    cases = [rmat_case(12, 16, 16)]
    u, v, w = dmf.gen_grid(32, 32, 6, 1, 1000)
    gg = dmf.build_csr(u, v, w)
    cases.append((gg, 0, gg.n - 1, (u, v, w)))
    for g, s, t, (u, v, w) in cases:
        say(f"[bottom-up] {env[0]}={env[1]} n={g.n} mode={mode}")
        with dmf.Engine(g, s, t) as eng:
            p0 = dmf.default_params(g, mode=mode, cascade=64)
            f = eng.static_solve(p0)
            check_engine(eng, g, g.cap, s, t, f, "static")
            cur = w.copy()
            applied = []
            for b, pct in enumerate((2, 6, 1, 4)):
                hi = 1000 if g is gg else 100
                applied.append(dmf.gen_updates(u, v, cur, len(u) * pct // 100, 1600 + b, cap_hi=hi))
                p = (dmf.default_params(g, mode=mode, cascade=64) if b % 2 == 0
                     else dmf.default_params(g, mode=mode, scheduler=dmf.SCHED_LEARNED))
                f = eng.apply_batch(*applied[-1], p)
                check_engine(eng, g, caps_after(g, applied), s, t, f, f"batch {b}")


@pytest.mark.parametrize("mode", [1, 2])
def test_cascade_parameter_on_rmat_and_grid(gpu, mode):
    """The cascade parameter (sub-rounds chained on the device, launched in
    checked chunks) under the fixed schedule and under the learned scheduler
    (which uses it by default), on an R-MAT graph and on a grid where excess
    travels far: exact after every batch."""
    dmf = gpu
    cases = [rmat_case(11, 16, 15)]
    u, v, w = dmf.gen_grid(40, 40, 6, 1, 1000)
    gg = dmf.build_csr(u, v, w)
    cases.append((gg, 0, gg.n - 1, (u, v, w)))
    for g, s, t, (u, v, w) in cases:
        for sched in ("fixed", "learned"):
            p = (dmf.default_params(g, mode=mode, cascade=64) if sched == "fixed"
                 else dmf.default_params(g, mode=mode, scheduler=dmf.SCHED_LEARNED))
            say(f"[cascade param] n={g.n} {sched} mode={mode}")
            with dmf.Engine(g, s, t) as eng:
                f = eng.static_solve(p)
                check_engine(eng, g, g.cap, s, t, f, "static")
                cur = w.copy()
                applied = []
                for b, pct in enumerate((2, 6, 3)):
                    hi = 1000 if g is cases[1][0] else 100
                    applied.append(dmf.gen_updates(u, v, cur, len(u) * pct // 100, 1500 + b, cap_hi=hi))
                    f = eng.apply_batch(*applied[-1], p)
                    check_engine(eng, g, caps_after(g, applied), s, t, f, f"{sched} batch {b}")


@pytest.mark.parametrize("interval", [1, 3])
def test_round_graphs_on_a_deep_grid(gpu, interval):
    """Round graphs (mode 3) on a grid whose relabels run deeper than the
    longest captured round (64 level launches): each such relabel stops the
    replay, goes on from the host through level segments and ends its round
    with direct launches; a fresh engine's first relabel does the same with
    the shortest round.  Exact after the static solve and every batch, with a
    relabel every round and every third round."""
    dmf = gpu
    u, v, w = dmf.gen_grid(40, 40, 8, 1, 1000)
    g = dmf.build_csr(u, v, w)
    s, t = 0, g.n - 1
    p = dmf.default_params(g, mode=3, relabel_interval=interval)
    say(f"[rounds grid] interval={interval}")
    with dmf.Engine(g, s, t) as eng:
        f = eng.static_solve(p)
        check_engine(eng, g, g.cap, s, t, f, "static")
        assert eng.counters()["global_relabels"] >= 1
        cur = w.copy()
        applied = []
        for b, pct in enumerate((3, 8)):
            applied.append(dmf.gen_updates(u, v, cur, len(u) * pct // 100, 2500 + b, cap_hi=1000))
            f = eng.apply_batch(*applied[-1], p)
            check_engine(eng, g, caps_after(g, applied), s, t, f, f"batch {b}")


def test_autotuned_schedule_keeps_results(gpu):
    """dmf_engine_autotune_schedule picks thresholds and, for the fixed
    schedule, a relabel interval among its candidates; results stay exact
    under the tuned parameters, and the learned scheduler's are unchanged."""
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(12, 16, 12)
    with dmf.Engine(g, s, t) as eng:
        p = dmf.default_params(g)
        lm, wt, iv = eng.autotune_schedule(p)
        assert lm > 0 and wt > 0 and iv in (1, 2, 3, 5, 8) and p.relabel_interval == iv
        say(f"[autotune schedule] lane_max={lm} wide_t={wt} interval={iv}")
        f = eng.static_solve(p)
        check_engine(eng, g, g.cap, s, t, f, "tuned static")
        cur = w.copy()
        src, dst, nw = dmf.gen_updates(u, v, cur, len(u) // 25, 78)
        f = eng.apply_batch(src, dst, nw, p)
        check_engine(eng, g, caps_after(g, [(src, dst, nw)]), s, t, f, "tuned batch")
        q = dmf.default_params(g, scheduler=dmf.SCHED_LEARNED)
        eng.autotune_schedule(q)
        assert q.relabel_interval == 1
