"""GPU parity at the full sizes of the benchmark configurations.

* RMAT-20 (scale 20, EF16, seed 2), 10 successive batches of 1%..10% of the
  input edges: after the static solve and after every batch, the flow value
  and S_min equal the oracle's Dinic, and the residual state passes
  check_state; after the last batch the preflow is completed into a flow that
  passes conservation.  Plain (the reference's schedule), learned, push-pull
  and topology-driven schedules, on the default driver.
* 256 uneven RMAT instances (the batched-instances workload, seed 4, 5 batches
  of 5%): every instance's flow and S_min equal Dinic and its solo engine
  after every batch; persistent and graph-replayed drivers.
* 4096 x 4096 grid (seed 3, one 10% batch) and RMAT-23 (seed 5, streaming
  batches): too large for Dinic in a test, so they are certified instead.  A
  feasible flow of value F (capacity and conservation, after completion) and
  a cut (S, V \\ S) with s in S, t not in S and capacity F prove both optimal
  (max-flow = min-cut); the state is checked before completion as well, and
  the value is compared with an independent from-scratch solve.
"""
import sys

import numpy as np
import pytest

import oracle
from util import (caps_sequence, multi_instance, multi_instance_specs, parallel_map, rmat_case,
                  serving_batches)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_exact(eng, g, cap, s, t, flow, expect, where):
    """Flow and S_min equal Dinic's; capacities and residual state consistent."""
    fd, side = expect
    assert flow == fd, f"{where}: flow {flow} != dinic {fd}"
    smin = eng.min_cut()
    assert np.array_equal(smin, side), f"{where}: S_min differs on {(smin != side).sum()} vertices"
    res, capd, ex, _ = eng.state()
    assert np.array_equal(capd, cap), f"{where}: capacities differ"
    bad = oracle.check_state(g, res, capd, ex)
    assert not bad, f"{where}: {bad}"


def certify(eng, g, cap, s, t, flow, where, state=True):
    """Optimality certificate without Dinic: S = min_cut() holds s, not t, and
    its cut capacity equals the flow value; the residual state is consistent."""
    side = eng.min_cut()
    assert side[s] == 1 and side[t] == 0, f"{where}: s in S {side[s]}, t in S {side[t]}"
    cc = oracle.cut_capacity(g, cap, side)
    assert cc == flow, f"{where}: cut capacity {cc} != flow {flow}"
    if state:
        res, capd, ex, _ = eng.state()
        assert np.array_equal(capd, cap), f"{where}: capacities differ"
        bad = oracle.check_state(g, res, capd, ex)
        assert not bad, f"{where}: {bad}"
    return side


def complete_and_check(eng, g, cap, sources, sinks, flow, where):
    """Preflow -> flow on the device: same value, conservation, same S_min."""
    before = eng.min_cut()
    fc = eng.complete_flow()
    assert fc == flow, f"{where}: completion changed the flow {flow} -> {fc}"
    res, capd, ex, _ = eng.state()
    assert np.array_equal(capd, cap), f"{where}: capacities differ"
    bad = oracle.check_flow(g, res, capd, ex, sources, sinks, flow)
    assert not bad, f"{where}: {bad}"
    after = eng.min_cut()
    assert np.array_equal(before, after), f"{where}: S_min moved by the completion"
    return res, ex


# ------------------------------------------------------------------ RMAT-20
@pytest.fixture(scope="module")
def rmat20(gpu):
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(20, 16, 2)
    cur = w.copy()
    # bench.py's workload: batch i (1..10) holds i% of the input edges
    batches = [dmf.gen_updates(u, v, cur, len(u) * pct // 100, 2 * 1000 + pct) for pct in range(1, 11)]
    caps = [g.cap.copy()] + caps_sequence(g, batches)
    say(f"[rmat20] n={g.n} m={g.m}: Dinic on {len(caps)} capacity sets")
    expect = parallel_map(lambda c: oracle.dinic(g, s, t, cap=c), caps, workers=6)
    return g, s, t, batches, caps, expect


@pytest.mark.parametrize("schedule", ["plain", "learned", "pp", "topo", "plain-rounds"])
def test_rmat20_every_batch_exact(gpu, rmat20, schedule):
    dmf = gpu
    g, s, t, batches, caps, expect = rmat20
    p = {"plain": dmf.default_params(g),
         "plain-rounds": dmf.default_params(g, mode=3),
         "learned": dmf.default_params(g, scheduler=dmf.SCHED_LEARNED),
         "pp": dmf.default_params(g, variant="pp"),
         "topo": dmf.default_params(g, variant="topo")}[schedule]
    with dmf.Engine(g, s, t) as eng:
        # every schedule starts from the push-pull static state, as in bench.py
        f = eng.static_solve(dmf.default_params(g, variant="pp"))
        check_exact(eng, g, caps[0], s, t, f, expect[0], f"{schedule} static")
        for i, b in enumerate(batches):
            say(f"[rmat20] {schedule} batch {i + 1} k={len(b[0])}")
            f = eng.apply_batch(*b, p)
            check_exact(eng, g, caps[i + 1], s, t, f, expect[i + 1], f"{schedule} batch {i + 1}")
        complete_and_check(eng, g, caps[-1], [s], [t], f, f"{schedule} completed")


# -------------------------------------------------- 256 uneven RMAT instances
@pytest.fixture(scope="module")
def instances256(gpu):
    dmf = gpu
    specs = multi_instance_specs()
    cases = parallel_map(multi_instance, specs)
    nv = sum(c[0].n for c in cases)
    na = sum(c[0].m for c in cases)
    nu = sum(len(b[0]) for c in cases for b in c[4])
    say(f"[256] vertices={nv} arcs={na} updates={nu}")
    # the totals dmf_multi_bench reports for this workload
    assert (nv, na, nu) == (4351489, 127438266, 16938100)
    caps = parallel_map(lambda c: [c[0].cap.copy()] + caps_sequence(c[0], c[4]), cases)
    jobs = [(i, j) for i in range(len(cases)) for j in range(6)]
    dres = parallel_map(lambda ij: oracle.dinic(cases[ij[0]][0], cases[ij[0]][1], cases[ij[0]][2],
                                                cap=caps[ij[0]][ij[1]]), jobs)
    expect = [[dres[i * 6 + j] for j in range(6)] for i in range(len(cases))]
    # solo engines, one after another (stepped driver, as the solo baseline)
    solo = []
    for i, (g, s, t, _, bl) in enumerate(cases):
        p = dmf.default_params(g, mode=1)
        with dmf.Engine(g, s, t) as eng:
            fl = [eng.static_solve(p)]
            fl += [eng.apply_batch(*b, p) for b in bl]
        solo.append(fl)
    say("[256] Dinic and solo engines done")
    return cases, caps, expect, solo


@pytest.mark.parametrize("mode", [2, 0, 3])
def test_256_instances_every_batch_exact(gpu, instances256, mode):
    dmf = gpu
    cases, caps, expect, solo = instances256
    k = len(cases)
    graphs = [c[0] for c in cases]
    with dmf.BatchEngine(graphs, [c[1] for c in cases], [c[2] for c in cases]) as be:
        p = dmf.default_params(graphs[0], mode=mode)
        lm, wt = be.autotune(p)
        say(f"[256] mode={mode} autotuned lane_max={lm} wide_t={wt}")

        def check(flows, j, where):
            sides = be.min_cut_all()
            for i in range(k):
                fd, side = expect[i][j]
                assert flows[i] == fd == solo[i][j], f"{where} instance {i}: {flows[i]} dinic {fd} solo {solo[i][j]}"
                assert np.array_equal(sides[i], side), f"{where} instance {i}: S_min differs"

        check(be.static_solve_all(p), 0, "static")
        for b in range(5):
            say(f"[256] mode={mode} batch {b + 1}")
            flows = be.apply_all([c[4][b] for c in cases], p)
            assert np.array_equal(flows, be.flows())
            check(flows, b + 1, f"batch {b + 1}")
        # completion of every instance at once: conservation in the union
        total = int(flows.sum())
        assert be.complete_flow() == total
        assert np.array_equal(be.flows(), flows)
        res, _, ex, _ = be.state()
        voff = be.offsets
        aoff = np.concatenate([[0], np.cumsum([g.m for g in graphs])])
        for i, g in enumerate(graphs):
            r = res[aoff[i]:aoff[i + 1]]
            x = ex[voff[i]:voff[i + 1]]
            bad = oracle.check_flow(g, r, caps[i][5], x, [cases[i][1]], [cases[i][2]], int(flows[i]))
            assert not bad, f"completed instance {i}: {bad}"


# ---------------------------------------------------------------- grid 4096^2
def test_grid4096_one_batch_certified(gpu):
    dmf = gpu
    u, v, w = dmf.gen_grid(4096, 4096, 3, 1, 1000)
    g = dmf.build_csr(u, v, w)
    s, t = 0, g.n - 1
    cur = w.copy()
    batch = dmf.gen_updates(u, v, cur, int(len(u) * 10 / 100), 3 * 1000 + 1, inc_frac=0.5, cap_lo=1, cap_hi=1000)
    cap1 = caps_sequence(g, [batch])[0]
    static_p = dmf.default_params(g, cascade=1024)
    say(f"[grid] n={g.n} m={g.m} k={len(batch[0])}: static solve")
    with dmf.Engine(g, s, t) as eng:
        f0 = eng.static_solve(static_p)
        certify(eng, g, g.cap, s, t, f0, "grid static")
        say("[grid] learned batch")
        f1 = eng.apply_batch(*batch, dmf.default_params(g, scheduler=dmf.SCHED_LEARNED))
        certify(eng, g, cap1, s, t, f1, "grid learned batch")
        complete_and_check(eng, g, cap1, [s], [t], f1, "grid completed")
        certify(eng, g, cap1, s, t, f1, "grid completed", state=False)
    # an independent from-scratch solve of the updated grid
    say("[grid] static solve of the updated grid")
    g2 = dmf.Graph(g.n, g.m, g.row, g.col, cap1, g.rev)
    with dmf.Engine(g2, s, t) as eng:
        assert eng.static_solve(static_p) == f1


# ---------------------------------------------------- RMAT-23 streaming batches
def test_rmat23_serving_batches_certified(gpu):
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(23, 16, 5)
    batches = serving_batches(u, v, w, 6)
    caps = caps_sequence(g, batches)
    say(f"[rmat23] n={g.n} m={g.m} batches={[len(b[0]) for b in batches]}")
    # served through the streaming server (learned scheduler, its default)
    with dmf.Server(dmf.server_config()) as srv:
        sid, f_static = srv.open(g, s, t, max(len(b[0]) for b in batches))
        tickets = [srv.submit(sid, *b) for b in batches]
        served = []
        for tk in tickets:
            r = srv.wait(tk)
            assert r.status == dmf.OK
            served.append(r.flow)
        srv.close(sid)
    say(f"[rmat23] served flows {served}")
    # the same batches on an engine, certified after each
    with dmf.Engine(g, s, t) as eng:
        p = dmf.default_params(g, scheduler=dmf.SCHED_LEARNED)
        f = eng.static_solve(p)
        assert f == f_static
        certify(eng, g, g.cap, s, t, f, "rmat23 static", state=False)
        for i, b in enumerate(batches):
            f = eng.apply_batch(*b, p)
            assert f == served[i], f"batch {i}: engine {f} != served {served[i]}"
            last = i == len(batches) - 1
            certify(eng, g, caps[i], s, t, f, f"rmat23 batch {i}", state=last or i == 0)
        complete_and_check(eng, g, caps[-1], [s], [t], f, "rmat23 completed")
    # the reference's dynamic == static check (code/dyn_data_bench.cu:535-539)
    allb = [np.concatenate([b[j] for b in batches]) for j in range(3)]
    stat, _ = dmf.static_recalculate(g, *allb, s, t, max(1, g.m // g.n), "data")
    assert stat == f
