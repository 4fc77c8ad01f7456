"""Shared helpers for the tests (graph builders and cap bookkeeping)."""
import numpy as np

import dyn_maxflow_amd as dmf


def rmat_case(scale, ef, seed):
    u, v, w = dmf.gen_rmat(scale, ef, seed)
    g = dmf.build_csr(u, v, w)
    s, t = dmf.pick_source_sink(g.n, u, v)
    return g, s, t, (u, v, w)


def arc_index(g):
    """dict (tail, head) -> arc id (last wins, as the reference's scan)."""
    tails = g.tails()
    return {(int(a), int(b)): e for e, (a, b) in enumerate(zip(tails, g.col))}


def caps_after(g, updates, base=None):
    """Capacities after applying update batches in order (last record wins)."""
    idx = arc_index(g)
    cap = (g.cap if base is None else base).copy()
    for src, dst, w in updates:
        for a, b, c in zip(src, dst, w):
            if a == b:
                continue
            cap[idx[(int(a), int(b))]] = c
    return cap


def small_random_graph(rng, n, m, cap_hi=20):
    u = rng.integers(0, n, m).astype(np.int32)
    v = rng.integers(0, n, m).astype(np.int32)
    keep = u != v
    u, v = u[keep], v[keep]
    w = rng.integers(1, cap_hi + 1, len(u)).astype(np.int32)
    # make sure ids 0..n-1 all appear so n is fixed
    u = np.concatenate([u, np.array([n - 1], np.int32)])
    v = np.concatenate([v, np.array([0], np.int32)])
    w = np.concatenate([w, np.array([1], np.int32)])
    return dmf.build_csr(u, v, w), (u, v, w)


def arc_positions(g, src, dst):
    """CSR arc index of every (src, dst) update record, vectorised (last arc of
    a repeated (tail, head) pair, as the reference's scan; every pair must exist)."""
    key = g.tails().astype(np.int64) * g.n + g.col
    order = np.argsort(key, kind="stable")
    sk = key[order]
    k = np.asarray(src, np.int64) * g.n + np.asarray(dst, np.int64)
    pos = np.searchsorted(sk, k, side="right") - 1
    assert (pos >= 0).all() and np.array_equal(sk[pos], k), "update names an absent arc"
    return order[pos]


def caps_sequence(g, batches):
    """Capacities after each of the successive batches (last record wins),
    vectorised for graphs of many million arcs: a list of len(batches) arrays."""
    cap = g.cap.copy()
    out = []
    for src, dst, w in batches:
        src, dst, w = (np.asarray(x, np.int32) for x in (src, dst, w))
        keep = src != dst
        cap[arc_positions(g, src[keep], dst[keep])] = w[keep]
        out.append(cap.copy())
    return out


def parallel_map(fn, items, workers=None):
    """fn over items on a thread pool (the oracle's and the library's ctypes
    calls release the interpreter lock), results in item order."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    workers = workers or max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(fn, items))


def _splitmix(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def multi_instance_specs(count=256, log_lo=10.0, log_hi=16.0, ef_lo=8, ef_hi=32, seed=4):
    """(n, edge factor, seed) of every instance of the batched-instances
    workload, drawn as dmf_multi_bench draws them (csrc/multi_bench.cpp)."""
    import math
    st = seed
    out = []
    for i in range(count):
        st, z = _splitmix(st)
        x = log_lo + (log_hi - log_lo) * ((z >> 11) * (1.0 / 9007199254740992.0))
        st, z = _splitmix(st)
        ef = ef_lo + int(math.floor(((z >> 11) * (1.0 / 9007199254740992.0)) * (ef_hi - ef_lo + 1)))
        out.append((max(2, int(math.floor(2.0 ** x))), ef, seed * 1000003 + i))
    return out


def multi_instance(spec, batches=5, pct=5.0):
    """One instance of the batched-instances workload: (g, s, t, (u, v, w), update batches)."""
    import math
    n, ef, iseed = spec
    u, v, w = dmf.gen_rmat_n(n, n * ef, iseed)
    g = dmf.build_csr(u, v, w)
    s, t = dmf.pick_source_sink(g.n, u, v)
    cur = w.copy()
    count = max(1, int(math.floor(len(u) * pct / 100.0)))
    bl = [dmf.gen_updates(u, v, cur, count, iseed * 31 + b + 1) for b in range(batches)]
    return g, s, t, (u, v, w), bl


def serving_batches(u, v, w, count, kmin=1000, kmax=1000000, inc=0.4, seed=5):
    """The first `count` update batches of the streaming-serving workload as
    dmf_serve_bench draws them (csrc/serve_bench.cpp): sizes log-uniform in
    [kmin, kmax], successive, inc of them increases."""
    import math
    st = seed * 7919 + 11
    cur = w.copy()
    l0, l1 = math.log(kmin), math.log(kmax)
    out = []
    for b in range(count):
        st, z = _splitmix(st)
        k = min(len(u), int(math.exp(l0 + (l1 - l0) * ((z >> 11) * (1.0 / 9007199254740992.0)))))
        out.append(dmf.gen_updates(u, v, cur, k, seed * 1000003 + b, inc_frac=inc))
    return out
