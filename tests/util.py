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
