"""Learned scheduler vs fixed-frequency global relabel on a high-diameter grid.

  python -m dyn_maxflow_amd.grid_bench [--side 4096] [--seed 3] [--pct 10]
         [--schedules learned,fixed1,fixed1c1024] [--mode 2]
         [--static-interval 1] [--static-cascade 1024]
         [--progress FILE] [--no-check]

Workload: side x side 4-neighbour grid (every neighbour pair in both
directions, independent capacities uniform [1,1000]), s = corner (0,0),
t = corner (side-1, side-1), seed 3; one batch updating pct% of the input
arcs (half increases, half decreases).  For every schedule a fresh engine is
built, solved statically (data variant, global relabel every
--static-interval rounds, at most --static-cascade push sub-rounds after
each round; the same for all schedules), and then takes the batch under
that schedule:
  learned   the device MLP decides when to relabel (and the variant);
  fixedN    data variant, global relabel every N rounds (fixed1 = reference);
  fixedNcK  the same with at most K cascade push sub-rounds after each round.
The batch is timed with HIP events on the engine's stream.  Checks: every
schedule reports the same flow, and a from-scratch static solve of the
updated graph (the reference's dynamic == static check) reports it too.
--progress FILE appends the engine's per-round trace lines to FILE, so a
long solve keeps writing while it runs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import dyn_maxflow_amd as dmf  # noqa: E402


def updated_caps(g, src, dst, w):
    tails = g.tails().astype(np.int64)
    key = tails * g.n + g.col
    order = np.argsort(key)
    k = src.astype(np.int64) * g.n + dst
    pos = order[np.searchsorted(key[order], k)]
    assert np.array_equal(key[pos], k), "update names an arc that is not in the grid"
    cap = g.cap.copy()
    cap[pos] = w
    return cap


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--side", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--pct", type=float, default=10.0)
    ap.add_argument("--schedules", default="learned,fixed1,fixed1c1024")
    ap.add_argument("--mode", type=int, default=2, help="driver: 1 stepped, 2 graph-replayed BFS levels")
    ap.add_argument("--static-interval", type=int, default=1)
    ap.add_argument("--static-cascade", type=int, default=1024)
    ap.add_argument("--progress", default=None)
    ap.add_argument("--no-check", action="store_true")
    a = ap.parse_args()
    if dmf.device_count() < 1:
        print("grid_bench: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    t0 = time.time()
    u, v, w = dmf.gen_grid(a.side, a.side, a.seed, 1, 1000)
    g = dmf.build_csr(u, v, w)
    s, t = 0, g.n - 1
    cur = w.copy()
    src, dst, nw = dmf.gen_updates(u, v, cur, int(len(u) * a.pct / 100), a.seed * 1000 + 1, inc_frac=0.5,
                                   cap_lo=1, cap_hi=1000)
    print(f"[grid] {a.side}x{a.side}: n={g.n} m={g.m} batch={len(src)} ({time.time() - t0:.1f}s)", flush=True)
    if a.progress:
        os.environ["DMF_TRACE"] = a.progress
    static_p = dmf.default_params(g, mode=a.mode, relabel_interval=a.static_interval, cascade=a.static_cascade)
    rows = []
    for name in a.schedules.split(","):
        if name == "learned":
            p = dmf.default_params(g, mode=a.mode, scheduler=dmf.SCHED_LEARNED)
        elif name.startswith("fixed"):
            iv, _, casc = name[5:].partition("c")
            p = dmf.default_params(g, mode=a.mode, relabel_interval=int(iv), cascade=int(casc or 0))
        else:
            raise SystemExit(f"unknown schedule {name}")
        with dmf.Engine(g, s, t) as eng:
            t1 = time.time()
            f0 = eng.static_solve(static_p)
            c0 = eng.counters()
            t2 = time.time()
            f1 = eng.apply_batch(src, dst, nw, p)
            c1 = eng.counters()
            t3 = time.time()
        row = {"schedule": name, "static_flow": f0, "flow": f1, "static_ms": round(c0["device_ms"], 2),
               "batch_ms": round(c1["device_ms"], 2), "batch_wall_s": round(t3 - t2, 2),
               "global_relabels": c1["global_relabels"], "bfs_levels": c1["bfs_levels"], "rounds": c1["rounds"],
               "arc_visits": c1["arc_visits"]}
        rows.append(row)
        print(f"[grid] {name}: static {c0['device_ms']:.1f} ms ({t2 - t1:.1f}s wall), batch {c1['device_ms']:.1f} ms, "
              f"relabels {c1['global_relabels']}, levels {c1['bfs_levels']}, rounds {c1['rounds']}, flow {f1}",
              flush=True)
    os.environ.pop("DMF_TRACE", None)
    flows = {r["flow"] for r in rows}
    check = None
    if not a.no_check:
        g2 = dmf.Graph(g.n, g.m, g.row, g.col, updated_caps(g, src, dst, nw), g.rev)
        with dmf.Engine(g2, s, t) as eng:
            check = eng.static_solve(static_p)
        print(f"[grid] static recompute of the updated grid: flow {check}", flush=True)
    out = {"workload": f"grid {a.side}x{a.side} seed {a.seed}, caps U[1,1000], one batch of {a.pct}% of arcs",
           "n": g.n, "m": g.m, "k": int(len(src)), "driver": {1: "stepped", 2: "graph-replayed BFS"}.get(a.mode),
           "static_interval": a.static_interval, "static_cascade": a.static_cascade, "schedules": rows,
           "flows_agree": len(flows) == 1 and (check is None or check in flows), "static_recompute_flow": check}
    base = [r for r in rows if r["schedule"] != "learned"]
    learned = [r for r in rows if r["schedule"] == "learned"]
    if base and learned:
        out["learned_speedup"] = {r["schedule"]: round(r["batch_ms"] / learned[0]["batch_ms"], 3) for r in base}
    print(json.dumps(out), flush=True)
    if not out["flows_agree"]:
        sys.exit(1)


if __name__ == "__main__":
    main()
