"""Trains the learned scheduler from traces the engine itself generates.

  python -m dyn_maxflow_amd.sched_train collect --out DIR [--workloads SET | --graph G.bin --source S --sink T]
                                                                             (needs the GPU)
  python -m dyn_maxflow_amd.sched_train train --traces DIR [DIR ...] [--teacher batch|workload]
                                              [--variants data,topo,pp] [--variant-teacher same|workload]
                                              (CPU; writes csrc/sched_weights.inc)
  python -m dyn_maxflow_amd.sched_train retrofit --traces DIR --out DIR
      (traces taken before feature 0 carried the graph's size at batch start)

collect: for one set of training workloads, every batch is solved under every
fixed schedule in {data, topo, pp} x {relabel every 1, 2, 3, 5, 8 rounds},
data-driven ones with the engine's cascade push sub-rounds (as the learned
scheduler runs them) and also with the lazy start (no relabel before the
first round), each from a push-pull static state so that every variant may
take the batches.  Sets (WORKLOAD_SETS): base = R-MAT scale 15-20 with seeds 11-16
and 128/256 grids; large = R-MAT scale 20 with seeds 17 and 18 and scale 21
with seed 19; grids = a 512 grid; xlarge = R-MAT scale 22 with seed 20;
small = R-MAT scale 21 and 22 with seeds 21 and 22.  Batches are 1-10% of the
input arcs with several increase/decrease mixes, except for small (0.002-0.5%,
40-60% increases).  None uses seed 2 (the RMAT-20 bench graph), seed 3 (the
evaluation grids) or the scale-23 serving graph.  The engine writes one trace
line per scheduler evaluation (the 8 device-computed features, the round,
whether that round started with a global relabel) and one summary line per
solve (device ms, variant, interval).

train: the fastest schedule of each batch, among the variants --variants
allows, is its teacher (--teacher workload: the schedule with the lowest total
time over all of a workload's batches).  The variant head learns that
schedule's variant from the batch-start features; the relabel head learns
whether the teacher relabelled before its first round (from the batch-start
features; data-driven schedules may start on the labels of the previous
solve) and, from the features at the end of each round of the teacher's run,
whether it relabelled before the next round.  A numpy MLP 8->32->32->4 (ReLU)
is fitted with Adam; the weights are written in the row-major SchedWeights
order of csrc/sched.hpp.  The shipped weights come from all five sets on the
stepped driver (collect --mode 1), all variants, per-batch teachers; train
--save-examples writes the examples it fitted (sched_examples.npz, committed)
and train --examples refits from that file alone, reproducing the weights.
"""
from __future__ import annotations

import argparse
import glob
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
WEIGHTS = os.path.join(HERE, "csrc", "sched_weights.inc")
FEAT, HID, OUT = 8, 32, 4
VARIANTS = ["data", "topo", "pp"]
INTERVALS = [1, 2, 3, 5, 8]
# cascade sub-round cap of the data-driven teacher schedules and of the
# static solves (the learned scheduler's default when its parameters give none)
COLLECT_CASCADE = 1024
EXAMPLES = os.path.join(HERE, "sched_examples.npz")


WORKLOAD_SETS = {
    # (kind, scale or side, seed)
    "base": [("rmat", 15, 11), ("rmat", 16, 12), ("rmat", 17, 13), ("rmat", 18, 14), ("rmat", 19, 15),
             ("rmat", 20, 16), ("grid", 128, 21), ("grid", 256, 22)],
    "large": [("rmat", 20, 17), ("rmat", 20, 18), ("rmat", 21, 19)],
    "grids": [("grid", 512, 23)],
    "xlarge": [("rmat", 22, 20)],
    "small": [("rmat", 21, 21), ("rmat", 22, 22)],
}
# batches as (percent of the input arcs, fraction of increases)
BATCHES = {"small": ((0.002, 0.4), (0.01, 0.4), (0.05, 0.4), (0.2, 0.4), (0.5, 0.4), (0.02, 0.6))}
DEFAULT_BATCHES = ((1, 0.5), (3, 0.5), (5, 0.2), (10, 0.5), (2, 0.8), (7, 0.5))


def _workloads(which="base"):
    import dyn_maxflow_amd as dmf
    for kind, size, seed in WORKLOAD_SETS[which]:
        if kind == "rmat":
            u, v, w = dmf.gen_rmat(size, 16, seed)
            yield (f"rmat{size}s{seed}", u, v, w, 1, 100)
        else:
            u, v, w = dmf.gen_grid(size, size, seed, 1, 1000)
            yield (f"grid{size}s{seed}", u, v, w, 1, 1000)


def graph_workload(path: str):
    """A user's graph (.bin, the reference's graph format) as a training
    workload: its arcs with positive capacity are the edges the update batches
    draw from, with new capacities in [1, max capacity]."""
    import dyn_maxflow_amd as dmf
    g = dmf.read_graph_bin(path)
    keep = g.cap > 0
    u = g.tails()[keep].astype(np.int32)
    v = g.col[keep].astype(np.int32)
    w = g.cap[keep].astype(np.int32)
    name = os.path.splitext(os.path.basename(path))[0]
    return name, u, v, w, 1, int(max(1, w.max(initial=1)))


def collect(out_dir: str, quick: bool = False, mode: int = 1, which: str = "base", graph: str = None,
            source: int = None, sink: int = None) -> None:
    """graph: collect on this .bin graph (source/sink given) instead of a set."""
    sys.path.insert(0, ROOT)
    import dyn_maxflow_amd as dmf
    os.makedirs(out_dir, exist_ok=True)
    workloads = [graph_workload(graph)] if graph else _workloads(which)
    for name, u, v, w, lo, hi in workloads:
        g = dmf.build_csr(u, v, w)
        if graph:
            s, t = source, sink
        elif name.startswith("grid"):
            s, t = 0, g.n - 1
        else:
            s, t = dmf.pick_source_sink(g.n, u, v)
        batches = []
        cur = w.copy()
        for i, (pct, inc) in enumerate(BATCHES.get("custom" if graph else which, DEFAULT_BATCHES)):
            src, dst, nw = dmf.gen_updates(u, v, cur, max(1, int(len(u) * pct / 100)), 7000 + i,
                                           inc_frac=inc, cap_lo=lo, cap_hi=hi)
            batches.append((src, dst, nw))
        trace = os.path.join(out_dir, f"{name}.jsonl")
        if os.path.exists(trace):
            os.remove(trace)
        os.environ["DMF_TRACE"] = trace
        with dmf.Engine(g, s, t) as eng:
            os.environ.pop("DMF_TRACE", None)
            # push-pull static state: every variant may take the batches
            base = dmf.default_params(g, variant="pp", mode=mode, cascade=COLLECT_CASCADE)
            for variant in VARIANTS:
                for interval in ([1, 3] if quick else INTERVALS):
                    for lazy in ((0, 1) if variant == "data" else (0,)):
                        eng.reset()
                        eng.static_solve(base)
                        p = dmf.default_params(g, variant=variant, relabel_interval=interval, lazy_start=lazy,
                                               mode=mode, cascade=COLLECT_CASCADE if variant == "data" else 0)
                        for src, dst, nw in batches:
                            eng.apply_batch(src, dst, nw, p)
                        print(f"[collect] {name} {variant} interval={interval} lazy={lazy}", flush=True)


def _graph_size(name: str) -> int:
    """Vertex count of a training workload, regenerated from its name."""
    import re
    import dyn_maxflow_amd as dmf
    m = re.match(r"rmat(\d+)s(\d+)$", name)
    if m:
        u, v, w = dmf.gen_rmat(int(m.group(1)), 16, int(m.group(2)))
        return dmf.build_csr(u, v, w).n
    m = re.match(r"grid(\d+)s(\d+)$", name)
    if m:
        return int(m.group(1)) ** 2
    raise ValueError(f"unknown workload {name}")


def retrofit(src_dir: str, out_dir: str) -> None:
    """Rewrites traces taken before feature 0 described the graph at batch
    start (it was 0 there): each batch row gets log2(1+n)/32, the value the
    engine now computes, with n regenerated from the workload's name."""
    os.makedirs(out_dir, exist_ok=True)
    for f in sorted(glob.glob(os.path.join(src_dir, "*.jsonl"))):
        name = os.path.basename(f)[:-len(".jsonl")]
        val = round(float(np.float32(np.log2(1.0 + _graph_size(name))) / np.float32(32.0)), 6)
        with open(f) as fi, open(os.path.join(out_dir, os.path.basename(f)), "w") as fo:
            for line in fi:
                rec = json.loads(line)
                if rec.get("what") == "batch":
                    rec["feat"][0] = val
                fo.write(json.dumps(rec, separators=(",", ":")) + "\n")
        print(f"[retrofit] {name}: feature 0 at batch start = {val}")


def _load(trace_dir: str):
    """Returns per-file lists of solves: dict(summary, rows)."""
    files = sorted(glob.glob(os.path.join(trace_dir, "*.jsonl")))
    data = {}
    for f in files:
        solves = {}
        for line in open(f):
            rec = json.loads(line)
            solves.setdefault(rec["solve"], {"rows": []})
            if rec["what"] == "summary":
                solves[rec["solve"]]["summary"] = rec
            else:
                solves[rec["solve"]]["rows"].append(rec)
        data[os.path.basename(f)] = [solves[k] for k in sorted(solves)]
    return data


def _examples(data, teacher="batch", variants=(0, 1, 2), variant_teacher="same"):
    """(X_relabel, y_relabel, X_variant, y_variant) from the teacher schedules.
    teacher="workload": the schedule with the lowest total time over all of a
    workload's batches teaches every batch of it (per-batch winners flip
    between near-equal schedules on timing noise); "batch": the fastest
    schedule of each batch.  variant_teacher="workload": every batch of a
    workload is taught the variant of that workload's best schedule, and its
    relabel decisions come from the fastest run of that variant on the batch
    (per-batch variant winners are noisier than the relabel timing)."""
    Xr, yr, Xv, yv = [], [], [], []
    for fname, solves in data.items():
        # batch solves grouped by schedule, in order; static solves (entry 0) skipped
        runs = {}
        for sv in solves:
            sm = sv.get("summary")
            if not sm or sm["entry"] != 1 or sm["scheduler"] != 0:
                continue
            if sm["variant"] not in variants:
                continue
            # data-driven teachers as the learned scheduler runs them: with
            # cascade sub-rounds (traces that record none predate them)
            if sm["variant"] == 0 and "cascade" in sm and sm["cascade"] <= 0:
                continue
            key = (sm["variant"], sm["interval"], sm.get("lazy", 0))
            runs.setdefault(key, []).append(sv)
        if not runs:
            continue
        nb = min(len(r) for r in runs.values())
        overall = min(runs, key=lambda k: sum(r["summary"]["device_ms"] for r in runs[k][:nb]))
        for b in range(nb):
            if variant_teacher == "workload":
                cand = [k for k in runs if k[0] == overall[0]]
                best = min(cand, key=lambda k: runs[k][b]["summary"]["device_ms"])
            else:
                best = overall if teacher == "workload" else min(runs, key=lambda k: runs[k][b]["summary"]["device_ms"])
            run = runs[best][b]
            batch_rows = [r for r in run["rows"] if r["what"] == "batch"]
            rows = [r for r in run["rows"] if r["what"] == "round"]
            if batch_rows:
                Xv.append(batch_rows[0]["feat"])
                yv.append(best[0])
                # batch start: did the teacher relabel before its first round?
                Xr.append(batch_rows[0]["feat"])
                yr.append(rows[0]["relabelled"] if rows else 1)
            for i, r in enumerate(rows):
                nxt = rows[i + 1]["relabelled"] if i + 1 < len(rows) else 1
                Xr.append(r["feat"])
                yr.append(nxt)
    return (np.array(Xr, np.float32), np.array(yr, np.float32), np.array(Xv, np.float32),
            np.array(yv, np.int64))


def _forward(P, X):
    h1 = np.maximum(0, X @ P["w1"] + P["b1"])
    h2 = np.maximum(0, h1 @ P["w2"] + P["b2"])
    return h1, h2, h2 @ P["w3"] + P["b3"]


def fit(Xr, yr, Xv, yv, steps=4000, lr=3e-3, seed=0):
    rng = np.random.default_rng(seed)
    P = {
        "w1": rng.normal(0, np.sqrt(2 / FEAT), (FEAT, HID)).astype(np.float32),
        "b1": np.zeros(HID, np.float32),
        "w2": rng.normal(0, np.sqrt(2 / HID), (HID, HID)).astype(np.float32),
        "b2": np.zeros(HID, np.float32),
        "w3": rng.normal(0, np.sqrt(1 / HID), (HID, OUT)).astype(np.float32),
        "b3": np.zeros(OUT, np.float32),
    }
    m = {k: np.zeros_like(v) for k, v in P.items()}
    s2 = {k: np.zeros_like(v) for k, v in P.items()}
    X = np.concatenate([Xr, Xv]) if len(Xv) else Xr
    nr = len(Xr)
    for it in range(1, steps + 1):
        h1, h2, o = _forward(P, X)
        g = np.zeros_like(o)
        # relabel head: logistic loss on rows [0, nr)
        if nr:
            pr = 1 / (1 + np.exp(-np.clip(o[:nr, 0], -30, 30)))
            g[:nr, 0] = (pr - yr) / nr
        # variant head: softmax over out[1:4] on rows [nr, ...)
        if len(Xv):
            z = o[nr:, 1:4]
            z = z - z.max(axis=1, keepdims=True)
            e = np.exp(z)
            sm = e / e.sum(axis=1, keepdims=True)
            sm[np.arange(len(yv)), yv] -= 1
            g[nr:, 1:4] = sm / len(yv)
        grads = {"w3": h2.T @ g, "b3": g.sum(0)}
        d2 = (g @ P["w3"].T) * (h2 > 0)
        grads["w2"] = h1.T @ d2
        grads["b2"] = d2.sum(0)
        d1 = (d2 @ P["w2"].T) * (h1 > 0)
        grads["w1"] = X.T @ d1
        grads["b1"] = d1.sum(0)
        for k in P:
            m[k] = 0.9 * m[k] + 0.1 * grads[k]
            s2[k] = 0.999 * s2[k] + 0.001 * grads[k] ** 2
            mh = m[k] / (1 - 0.9 ** it)
            sh = s2[k] / (1 - 0.999 ** it)
            P[k] = (P[k] - lr * mh / (np.sqrt(sh) + 1e-8)).astype(np.float32)
    _, _, o = _forward(P, X)
    acc_r = float(((o[:nr, 0] > 0) == (yr > 0.5)).mean()) if nr else float("nan")
    acc_v = float((o[nr:, 1:4].argmax(1) == yv).mean()) if len(Xv) else float("nan")
    return P, acc_r, acc_v


def write_weights(P, note: str, path: str | None = None) -> None:
    flat = np.concatenate([P["w1"].ravel(), P["b1"], P["w2"].ravel(), P["b2"], P["w3"].ravel(), P["b3"]])
    lines = ["// Scheduler weights for sched.hpp, row-major in SchedWeights order.",
             f"// Trained by dyn_maxflow_amd/sched_train.py: {note}",
             "#define DMF_SCHED_TRAINED 1",
             f"static const float kSchedWeightsInit[{len(flat)}] = {{"]
    for i in range(0, len(flat), 6):
        lines.append("  " + ", ".join(f"{float(x)!r}f" for x in flat[i:i + 6]) + ",")
    lines.append("};")
    with open(path or WEIGHTS, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def train(trace_dirs, teacher="batch", variants=(0, 1, 2), variant_teacher="same", examples=None,
          save_examples=None, out=None) -> None:
    """From traces (trace_dirs), or from a saved example set (examples)."""
    if examples:
        with np.load(examples, allow_pickle=False) as z:
            Xr, yr, Xv, yv = z["Xr"], z["yr"], z["Xv"], z["yv"]
    else:
        data = {}
        for d in trace_dirs:
            data.update({f"{d}/{k}": v for k, v in _load(d).items()})
        Xr, yr, Xv, yv = _examples(data, teacher, variants, variant_teacher)
    if save_examples:
        np.savez_compressed(save_examples, Xr=Xr, yr=yr, Xv=Xv, yv=yv)
        print(f"[train] wrote {save_examples}")
    print(f"[train] relabel examples {len(Xr)} (positive {yr.mean() if len(yr) else 0:.2f}), "
          f"variant examples {len(Xv)} (counts {np.bincount(yv, minlength=3).tolist()})")
    P, acc_r, acc_v = fit(Xr, yr, Xv, yv)
    print(f"[train] accuracy: relabel {acc_r:.3f} variant {acc_v:.3f}")
    write_weights(P, f"{len(Xr)} relabel / {len(Xv)} variant examples, train acc "
                     f"{acc_r:.3f} / {acc_v:.3f}", out)
    print(f"[train] wrote {out or WEIGHTS}")


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--out", required=True)
    c.add_argument("--quick", action="store_true")
    c.add_argument("--mode", type=int, default=1, help="driver the traces are taken on (1 stepped, 2 graph)")
    c.add_argument("--workloads", choices=sorted(WORKLOAD_SETS), default="base")
    c.add_argument("--graph", default=None, help="collect on this .bin graph instead of a workload set")
    c.add_argument("--source", type=int, default=None)
    c.add_argument("--sink", type=int, default=None)
    rf = sub.add_parser("retrofit")
    rf.add_argument("--traces", required=True)
    rf.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--traces", nargs="+", default=[])
    t.add_argument("--examples", default=None,
                   help="fit from this saved example set instead of traces (e.g. the committed sched_examples.npz)")
    t.add_argument("--save-examples", default=None, help="write the examples fitted to this .npz file")
    t.add_argument("--teacher", choices=["workload", "batch"], default="batch")
    t.add_argument("--variant-teacher", choices=["same", "workload"], default="same",
                   help="workload: each workload's best variant teaches all its batches")
    t.add_argument("--variants", default="data,topo,pp",
                   help="variants a teacher schedule may use (comma-separated)")
    a = ap.parse_args()
    if a.cmd == "retrofit":
        retrofit(a.traces, a.out)
    elif a.cmd == "collect":
        if a.graph and (a.source is None or a.sink is None):
            ap.error("--graph needs --source and --sink")
        collect(a.out, a.quick, a.mode, a.workloads, a.graph, a.source, a.sink)
    elif a.cmd == "train":
        if not a.traces and not a.examples:
            ap.error("train needs --traces or --examples")
        train(a.traces, a.teacher, tuple(VARIANTS.index(x) for x in a.variants.split(",")), a.variant_teacher,
              a.examples, a.save_examples)


if __name__ == "__main__":
    main()
