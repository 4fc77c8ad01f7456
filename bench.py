"""Benchmark: ms per update batch on RMAT-20 (1%..10% successive mixed batches).

Workload (one GPU): directed R-MAT scale 20, edge factor 16, a/b/c/d =
0.57/0.19/0.19/0.05, self-loops and duplicates dropped, seed 2, capacities
uniform [1,100]; s = max out-degree vertex, t = max in-degree vertex != s.
One repeat = reset to the original graph, static solve (untimed; push-pull by
default, for every pass alike, so that the scheduler may send a batch to any
variant), then 10 successive update batches of 1%, 2%, ..., 10% of the input
edges (50/50 increase/decrease).  A step is one batch: apply + repair +
re-converge, timed with HIP events on the engine's stream, inputs already in
HBM.

Passes: the learned scheduler (headline `value`), the plain fixed schedule of
the reference for the data-driven, push-pull and topology-driven variants
(global relabel every round), and a profiled plain pass that times every BFS
level launch of the default driver for the roofline.  After the timed runs
every batch's flow and S_min are checked against the oracle's Dinic, and the
oracle's CPU push-relabel is timed on a bounded sample (single-threaded and
with OpenMP).

Prints one JSON line (see build_line).  --gpus > 1 is not supported on this
build (one GPU): it exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_ARC, BYTES_VERTEX, BYTES_UPDATE = 12, 16, 24
# HBM traffic of the dominant kernel measured with rocprofv3 --pmc (FETCH_SIZE,
# WRITE_SIZE per dispatch, scaled by a plain copy of known size); written from
# the committed profile summary of this workload
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_bfs_level_dev_rmat20_traffic.json")
DRIVERS = {0: "persistent", 1: "stepped", 2: "stepped, BFS levels replayed from hipGraphs",
           3: "round graphs (each fixed-schedule round replayed from one hipGraph)"}
LEVEL_KERNEL = {0: "k_solve_persistent", 1: "k_bfs_level", 2: "k_bfs_level_dev", 3: "k_bfs_level_dev"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def nearest_rank(xs, q):
    xs = sorted(xs)
    if not xs:
        return float("nan")
    k = max(1, int(np.ceil(q / 100.0 * len(xs))))
    return float(xs[k - 1])


def make_workload(scale, ef, seed, npct):
    import dyn_maxflow_amd as dmf
    t0 = time.time()
    u, v, w = dmf.gen_rmat(scale, ef, seed)
    g = dmf.build_csr(u, v, w)
    s, t = dmf.pick_source_sink(g.n, u, v)
    cur = w.copy()
    batches = []
    for pct in range(1, npct + 1):
        src, dst, nw = dmf.gen_updates(u, v, cur, len(u) * pct // 100, seed * 1000 + pct)
        batches.append((src, dst, nw))
    log(f"[bench] workload rmat-{scale} ef{ef} seed {seed}: n={g.n} m={g.m} m_in={len(u)} "
        f"s={s} t={t} batches={[len(b[0]) for b in batches]} ({time.time() - t0:.1f}s)")
    return g, s, t, (u, v, w), batches


class HipBuffers:
    """Device buffers through the HIP runtime libdmf itself is linked against
    (a second HIP runtime in the process, such as the one bundled with a
    PyTorch wheel, cannot share the device with it)."""

    def __init__(self):
        import ctypes
        self.C = ctypes
        for name in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so.7", "libamdhip64.so"):
            try:
                self.hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a, dtype=np.int32)
        p = self.C.c_void_p()
        if self.hip.hipMalloc(self.C.byref(p), max(4, a.nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        if a.nbytes and self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) != 0:  # host -> device
            raise RuntimeError("hipMemcpy failed")
        self.ptrs.append(p)
        return p.value

    def sync(self):
        self.hip.hipDeviceSynchronize()

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def run_gpu(g, s, t, batches, params, steps, warmup, label, profile=False, prof=None, static_params=None,
            sides=None):
    """Returns per-batch records for `steps` timed batches after `warmup` untimed.
    With profile, every BFS level launch is timed by HIP events and the BFS
    counters of every solve (static ones included) are summed into prof.
    sides (a dict) receives S_min after each batch of the first repeat, read
    outside the timed span."""
    import dyn_maxflow_amd as dmf
    bufs = HipBuffers()
    dev = [tuple(bufs.upload(a) for a in b) for b in batches]
    bufs.sync()
    recs, flows = [], {}
    done = 0
    total = steps + warmup
    with dmf.Engine(g, s, t) as eng:
        eng.set_profiling(profile)

        def account(c):
            if prof is not None:
                for key in ("bfs_level_ms", "bfs_level_launches", "bfs_arc_visits", "bfs_vertex_visits"):
                    prof[key] = prof.get(key, 0) + c[key]
        while done < total:
            eng.reset()
            eng.static_solve(static_params if static_params is not None else params)
            account(eng.counters())
            for bi, (ds, dd, dw) in enumerate(dev):
                if done >= total:
                    break
                bufs.sync()
                w0 = time.perf_counter()
                f = eng.apply_batch_device(len(batches[bi][0]), ds, dd, dw, params)
                wall = (time.perf_counter() - w0) * 1e3
                c = eng.counters()
                account(c)
                if bi in flows and flows[bi] != f:
                    raise RuntimeError(f"{label}: batch {bi} flow {f} differs from earlier repeat {flows[bi]}")
                flows[bi] = f
                if sides is not None and bi not in sides:
                    sides[bi] = eng.min_cut()
                if done >= warmup:
                    c["wall_ms"] = wall
                    c["batch"] = bi
                    c["flow"] = f
                    recs.append(c)
                done += 1
            log(f"[bench] {label}: {done}/{total} batches")
    bufs.free()
    return recs, flows


def summarize(recs):
    ms = [r["device_ms"] for r in recs]
    arcs = sum(r["arc_visits"] for r in recs)
    verts = sum(r["vertex_visits"] for r in recs)
    upds = sum(r["updates"] for r in recs)
    tsec = sum(ms) / 1e3
    bytes_alg = BYTES_ARC * arcs + BYTES_VERTEX * verts + BYTES_UPDATE * upds
    return {
        "p50_ms": nearest_rank(ms, 50),
        "p99_ms": nearest_rank(ms, 99),
        "mean_ms": float(np.mean(ms)),
        "wall_p50_ms": nearest_rank([r["wall_ms"] for r in recs], 50),
        "gteps": arcs / tsec / 1e9 if tsec > 0 else 0.0,
        "achieved_gbs": bytes_alg / tsec / 1e9 if tsec > 0 else 0.0,
        "bytes_per_batch": bytes_alg / max(1, len(recs)),
        "rounds_mean": float(np.mean([r["rounds"] for r in recs])),
        "launches_per_batch": float(np.mean([r["launches"] for r in recs])),
        "global_relabels_mean": float(np.mean([r["global_relabels"] for r in recs])),
        # atomic operations issued by the kernels (claims, excess transfers,
        # queue reservations), counted on the device
        "atomics_per_batch": float(np.mean([r["atomics"] for r in recs])),
        "atomic_gops": sum(r["atomics"] for r in recs) / tsec / 1e9 if tsec > 0 else 0.0,
    }


def pass_object(S):
    """The per-pass object of the JSON line."""
    return {"p50_ms": round(S["p50_ms"], 4), "p99_ms": round(S["p99_ms"], 4), "mean_ms": round(S["mean_ms"], 4),
            "gteps": round(S["gteps"], 4), "rounds_per_batch": S["rounds_mean"],
            "global_relabels_per_batch": S["global_relabels_mean"],
            "launches_per_batch": S["launches_per_batch"],
            "achieved_gbs": round(S["achieved_gbs"], 2),
            "roofline_frac": round(S["achieved_gbs"] / HBM_PEAK_GBS, 4),
            "atomics_per_batch": round(S["atomics_per_batch"]),
            "atomic_gops": round(S["atomic_gops"], 4)}


def caps_sequence(g, batches):
    """Capacities after each successive batch (last record wins)."""
    key = g.tails().astype(np.int64) * g.n + g.col
    order = np.argsort(key, kind="stable")
    sk = key[order]
    cap = g.cap.copy()
    out = []
    for src, dst, w in batches:
        keep = src != dst
        k = src[keep].astype(np.int64) * g.n + dst[keep]
        pos = order[np.searchsorted(sk, k, side="right") - 1]
        cap[pos] = w[keep]
        out.append(cap.copy())
    return out


def verify(g, s, t, batches, flows, sides):
    """Flow and S_min after every successive batch vs the oracle's Dinic (CPU,
    one thread per batch)."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    caps = caps_sequence(g, batches)
    with ThreadPoolExecutor(max_workers=min(len(caps), 6)) as ex:
        ref = list(ex.map(lambda c: oracle.dinic(g, s, t, cap=c), caps))
    ok = True
    for bi, (fd, side) in enumerate(ref):
        same_cut = bi in sides and np.array_equal(sides[bi], side)
        ok = ok and fd == flows[bi] and same_cut
        log(f"[bench] verify batch {bi + 1}: gpu {flows[bi]} dinic {fd} S_min equal {same_cut}")
    return ok


def cpu_baseline(g, s, t, batches, kc, seconds_cap, threads):
    """The oracle's push-relabel with the same update repair, data variant, on
    this host: a bounded sample of the same successive batches, timed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    pr = oracle.PushRelabel(g, s, t, threads=threads)
    try:
        pr.static("data", kc)
        times = []
        t_start = time.time()
        for src, dst, w in batches:
            t0 = time.perf_counter()
            pr.apply(src, dst, w, "data", kc)
            times.append((time.perf_counter() - t0) * 1e3)
            if time.time() - t_start > seconds_cap:
                break
    finally:
        pr.close()
    return {
        "value": round(float(np.mean(times)), 3),
        "unit": "ms/batch",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle push-relabel (data variant, {'OpenMP' if threads > 1 else 'sequential'}, kc={kc}) on the "
                  f"same RMAT-20 graph: static solve untimed, then the first {len(times)} successive batches "
                  f"({', '.join(str(i + 1) + '%' for i in range(len(times)))}) timed",
    }


def roofline(prof, mode):
    """Dominant kernel (the global relabel's level expansion) of the plain
    path on the bench's driver: algorithmic bytes per launch (12 B per arc
    visit + 16 B per vertex visit, device counters) / mean launch duration
    (HIP events on the engine's stream around every graph-replayed segment of
    level launches, or around every level launch on the stepped driver)."""
    n = prof.get("bfs_level_launches", 0)
    ms = prof.get("bfs_level_ms", 0.0)
    if not n or ms <= 0:
        return None
    bytes_alg = BYTES_ARC * prof["bfs_arc_visits"] + BYTES_VERTEX * prof["bfs_vertex_visits"]
    gbs = bytes_alg / (ms / 1e3) / 1e9
    tr = traffic_from_profile()
    if tr and tr.get("kernel") != LEVEL_KERNEL[mode]:
        tr = None  # the committed counter summary is of another driver's level kernel
    return {
        "bound": "hbm",
        "kernel": LEVEL_KERNEL[mode],
        "achieved": round(gbs, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(gbs / HBM_PEAK_GBS, 4),
        "traffic": tr["traffic_bytes_per_launch"] if tr else None,
        "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT) if tr else None,
        # L2 atomic requests per launch (rocprofv3 TCC_ATOMIC, same summary)
        # over this pass's mean launch duration
        "tcc_atomic_per_launch": tr.get("tcc_atomic_per_dispatch") if tr else None,
        "atomic_gops": round(tr["tcc_atomic_per_dispatch"] / (ms / n) / 1e6, 4)
        if tr and tr.get("tcc_atomic_per_dispatch") else None,
        "bytes_alg_per_launch": round(bytes_alg / n),
        "avg_launch_us": round(ms * 1e3 / n, 2),
        "launches": n,
        "pass": f"profiled plain pass on the {DRIVERS[mode]} driver: static solves and batches, every level launch",
    }


def copy_probe(mib):
    """A plain copy of known size: FETCH_SIZE / WRITE_SIZE of its dispatches
    give the scale of those counters."""
    bufs = HipBuffers()
    n = mib << 18  # int32 elements
    a = bufs.upload(np.arange(n, dtype=np.int32))
    b = bufs.upload(np.zeros(n, dtype=np.int32))
    for _ in range(5):
        if bufs.hip.hipMemcpy(bufs.C.c_void_p(b), bufs.C.c_void_p(a), n * 4, 3) != 0:  # device -> device
            raise RuntimeError("hipMemcpy failed")
    bufs.sync()
    bufs.free()
    print(json.dumps({"copy_probe_bytes": n * 4, "copies": 5}), flush=True)


def traffic_from_profile():
    """HBM bytes per level launch from the committed PMC summary, or None."""
    try:
        with open(TRAFFIC_FILE) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return None


def build_line(args, g, s, t, kc, S, L, variants, roof, cpu, cpu1, verified):
    """The one JSON line bench.py prints.  S: plain data-driven summary, L:
    learned summary (or None), variants: {name: summary} of the plain runs."""
    head = L if L is not None else S
    out = {
        "metric": "p50 ms per update batch (RMAT-20, successive 1-10% mixed batches)",
        "value": round(head["p50_ms"], 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["mean_ms"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32/int64",
        "data": "synthetic (seeded R-MAT generator)",
        "config": {
            "workload": f"rmat-{args.scale} ef{args.ef} seed {args.seed}, 10 successive batches 1%..10% "
                        "of the input edges, 50/50 inc/dec, caps U[1,100]",
            "n": g.n, "m": g.m, "s": s, "t": t,
            "variant": "data", "kc": kc, "relabel_interval": args.interval,
            "static_solve_variant": args.static_variant,
            "driver": DRIVERS[args.mode],
        },
        "scheduler": "learned" if L is not None else "fixed",
        "p99_ms": round(head["p99_ms"], 4),
        "wall_p50_ms": round(head["wall_p50_ms"], 4),
        "gteps": round(head["gteps"], 4),
        "rounds_per_batch": head["rounds_mean"],
        "atomics_per_batch": round(head["atomics_per_batch"]),
        "atomic_gops": round(head["atomic_gops"], 4),
        "plain": pass_object(S),
        "learned": pass_object(L) if L is not None else None,
        "plain_variants": {k: pass_object(v) for k, v in variants.items()},
        "learned_speedup_p50": round(S["p50_ms"] / L["p50_ms"], 3) if L is not None else None,
        # the same ratio against the fixed schedule on the round-graph driver
        # (which the learned scheduler cannot use: its rounds end on decisions
        # taken by the host)
        "learned_speedup_p50_vs_plain_rounds": round(variants["data@rounds"]["p50_ms"] / L["p50_ms"], 3)
        if L is not None and "data@rounds" in variants else None,
        "flows_and_smin_verified_vs_dinic": verified,
        "roofline": roof,
        "all_kernels": {"achieved_gbs": round(S["achieved_gbs"], 2),
                        "frac": round(S["achieved_gbs"] / HBM_PEAK_GBS, 4),
                        "bytes_alg_per_batch": round(S["bytes_per_batch"])},
        "cpu_baseline": cpu,
        "cpu_baseline_single_thread": cpu1,
    }
    if cpu is not None:
        out["cpu_baseline"] = dict(cpu, host_nproc=os.cpu_count())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--ef", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--mode", type=int, default=2,
                    help="driver: 0 persistent one-launch solve, 1 stepped, 2 stepped with graph-replayed BFS levels, "
                         "3 round graphs (fixed schedule only; others run as 2)")
    ap.add_argument("--interval", type=int, default=1, help="rounds between global relabels of the plain passes")
    ap.add_argument("--sched", default="both", choices=["both", "plain"],
                    help="time the learned scheduler beside the plain fixed schedule")
    ap.add_argument("--variants", default="data,pp,topo,data+cascade,data@rounds",
                    help="variants of the plain fixed-schedule passes (the first is the headline's plain; "
                         "+cascade adds the engine's cascade sub-rounds; @rounds runs the pass on the "
                         "round-graph driver, mode 3, instead of --mode)")
    ap.add_argument("--cascade", type=int, default=1024,
                    help="cascade sub-round cap of the +cascade passes")
    ap.add_argument("--static-variant", default="pp",
                    help="variant of the untimed static solve that every repeat starts from (the same for "
                         "every pass; push-pull lets every variant take the batches)")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--roofline-steps", type=int, default=30,
                    help="batches of the profiled plain pass that measures the dominant kernel")
    ap.add_argument("--roofline-only", action="store_true",
                    help="run only the profiled pass (the command its rocprofv3 summaries are taken from)")
    ap.add_argument("--only", default=None,
                    help="run only one pass (data, pp, topo, data+cascade or learned) and print its object "
                         "(the command per-pass rocprofv3 summaries are taken from)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (PMC scaling reference) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is described, not built).", file=sys.stderr)
        sys.exit(2)

    import dyn_maxflow_amd as dmf
    if dmf.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    g, s, t, _, batches = make_workload(args.scale, args.ef, args.seed, 10)
    kc = max(1, g.m // g.n)
    static = dmf.default_params(g, variant=args.static_variant, mode=args.mode)

    def plain_params(variant):
        # "<variant>@rounds": the same fixed schedule on the round-graph driver
        base, _, drv = variant.partition("@")
        mode = 3 if drv == "rounds" else args.mode
        return dmf.default_params(g, variant=base, mode=mode, relabel_interval=args.interval)

    if args.only:
        if args.only == "learned":
            p1 = dmf.default_params(g, mode=args.mode if args.mode else 1, scheduler=dmf.SCHED_LEARNED)
        else:
            p1 = plain_params(args.only.split("+")[0])
            if args.only.endswith("+cascade"):
                p1.cascade = args.cascade
        recs, _ = run_gpu(g, s, t, batches, p1, args.steps, args.warmup, args.only, static_params=static)
        print(json.dumps({"pass": args.only, "summary": pass_object(summarize(recs))}), flush=True)
        return
    # Profiled pass: the plain schedule on the bench's driver, with every level
    # launch (graph driver: every replayed segment of them) timed by HIP events.
    prof = {}
    precs, _ = run_gpu(g, s, t, batches, plain_params("data"), args.roofline_steps, 2, "profiled plain",
                       profile=True, prof=prof, static_params=static)
    roof = roofline(prof, args.mode)
    if args.roofline_only:
        print(json.dumps({"roofline": roof, "profiled_p50_ms": nearest_rank([r["device_ms"] for r in precs], 50)}),
              flush=True)
        return
    names = [v for v in args.variants.split(",") if v]
    sides = {}
    variants, flows = {}, None
    for vi, var in enumerate(names):
        # "data+cascade": the reference's fixed schedule on the engine's cascade
        # sub-rounds and early-exit relabels (the mechanisms the learned
        # scheduler runs on), to separate them from the scheduler's decisions
        base = var.split("+")[0]
        pp = plain_params(base)
        if "+cascade" in var:
            pp.cascade = args.cascade
        recs, fl = run_gpu(g, s, t, batches, pp, args.steps, args.warmup, f"plain {var}",
                           static_params=static, sides=sides if vi == 0 else None)
        variants[var] = summarize(recs)
        if flows is None:
            flows = fl
        elif fl != flows:
            raise RuntimeError(f"plain {var} produced different flows from plain {names[0]}")
    S = variants[names[0]]
    L = None
    if args.sched != "plain":
        learned = dmf.default_params(g, mode=args.mode if args.mode else 1, scheduler=dmf.SCHED_LEARNED)
        lrecs, lflows = run_gpu(g, s, t, batches, learned, args.steps, args.warmup, "learned", static_params=static)
        L = summarize(lrecs)
        if lflows != flows:
            raise RuntimeError("learned scheduler produced different flows from the plain path")
    verified = None if args.no_verify else verify(g, s, t, batches, flows, sides)
    cpu = cpu1 = None
    if not args.no_cpu:
        # the GPU box gives this process a 16-core share (OMP_NUM_THREADS)
        threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(g, s, t, batches, kc, args.cpu_seconds, threads)
        cpu1 = cpu_baseline(g, s, t, batches, kc, args.cpu_seconds, 1)
    print(json.dumps(build_line(args, g, s, t, kc, S, L, variants, roof, cpu, cpu1, verified)), flush=True)


if __name__ == "__main__":
    main()
