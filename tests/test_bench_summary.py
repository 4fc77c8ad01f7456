"""bench.py's per-pass summary on synthetic counter records (CPU): every
field the JSON line carries is computed from the counters an engine returns."""
import os
import sys

import dyn_maxflow_amd as dmf

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def _record(ms, arcs, atomics):
    r = dmf.Counters().as_dict()  # exactly the fields counters() returns
    counted = dict(device_ms=ms, arc_visits=arcs, vertex_visits=arcs // 4, updates=100, rounds=3, launches=40,
                   global_relabels=2, atomics=atomics)
    assert set(counted) <= set(r), set(counted) - set(r)
    r.update(counted)
    r.update(wall_ms=ms + 0.1, batch=0, flow=1)  # added by bench.run_gpu
    return r


def test_summary_fields_and_rates():
    recs = [_record(2.0, 4_000_000, 1_000_000), _record(4.0, 8_000_000, 3_000_000)]
    S = bench.summarize(recs)
    assert S["p50_ms"] == 2.0 and S["p99_ms"] == 4.0
    assert abs(S["gteps"] - 12e6 / 6e-3 / 1e9) < 1e-9
    assert S["atomics_per_batch"] == 2_000_000
    assert abs(S["atomic_gops"] - 4e6 / 6e-3 / 1e9) < 1e-9


def test_json_line_carries_every_pass_field():
    """The line bench.py prints: headline fields, the atomic figures of the
    headline, plain, learned and per-variant plain objects, and both CPU
    baselines, so that a field dropped from the line fails here."""
    import argparse
    recs = [_record(2.0, 4_000_000, 1_000_000), _record(4.0, 8_000_000, 3_000_000)]
    S = bench.summarize(recs)
    L = bench.summarize([_record(1.0, 4_000_000, 1_000_000)])
    args = argparse.Namespace(steps=2, warmup=1, scale=20, ef=16, seed=2, interval=1, static_variant="pp", mode=2)
    g = dmf.Graph(4, 0, None, None, None, None)
    cpu = {"value": 900.0, "unit": "ms/batch", "cores": 16, "kind": "port", "sample": "x"}
    cpu1 = dict(cpu, cores=1)
    line = bench.build_line(args, g, 0, 1, 29, S, L, {"data": S, "pp": S, "topo": S}, None, cpu, cpu1, True)
    import json
    line = json.loads(json.dumps(line))
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in line, key
    assert line["value"] == 1.0 and line["learned_speedup_p50"] == 2.0  # nearest-rank p50 of (2, 4) ms
    assert line["atomics_per_batch"] == 1_000_000
    for obj in (line["plain"], line["learned"], *line["plain_variants"].values()):
        for key in ("p50_ms", "p99_ms", "gteps", "atomics_per_batch", "atomic_gops", "roofline_frac"):
            assert key in obj, key
    assert set(line["plain_variants"]) == {"data", "pp", "topo"}
    assert line["cpu_baseline"]["cores"] == 16 and "host_nproc" in line["cpu_baseline"]
    assert line["cpu_baseline_single_thread"]["cores"] == 1
