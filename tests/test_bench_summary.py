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
