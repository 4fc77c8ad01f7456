"""Scheduler training on synthetic traces (CPU): which schedule teaches each
batch, the variant restriction, and the weight file the engine compiles in."""
import json
import os

import numpy as np

from dyn_maxflow_amd import sched_train as st

# two fixed schedules over three batches; per batch the fastest is A, B, A,
# over the whole workload A (12 ms against 13 ms)
A = (0, 1, 0)  # data, relabel every round, no lazy start
B = (2, 2, 0)  # pp, relabel every 2 rounds
TIMES = {A: [1.0, 10.0, 1.0], B: [5.0, 2.0, 6.0]}


def _write_trace(path):
    solve = 0
    with open(path, "w") as fh:
        for key, times in TIMES.items():
            for b, ms in enumerate(times):
                solve += 1
                variant, interval, lazy = key
                feat = [float(b), float(variant), 0, 0, 0.5, 0.5, 0, 1.0]
                fh.write(json.dumps({"solve": solve, "what": "batch", "round": 0, "variant": variant,
                                     "relabelled": 1, "feat": feat}) + "\n")
                for rnd in (1, 2):
                    fh.write(json.dumps({"solve": solve, "what": "round", "round": rnd, "variant": variant,
                                         "relabelled": int(rnd % interval == 0), "feat": feat}) + "\n")
                fh.write(json.dumps({"solve": solve, "what": "summary", "entry": 1, "scheduler": 0,
                                     "variant": variant, "interval": interval, "lazy": lazy,
                                     "device_ms": ms}) + "\n")


def test_teacher_per_batch_and_per_workload(tmp_path):
    _write_trace(tmp_path / "w.jsonl")
    data = st._load(str(tmp_path))
    _, _, _, yv = st._examples(data, "batch")
    assert yv.tolist() == [A[0], B[0], A[0]]
    _, _, _, yv = st._examples(data, "workload")
    assert yv.tolist() == [A[0], A[0], A[0]]


def test_teacher_variant_restriction(tmp_path):
    _write_trace(tmp_path / "w.jsonl")
    data = st._load(str(tmp_path))
    _, _, _, yv = st._examples(data, "batch", variants=(2,))
    assert yv.tolist() == [B[0]] * 3
    # relabel examples: one at batch start plus one per round of each teacher run
    Xr, yr, _, _ = st._examples(data, "batch", variants=(0,))
    assert len(Xr) == 3 * (1 + 2) and yr.tolist() == [1, 1, 1] * 3


def test_weights_file_layout(tmp_path, monkeypatch):
    rng = np.random.default_rng(0)
    P = {"w1": rng.normal(size=(st.FEAT, st.HID)), "b1": rng.normal(size=st.HID),
         "w2": rng.normal(size=(st.HID, st.HID)), "b2": rng.normal(size=st.HID),
         "w3": rng.normal(size=(st.HID, st.OUT)), "b3": rng.normal(size=st.OUT)}
    out = tmp_path / "w.inc"
    monkeypatch.setattr(st, "WEIGHTS", str(out))
    st.write_weights(P, "test")
    txt = out.read_text()
    body = txt[txt.index("{") + 1:txt.rindex("}")]
    vals = [float(x.strip().rstrip("f")) for x in body.split(",") if x.strip()]
    n = st.FEAT * st.HID + st.HID + st.HID * st.HID + st.HID + st.HID * st.OUT + st.OUT
    assert len(vals) == n == 1476
    flat = np.concatenate([P[k].ravel() for k in ("w1", "b1", "w2", "b2", "w3", "b3")])
    assert np.allclose(vals, flat.astype(np.float32))


def test_shipped_weights_parse():
    path = os.path.join(os.path.dirname(st.__file__), "csrc", "sched_weights.inc")
    txt = open(path).read()
    assert "#define DMF_SCHED_TRAINED 1" in txt
    body = txt[txt.index("{") + 1:txt.rindex("}")]
    vals = np.array([float(x.strip().rstrip("f")) for x in body.split(",") if x.strip()])
    assert len(vals) == 1476 and np.all(np.isfinite(vals))


def test_graph_workload_from_bin(tmp_path, golden_dir):
    """A user's .bin graph becomes a training workload: its positive-capacity
    arcs, unchanged, as the edges batches draw from."""
    import dyn_maxflow_amd as dmf
    path = os.path.join(golden_dir, "johnson.bin")
    name, u, v, w, lo, hi = st.graph_workload(path)
    g = dmf.read_graph_bin(path)
    assert name == "johnson" and lo == 1 and hi == int(g.cap.max())
    assert len(u) == int((g.cap > 0).sum()) and np.all(w > 0)
    g2 = dmf.build_csr(u, v, w)
    assert g2.n == g.n and int(g2.cap.sum()) == int(g.cap.sum())


def test_variant_teacher_per_workload(tmp_path):
    """Variant from the workload's best schedule (A: data) for every batch;
    relabel examples from the fastest run of that variant on each batch."""
    _write_trace(tmp_path / "w.jsonl")
    data = st._load(str(tmp_path))
    Xr, yr, Xv, yv = st._examples(data, "batch", variant_teacher="workload")
    assert yv.tolist() == [A[0]] * 3
    assert len(Xr) == 3 * 3 and yr.tolist() == [1, 1, 1] * 3  # A relabels every round


def test_retrofit_sets_graph_size_at_batch_start(tmp_path):
    src = tmp_path / "src"
    src.mkdir()
    _write_trace(src / "grid64s9.jsonl")  # a 64x64 grid: n = 4096
    st.retrofit(str(src), str(tmp_path / "out"))
    rows = [json.loads(l) for l in open(tmp_path / "out" / "grid64s9.jsonl")]
    orig = [json.loads(l) for l in open(src / "grid64s9.jsonl")]
    want = round(float(np.float32(np.log2(4097.0)) / np.float32(32.0)), 6)
    assert len(rows) == len(orig)
    for r, o in zip(rows, orig):
        if r["what"] == "batch":
            assert r["feat"][0] == want and r["feat"][1:] == o["feat"][1:]
        else:
            assert r == o  # round and summary rows are unchanged


def test_committed_examples_rebuild_the_shipped_weights(tmp_path):
    """The shipped weights (csrc/sched_weights.inc) are the fit of the committed
    example set (sched_examples.npz, from the engine's own traces): refitting
    from that file alone writes the same file byte for byte."""
    from dyn_maxflow_amd import sched_train as st
    out = tmp_path / "w.inc"
    st.train([], examples=st.EXAMPLES, out=str(out))
    assert out.read_bytes() == open(st.WEIGHTS, "rb").read()
