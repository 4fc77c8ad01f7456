"""Graph / update .bin formats and the CSR builder, pinned against the
reference's own tools (input_create/create_bin.cpp, create_update_bin.cpp and
the code/graph.hpp + code/update.hpp readers, compiled into oracle/_ref) and
against fixtures those tools wrote (tests/golden)."""
import os
import subprocess

import numpy as np
import pytest

import dyn_maxflow_amd as dmf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = os.path.join(ROOT, "oracle", "_ref")
HAVE_REF = os.path.exists(os.path.join(REF_BIN, "create_bin"))


def load_txt_edges(path):
    a = np.loadtxt(path, dtype=np.int64).reshape(-1, 3)
    return a[:, 0].astype(np.int32), a[:, 1].astype(np.int32), a[:, 2].astype(np.int32)


def test_read_reference_written_graph(golden_dir, expected):
    g = dmf.read_graph_bin(os.path.join(golden_dir, "johnson.bin"))
    assert (g.n, g.m) == (expected["johnson"]["n"], expected["johnson"]["m"])
    assert g.row[0] == 0 and g.row[-1] == g.m
    assert np.array_equal(g.rev[g.rev], np.arange(g.m))
    tails = g.tails()
    assert np.array_equal(tails[g.rev], g.col)  # rev pairs u->v with v->u


@pytest.mark.parametrize("name", ["johnson", "rmat8"])
def test_builder_matches_reference_create_bin_bytes(golden_dir, tmp_path, name):
    txt = "johnson_w.txt" if name == "johnson" else "rmat8.txt"
    u, v, w = load_txt_edges(os.path.join(golden_dir, txt))
    g = dmf.build_csr(u, v, w)
    out = tmp_path / "mine.bin"
    dmf.write_graph_bin(str(out), g)
    ref = open(os.path.join(golden_dir, f"{name}.bin"), "rb").read()
    assert out.read_bytes() == ref


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference tree absent)")
def test_reader_matches_reference_loader(golden_dir):
    path = os.path.join(golden_dir, "rmat8.bin")
    dump = subprocess.run([os.path.join(REF_BIN, "load_dump"), "graph", path], check=True,
                          capture_output=True, text=True).stdout.split("\n")
    n, m = map(int, dump[0].split())
    g = dmf.read_graph_bin(path)
    assert (g.n, g.m) == (n, m)
    for line, arr in zip(dump[1:5], (g.row, g.col, g.cap, g.rev)):
        assert np.array_equal(np.array(line.split(), dtype=np.int64), arr)


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference tree absent)")
def test_update_reader_matches_reference_loader(golden_dir):
    path = os.path.join(golden_dir, "johnson_p2.bin")
    out = subprocess.run([os.path.join(REF_BIN, "load_dump"), "updates", path], check=True,
                         capture_output=True, text=True).stdout.strip().split("\n")
    size = int(out[0])
    recs = np.array([list(map(int, l.split())) for l in out[1:]], dtype=np.int64)
    src, dst, w = dmf.read_updates_bin(path)
    k = len(src)
    # code/update.hpp:72-83: k zero records, then the real ones minus self-loops
    nself = int((src == dst).sum())
    assert size == 2 * k - nself
    assert (recs[:k] == 0).all()
    real = recs[k:]
    assert (real[:, 0] == ord("a")).all()
    keep = src != dst
    assert np.array_equal(real[:, 1], src[keep])
    assert np.array_equal(real[:, 2], dst[keep])
    assert np.array_equal(real[:, 3], w[keep])


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference tree absent)")
def test_update_writer_matches_reference_create_update_bin(tmp_path):
    rng = np.random.default_rng(5)
    k = 37
    src = rng.integers(0, 50, k).astype(np.int32)
    dst = rng.integers(0, 50, k).astype(np.int32)
    w = rng.integers(0, 101, k).astype(np.int32)
    txt = tmp_path / "u.txt"
    txt.write_text(f"{k}\n" + "".join(f"{a} {b} {c}\n" for a, b, c in zip(src, dst, w)))
    ref_out = tmp_path / "ref.bin"
    subprocess.run([os.path.join(REF_BIN, "create_update_bin"), str(txt), str(ref_out)], check=True)
    mine = tmp_path / "mine.bin"
    dmf.write_updates_bin(str(mine), src, dst, w)
    assert mine.read_bytes() == ref_out.read_bytes()


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference tree absent)")
def test_builder_matches_create_bin_on_random_inputs(tmp_path):
    rng = np.random.default_rng(9)
    for trial in range(5):
        n, k = 40 + trial * 10, 300
        u = rng.integers(0, n, k).astype(np.int32)
        v = rng.integers(0, n, k).astype(np.int32)
        w = rng.integers(1, 100, k).astype(np.int32)
        txt = tmp_path / f"e{trial}.txt"
        txt.write_text("".join(f"{a} {b} {c}\n" for a, b, c in zip(u, v, w)))
        ref = tmp_path / f"r{trial}.bin"
        subprocess.run([os.path.join(REF_BIN, "create_bin"), str(txt), str(ref)], check=True,
                       stdout=subprocess.DEVNULL)
        mine = tmp_path / f"m{trial}.bin"
        dmf.write_graph_bin(str(mine), dmf.build_csr(u, v, w))
        assert mine.read_bytes() == ref.read_bytes()


def test_update_roundtrip(tmp_path):
    src = np.array([1, 2, 3, 3], np.int32)
    dst = np.array([2, 3, 1, 3], np.int32)
    w = np.array([5, 0, 100, 7], np.int32)
    p = tmp_path / "u.bin"
    dmf.write_updates_bin(str(p), src, dst, w)
    assert p.stat().st_size == 4 * (1 + 3 * 4)
    a, b, c = dmf.read_updates_bin(str(p))
    assert np.array_equal(a, src) and np.array_equal(b, dst) and np.array_equal(c, w)


def test_graph_roundtrip(tmp_path):
    u, v, w = dmf.gen_rmat(9, 8, 4)
    g = dmf.build_csr(u, v, w)
    p = tmp_path / "g.bin"
    dmf.write_graph_bin(str(p), g)
    h = dmf.read_graph_bin(str(p))
    for a in ("row", "col", "cap", "rev"):
        assert np.array_equal(getattr(g, a), getattr(h, a))


def test_malformed_files_are_reported(tmp_path, golden_dir):
    with pytest.raises(dmf.DmfError) as e:
        dmf.read_graph_bin(str(tmp_path / "absent.bin"))
    assert e.value.status == dmf.EIO
    good = open(os.path.join(golden_dir, "johnson.bin"), "rb").read()
    (tmp_path / "short.bin").write_bytes(good[:-4])
    with pytest.raises(dmf.DmfError) as e:
        dmf.read_graph_bin(str(tmp_path / "short.bin"))
    assert e.value.status == dmf.EIO
    neg = np.array([-3, 5], np.int32).tobytes()
    (tmp_path / "neg.bin").write_bytes(neg)
    with pytest.raises(dmf.DmfError) as e:
        dmf.read_graph_bin(str(tmp_path / "neg.bin"))
    assert e.value.status == dmf.EFORMAT
    arr = np.frombuffer(good, dtype=np.int32).copy()
    n, m = arr[0], arr[1]
    arr[2 + (n + 1) + 3 * m - 1] = (arr[2 + (n + 1) + 3 * m - 1] + 1) % m  # corrupt one rev entry
    (tmp_path / "rev.bin").write_bytes(arr.tobytes())
    with pytest.raises(dmf.DmfError) as e:
        dmf.read_graph_bin(str(tmp_path / "rev.bin"))
    assert e.value.status == dmf.EFORMAT
    (tmp_path / "ushort.bin").write_bytes(np.array([5, 1, 2], np.int32).tobytes())
    with pytest.raises(dmf.DmfError) as e:
        dmf.read_updates_bin(str(tmp_path / "ushort.bin"))
    assert e.value.status == dmf.EIO


def test_csr_check_rejects_each_malformation(golden_dir):
    """Every golden graph passes; each single corruption that would send a
    kernel out of range, or pair an arc with the wrong reverse, is EFORMAT."""
    for name in ("johnson.bin", "rmat8.bin"):  # graphs written by the reference's create_bin
        dmf.csr_check(dmf.read_graph_bin(os.path.join(golden_dir, name)))
    g = dmf.build_csr([0, 1, 2, 0], [1, 2, 3, 2], [5, 4, 3, 2])
    dmf.csr_check(g)
    tails = g.tails()
    # an arc whose reverse is itself a valid involution but heads to the wrong tail
    e = int(np.flatnonzero(tails == 0)[0])
    other = int(np.flatnonzero((tails != 0) & (g.col != 0))[0])
    swap = g.rev.copy()
    a, b = e, other
    ra, rb = swap[a], swap[b]
    swap[a], swap[b], swap[ra], swap[rb] = rb, ra, b, a  # still rev[rev[x]] == x
    assert np.array_equal(swap[swap], np.arange(g.m))
    bad = {
        "row start": ("row", 0, 1),
        "row end": ("row", g.n, g.m - 1),
        "row order": ("row", 2, int(g.row[3]) + 1),
        "head range": ("col", 1, g.n),
        "negative head": ("col", 1, -1),
        "negative cap": ("cap", 0, -1),
        "rev range": ("rev", 0, g.m),
        "rev involution": ("rev", 0, (int(g.rev[0]) + 1) % g.m),
    }
    for what, (field, i, val) in bad.items():
        h = dmf.Graph(g.n, g.m, g.row.copy(), g.col.copy(), g.cap.copy(), g.rev.copy())
        getattr(h, field)[i] = val
        with pytest.raises(dmf.DmfError) as err:
            dmf.csr_check(h)
        assert err.value.status == dmf.EFORMAT, what
    h = dmf.Graph(g.n, g.m, g.row.copy(), g.col.copy(), g.cap.copy(), swap)
    with pytest.raises(dmf.DmfError) as err:
        dmf.csr_check(h)
    assert err.value.status == dmf.EFORMAT
    with pytest.raises(dmf.DmfError) as err:
        dmf.csr_check(dmf.Graph(g.n, g.m, g.row[:-1], g.col, g.cap, g.rev))
    assert err.value.status == dmf.EINVAL


def test_csr_check_threaded_ranges():
    """Above 2^22 arcs the arc pass is split over threads: a corruption in the
    first, a middle and the last arc is found whichever range holds it."""
    u, v, w = dmf.gen_rmat(18, 16, 3)
    g = dmf.build_csr(u, v, w)
    assert g.m >= 1 << 22
    dmf.csr_check(g)
    for e in (0, g.m // 2 + 1, g.m - 1):
        rev = g.rev.copy()
        rev[e] = (rev[e] + 1) % g.m
        with pytest.raises(dmf.DmfError) as err:
            dmf.csr_check(dmf.Graph(g.n, g.m, g.row, g.col, g.cap, rev))
        assert err.value.status == dmf.EFORMAT


def test_engine_rejects_malformed_csr_before_any_device_call():
    """dmf_engine_create, dmf_batch_create and the recalculate entry point run
    the CSR check first, so a bad graph is EFORMAT with or without a GPU."""
    g = dmf.build_csr([0, 1, 2, 0], [1, 2, 3, 2], [5, 4, 3, 2])
    g.col[1] = g.n + 5
    with pytest.raises(dmf.DmfError) as err:
        dmf.Engine(g, 0, 3)
    assert err.value.status == dmf.EFORMAT
    with pytest.raises(dmf.DmfError) as err:
        dmf.BatchEngine([g], [0], [3])
    assert err.value.status == dmf.EFORMAT
    with pytest.raises(dmf.DmfError) as err:
        dmf.recalculate_max_flow(g, [0], [1], [7], 0, 3, kc=1)
    assert err.value.status == dmf.EFORMAT


def test_rmat_generator_properties():
    u, v, w = dmf.gen_rmat(10, 16, 1)
    u2, v2, w2 = dmf.gen_rmat(10, 16, 1)
    assert np.array_equal(u, u2) and np.array_equal(v, v2) and np.array_equal(w, w2)
    u3, _, _ = dmf.gen_rmat(10, 16, 2)
    assert not np.array_equal(u[: len(u3)], u3[: len(u)])
    assert (u != v).all()
    keys = u.astype(np.int64) * (1 << 20) + v
    assert len(np.unique(keys)) == len(keys)
    assert len(u) <= (1 << 10) * 16
    assert w.min() >= 1 and w.max() <= 100
    # a = 0.57 skews both ends to low ids: vertex 0 is the heaviest tail
    outd = np.bincount(u, minlength=1024)
    assert outd.argmax() == 0


def test_grid_generator_counts():
    rows, cols = 5, 7
    u, v, w = dmf.gen_grid(rows, cols, 3, 1, 1000)
    assert len(u) == 2 * (rows * (cols - 1) + cols * (rows - 1))
    g = dmf.build_csr(u, v, w)
    assert g.n == rows * cols
    assert g.m == len(u)  # both directions are input edges, so no zero-capacity reverses
    d = np.abs(u.astype(int) - v.astype(int))
    assert set(d.tolist()) <= {1, cols}


def test_pick_source_sink_rules():
    u = np.array([3, 3, 1, 1, 0, 2], np.int32)
    v = np.array([0, 2, 0, 2, 2, 0], np.int32)
    s, t = dmf.pick_source_sink(4, u, v)
    assert (s, t) == (1, 0)  # out-degree tie 1 vs 3 -> lowest id; in-degree tie 0 vs 2 -> 0
    u = np.array([0, 0, 1], np.int32)
    v = np.array([1, 2, 0], np.int32)
    s, t = dmf.pick_source_sink(3, u, v)
    assert s == 0 and t == 1  # t != s even though... in-degree ties broken by lowest id


def test_update_generator_rules():
    u, v, w = dmf.gen_rmat(9, 8, 2)
    cur = w.copy()
    src, dst, nw = dmf.gen_updates(u, v, cur, 200, seed=3, inc_frac=0.5, apply=True)
    assert len(src) == 200
    keys = src.astype(np.int64) * (1 << 20) + dst
    assert len(np.unique(keys)) == len(keys)  # distinct edges
    look = {(a, b): i for i, (a, b) in enumerate(zip(u.tolist(), v.tolist()))}
    inc = dec = 0
    for a, b, c in zip(src.tolist(), dst.tolist(), nw.tolist()):
        i = look[(a, b)]
        old = w[i]
        assert 1 <= c <= 100 and c != old
        inc += c > old
        dec += c < old
        assert cur[i] == c
    assert inc > 50 and dec > 50
    # decreases only
    cur2 = w.copy()
    _, _, nw2 = dmf.gen_updates(u, v, cur2, 100, seed=4, inc_frac=0.0, apply=False)
    assert np.array_equal(cur2, w)
