"""Input tool (dmf_input) and bench-driver host logic that needs no GPU."""
import os
import subprocess
import sys

import numpy as np
import pytest

import dyn_maxflow_amd as dmf

BIN = os.path.join(os.path.dirname(dmf.__file__), "bin")
REF_BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")


def tool(*args, **kw):
    return subprocess.run([os.path.join(BIN, "dmf_input"), *map(str, args)], capture_output=True,
                          text=True, **kw)


def test_rmat_command_writes_reference_layouts(tmp_path):
    pre = tmp_path / "g"
    r = tool("rmat", 9, 8, 3, pre)
    assert r.returncode == 0, r.stderr
    s, t = map(int, r.stdout.split())
    u, v, w = dmf.gen_rmat(9, 8, 3)
    g = dmf.read_graph_bin(str(pre) + ".bin")
    h = dmf.build_csr(u, v, w)
    assert np.array_equal(g.col, h.col) and np.array_equal(g.rev, h.rev)
    assert (s, t) == dmf.pick_source_sink(g.n, u, v)
    txt = np.loadtxt(str(pre) + ".txt", dtype=np.int64)
    assert np.array_equal(txt[:, 0], u) and np.array_equal(txt[:, 2], w)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_BIN, "create_bin")),
                    reason="oracle/_ref not built")
def test_csr_and_upd2bin_commands_match_reference_tools(tmp_path, golden_dir):
    out = tmp_path / "j.bin"
    assert tool("csr", os.path.join(golden_dir, "johnson_w.txt"), out).returncode == 0
    assert out.read_bytes() == open(os.path.join(golden_dir, "johnson.bin"), "rb").read()
    upd = tmp_path / "u.txt"
    upd.write_text("3\n1 2 5\n2 3 0\n4 1 9\n")
    mine = tmp_path / "u.bin"
    ref = tmp_path / "r.bin"
    assert tool("upd2bin", upd, mine).returncode == 0
    subprocess.run([os.path.join(REF_BIN, "create_update_bin"), str(upd), str(ref)], check=True)
    assert mine.read_bytes() == ref.read_bytes()


def test_update_files_are_nested_prefixes(tmp_path, golden_dir):
    pre = tmp_path / "upd"
    r = tool("updates", os.path.join(golden_dir, "johnson_w.txt"), pre, 5, 50, 11)
    assert r.returncode == 0, r.stderr
    prev = None
    for i in range(1, 11):
        a, b, c = dmf.read_updates_bin(f"{pre}_p{i}.bin")
        assert len(a) == i * (210 * 5 // 100)
        if prev is not None:  # cut_p.cpp:32-38: file i is a prefix of file i+1
            assert np.array_equal(a[: len(prev[0])], prev[0])
            assert np.array_equal(c[: len(prev[2])], prev[2])
        prev = (a, b, c)
        txt = open(f"{pre}_p{i}.txt").read().split("\n")
        assert int(txt[0]) == len(a)


def test_successive_batches_grow_by_one_percent(tmp_path):
    pre = tmp_path / "g"
    tool("rmat", 10, 8, 2, pre)
    r = tool("successive", str(pre) + ".txt", tmp_path / "b", 50, 4)
    assert r.returncode == 0, r.stderr
    m_in = len(np.loadtxt(str(pre) + ".txt", dtype=np.int64))
    for i in range(1, 11):
        a, _, _ = dmf.read_updates_bin(str(tmp_path / f"b_b{i}.bin"))
        assert len(a) == m_in * i // 100


def test_weight_command_is_seeded(tmp_path, golden_dir):
    mtx = tmp_path / "e.mtx"
    mtx.write_text("%%header\n1 2\n2 3\n3 1\n")
    o1, o2 = tmp_path / "a.txt", tmp_path / "b.txt"
    assert tool("weight", mtx, o1, 5).returncode == 0
    assert tool("weight", mtx, o2, 5).returncode == 0
    assert o1.read_text() == o2.read_text()
    rows = [list(map(int, l.split())) for l in o1.read_text().strip().split("\n")]
    assert [r[:2] for r in rows] == [[1, 2], [2, 3], [3, 1]]
    assert all(1 <= r[2] <= 100 for r in rows)


@pytest.mark.parametrize("name", ["dyn_data_bench", "dyn_topo_bench", "dyn_pp_bench",
                                  "dyn_data_bench_large", "dyn_topo_bench_large", "dyn_pp_bench_large"])
def test_driver_argument_errors(name, tmp_path):
    exe = os.path.join(BIN, name)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    args = [str(tmp_path / "absent.bin"), "pre", "0", "1"] + (["1"] if name.endswith("large") else [])
    r = subprocess.run([exe, *args], capture_output=True, text=True)
    assert r.stdout.startswith("source:0 sink:1")
    assert r.returncode != 0 and "i/o error" in r.stderr
    r = subprocess.run([exe, *args, "--bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "unknown flag" in r.stderr


def test_rmat_n_matches_power_of_two_and_stays_in_range():
    u, v, w = dmf.gen_rmat(10, 8, 5)
    u2, v2, w2 = dmf.gen_rmat_n(1024, 1024 * 8, 5)
    assert np.array_equal(u, u2) and np.array_equal(v, v2) and np.array_equal(w, w2)
    for n in (1025, 1500, 3000, 40000):
        u, v, w = dmf.gen_rmat_n(n, n * 12, 7)
        assert len(u) > n  # enough arcs survive the range filter
        assert u.max() < n and v.max() < n and u.min() >= 0 and v.min() >= 0
        assert not np.any(u == v)
        keys = u.astype(np.int64) * n + v
        assert len(np.unique(keys)) == len(keys)
        assert w.min() >= 1 and w.max() <= 100
    with pytest.raises(dmf.DmfError):
        dmf.gen_rmat_n(1, 10, 1)


@pytest.mark.parametrize("name,bad", [("dmf_multi_bench", "--instances=0"), ("dmf_serve_bench", "--kmax=1"),
                                      ("dmf_multi_bench", "--bogus"), ("dmf_serve_bench", "--bogus")])
def test_workload_driver_argument_errors(name, bad):
    r = subprocess.run([os.path.join(BIN, name), bad], capture_output=True, text=True)
    assert r.returncode == 2 and ("invalid" in r.stderr or "unknown argument" in r.stderr)


def test_workload_drivers_refuse_without_device():
    if dmf.device_count() > 0:
        pytest.skip("a device is visible")
    for name in ("dmf_multi_bench", "dmf_serve_bench"):
        r = subprocess.run([os.path.join(BIN, name)], capture_output=True, text=True)
        assert r.returncode == 2 and "no GPU" in r.stderr
    r = subprocess.run([sys.executable, "-m", "dyn_maxflow_amd.grid_bench", "--side", "8"], capture_output=True,
                       text=True, cwd=os.path.dirname(os.path.dirname(BIN)))
    assert r.returncode != 0 and "no HIP device" in r.stderr


def test_tree_command_writes_the_update_type_layout(tmp_path, golden_dir):
    """dmf_input tree writes input_create's layout (create_txt_updates.sh:17-49):
    <dir>/{dec,inc,mixed}/<prefix>_p1..p10, nested prefixes, with only
    decreases, only increases, and half of each."""
    g = dmf.read_graph_bin(os.path.join(golden_dir, "johnson.bin"))
    edges = np.loadtxt(os.path.join(golden_dir, "johnson_w.txt"), dtype=np.int64)
    cap = {(int(a), int(b)): int(c) for a, b, c in edges}
    r = tool("tree", os.path.join(golden_dir, "johnson_w.txt"), tmp_path / "updates", "johnson", 5, 7)
    assert r.returncode == 0, r.stderr
    for kind in ("dec", "inc", "mixed"):
        prev = None
        ups = 0
        for i in range(1, 11):
            a, b, c = dmf.read_updates_bin(str(tmp_path / "updates" / kind / f"johnson_p{i}.bin"))
            assert len(a) == i * (210 * 5 // 100)
            if prev is not None:
                assert np.array_equal(a[: len(prev)], prev)
            prev = a
            if i == 10:
                ups = sum(int(w) > cap[(int(x), int(y))] for x, y, w in zip(a, b, c))
                if kind == "dec":
                    assert ups == 0
                elif kind == "inc":
                    assert ups == len(a)
                else:
                    assert 0 < ups < len(a)
    assert g.n > 0
