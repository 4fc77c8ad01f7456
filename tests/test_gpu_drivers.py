"""The reference-compatible drivers (bin/dyn_{data,topo,pp}_bench{,_large}) on
the golden johnson8-2-4 fixtures: the per-file stdout contract, every
'Dynamic flow' equal to the expected value (each file applied to the original
graph, code/README.md:24), and no PANIC line.  --mode=successive is checked
against Dinic on the accumulated updates."""
import os
import re
import shutil
import subprocess

import pytest

import oracle
import dyn_maxflow_amd as dmf
from util import caps_after

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(dmf.__file__), "bin")


def blocks(out):
    """{file: [dynamic flows]} in stdout order."""
    res, cur = {}, None
    for line in out.splitlines():
        m = re.match(r"Processing file: (.*)$", line)
        if m:
            cur = os.path.basename(m.group(1))
            res[cur] = []
        m = re.match(r"Dynamic flow:(-?\d+)$", line)
        if m and cur is not None:
            res[cur].append(int(m.group(1)))
    return res


def stage(tmp_path, golden_dir):
    for f in ("johnson.bin", "johnson_p1.bin", "johnson_p2.bin", "johnson_p3.bin"):
        shutil.copy(os.path.join(golden_dir, f), tmp_path / f)
    return str(tmp_path / "johnson.bin"), str(tmp_path / "johnson")


@pytest.mark.parametrize("variant", ["data", "topo", "pp"])
def test_bench_driver_independent_files(gpu, tmp_path, golden_dir, expected, variant):
    J = expected["johnson"]
    graph, prefix = stage(tmp_path, golden_dir)
    print(f"[driver] dyn_{variant}_bench", flush=True)
    r = subprocess.run([os.path.join(BIN, f"dyn_{variant}_bench"), graph, prefix, str(J["s"]), str(J["t"])],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith(f"source:{J['s']} sink:{J['t']}")
    assert "PANIC" not in r.stdout
    got = blocks(r.stdout)
    want = {b["file"]: b["flow"] for b in J["batches"]}
    assert len(got) == 10  # files _p1 .. _p10; the missing ones run as empty batches
    for name, flows in got.items():
        assert len(flows) >= 2, (name, flows)  # the kernel_cycles sweep
        assert set(flows) == {want.get(name, J["static_flow"])}, (name, flows)
    for line in ("updates size:", "=======kernel cycles", "total GPU Time:", "DYN TIME:", "UPDATE GPU Time:"):
        assert line in r.stdout
    # _large: one file by index
    r = subprocess.run([os.path.join(BIN, f"dyn_{variant}_bench_large"), graph, prefix, str(J["s"]), str(J["t"]), "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = blocks(r.stdout)
    assert list(got) == ["johnson_p2.bin"] and set(got["johnson_p2.bin"]) == {want["johnson_p2.bin"]}


def test_bench_driver_successive_mode(gpu, tmp_path, golden_dir, expected):
    J = expected["johnson"]
    graph, prefix = stage(tmp_path, golden_dir)
    r = subprocess.run([os.path.join(BIN, "dyn_data_bench"), graph, prefix, str(J["s"]), str(J["t"]),
                        "--mode=successive", "--kc=14"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "PANIC" not in r.stdout
    got = blocks(r.stdout)
    g = dmf.read_graph_bin(graph)
    applied = []
    for i in range(1, 4):
        applied.append(dmf.read_updates_bin(f"{prefix}_p{i}.bin"))
        want = oracle.dinic(g, J["s"], J["t"], cap=caps_after(g, applied))[0]
        assert got[f"johnson_p{i}.bin"] == [want], (i, got)
