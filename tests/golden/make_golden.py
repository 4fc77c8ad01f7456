"""Regenerates the golden fixtures in this directory.

Runs in the build container only (needs /root/reference and oracle/_ref):
  * johnson_w.txt       reference weight tool (input_create/weight.cpp) on the
                        reference's sample data/raw_dataset/johnson8-2-4.mtx
  * johnson.bin         reference create_bin (input_create/create_bin.cpp) on it
  * johnson_p{i}.bin    reference create_update_bin on update lists written by
                        this project's seeded generator (nested prefixes, as
                        input_create/cut_p.cpp slices them)
  * rmat8.txt/.bin      a scale-8 R-MAT edge list (this project's generator,
                        with both directions of some pairs and repeated pairs
                        added) and the reference create_bin output for it
  * expected.json       max-flow values and S_min sets from the oracle's Dinic,
                        each cross-checked here against networkx and scipy

The s=25, t=2 pair for johnson8-2-4 is the one input_create/README.md:61-65 gives.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import dyn_maxflow_amd as dmf  # noqa: E402
import oracle  # noqa: E402

REF = "/root/reference"
REF_BIN = os.path.join(ROOT, "oracle", "_ref")


def run(cmd):
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)


def nx_flow(g, s, t, cap):
    import networkx as nx
    G = nx.DiGraph()
    G.add_nodes_from(range(g.n))
    tails = g.tails()
    for e in range(g.m):
        a, b = int(tails[e]), int(g.col[e])
        c = int(cap[e])
        if G.has_edge(a, b):
            G[a][b]["capacity"] += c
        else:
            G.add_edge(a, b, capacity=c)
    val, (S, _) = nx.minimum_cut(G, s, t)
    return val


def scipy_flow(g, s, t, cap):
    import scipy.sparse as sp
    from scipy.sparse.csgraph import maximum_flow
    tails = g.tails()
    M = sp.csr_matrix((cap.astype(np.int32), (tails, g.col)), shape=(g.n, g.n))
    M.sum_duplicates()
    return maximum_flow(M, s, t, method="dinic").flow_value


def main():
    oracle.build_ref()
    mtx = os.path.join(REF, "data", "raw_dataset", "johnson8-2-4.mtx")
    wtxt = os.path.join(HERE, "johnson_w.txt")
    jbin = os.path.join(HERE, "johnson.bin")
    run([os.path.join(REF_BIN, "weight"), mtx, wtxt])
    run([os.path.join(REF_BIN, "create_bin"), wtxt, jbin])
    g = dmf.read_graph_bin(jbin)
    s, t = 25, 2
    exp = {"johnson": {"s": s, "t": t, "n": g.n, "m": g.m}}
    f0, side0 = oracle.dinic(g, s, t)
    assert f0 == nx_flow(g, s, t, g.cap) == scipy_flow(g, s, t, g.cap)
    exp["johnson"]["static_flow"] = f0
    exp["johnson"]["static_smin"] = np.flatnonzero(side0).tolist()
    # update files: nested prefixes of one shuffled list relative to the
    # original capacities (the reference's independent-file semantics)
    lines = np.loadtxt(wtxt, dtype=np.int64).reshape(-1, 3)
    u, v, w = (lines[:, i].astype(np.int32) for i in range(3))
    cur = w.copy()
    B = max(1, len(u) // 100 * 5)  # 5% per step so the tiny graph sees real changes
    src, dst, nw = dmf.gen_updates(u, v, cur, 3 * B, seed=7, apply=False)
    batches = []
    for i in (1, 2, 3):
        k = i * B
        txt = os.path.join(HERE, f"johnson_p{i}.txt")
        with open(txt, "w") as fh:
            fh.write(f"{k}\n")
            for a, b, c in zip(src[:k], dst[:k], nw[:k]):
                fh.write(f"{a} {b} {c}\n")
        pbin = os.path.join(HERE, f"johnson_p{i}.bin")
        run([os.path.join(REF_BIN, "create_update_bin"), txt, pbin])
        os.remove(txt)
        cap = g.cap.copy()
        tails = g.tails()
        for a, b, c in zip(src[:k], dst[:k], nw[:k]):
            idx = np.flatnonzero((tails == a) & (g.col == b))
            cap[idx[-1]] = c
        f, side = oracle.dinic(g, s, t, cap=cap)
        assert f == nx_flow(g, s, t, cap) == scipy_flow(g, s, t, cap)
        batches.append({"file": f"johnson_p{i}.bin", "k": int(k), "flow": f,
                        "smin": np.flatnonzero(side).tolist()})
    exp["johnson"]["batches"] = batches

    # R-MAT fixture for the CSR builder: duplicates and reverse pairs included
    ru, rv, rw = dmf.gen_rmat(8, 8, seed=11)
    extra = [(int(rv[i]), int(ru[i]), int(rw[i]) % 50 + 1) for i in range(0, len(ru), 7)]
    extra += [(int(ru[i]), int(rv[i]), 3) for i in range(0, len(ru), 13)]
    extra += [(5, 5, 9)]  # self-loop, dropped by create_bin
    rtxt = os.path.join(HERE, "rmat8.txt")
    with open(rtxt, "w") as fh:
        for a, b, c in zip(ru, rv, rw):
            fh.write(f"{a} {b} {c}\n")
        for a, b, c in extra:
            fh.write(f"{a} {b} {c}\n")
    rbin = os.path.join(HERE, "rmat8.bin")
    run([os.path.join(REF_BIN, "create_bin"), rtxt, rbin])
    rg = dmf.read_graph_bin(rbin)
    allu = np.concatenate([ru, np.array([e[0] for e in extra], np.int32)])
    allv = np.concatenate([rv, np.array([e[1] for e in extra], np.int32)])
    rs, rt = dmf.pick_source_sink(rg.n, allu, allv)
    rf, rside = oracle.dinic(rg, rs, rt)
    assert rf == nx_flow(rg, rs, rt, rg.cap) == scipy_flow(rg, rs, rt, rg.cap)
    exp["rmat8"] = {"s": rs, "t": rt, "n": rg.n, "m": rg.m, "flow": rf,
                    "smin": np.flatnonzero(rside).tolist()}
    with open(os.path.join(HERE, "expected.json"), "w") as fh:
        json.dump(exp, fh, indent=1)
    print("golden fixtures written")


if __name__ == "__main__":
    main()
