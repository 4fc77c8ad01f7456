"""ctypes wrapper of oracle/liboracle.so — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only to compute the expected results the engine is compared
with or as the timed CPU baseline.  The engine in dyn_maxflow_amd/ never
imports it.

See oracle.c for what each function restates (reference file:line) and how it
is pinned.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_U8P = C.POINTER(C.c_uint8)
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def build_ref() -> None:
    """Compile the reference's own input tools into oracle/_ref (needs /root/reference)."""
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_dinic.restype = C.c_int
        L.orc_dinic.argtypes = [C.c_int32, C.c_int32, _I32P, _I32P, _I32P, _I32P, C.c_int32,
                                C.c_int32, _I64P, _U8P]
        L.orc_pr_create.restype = C.c_void_p
        L.orc_pr_create.argtypes = [C.c_int32, C.c_int32, _I32P, _I32P, _I32P, _I32P, C.c_int32,
                                    C.c_int32]
        L.orc_pr_destroy.argtypes = [C.c_void_p]
        L.orc_pr_set_threads.argtypes = [C.c_void_p, C.c_int32]
        L.orc_pr_set_max_rounds.argtypes = [C.c_void_p, C.c_int32]
        L.orc_pr_reset.argtypes = [C.c_void_p]
        L.orc_pr_static.restype = C.c_int
        L.orc_pr_static.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.orc_pr_apply.restype = C.c_int
        L.orc_pr_apply.argtypes = [C.c_void_p, C.c_int32, _I32P, _I32P, _I32P, C.c_int32, C.c_int32]
        L.orc_pr_set_caps.restype = C.c_int
        L.orc_pr_set_caps.argtypes = [C.c_void_p, C.c_int32, _I32P, _I32P, _I32P]
        L.orc_pr_flow.restype = C.c_int64
        L.orc_pr_flow.argtypes = [C.c_void_p]
        L.orc_pr_excess_t.restype = C.c_int64
        L.orc_pr_excess_t.argtypes = [C.c_void_p]
        L.orc_pr_min_cut.argtypes = [C.c_void_p, _U8P]
        L.orc_pr_read_state.argtypes = [C.c_void_p, _I32P, _I32P, _I32P, _I64P, _I32P]
        L.orc_pr_counters.argtypes = [C.c_void_p, _I64P]
        L.orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a):
    return a.ctypes.data_as(_I32P)


def dinic(g, s: int, t: int, cap=None):
    """(max-flow value, S_min as uint8[n]) by int64 Dinic on the CSR g."""
    row, col, rev = _i32(g.row), _i32(g.col), _i32(g.rev)
    capa = _i32(g.cap if cap is None else cap)
    side = np.zeros(g.n, dtype=np.uint8)
    f = C.c_int64()
    rc = lib().orc_dinic(g.n, g.m, _p(row), _p(col), _p(capa), _p(rev), s, t, C.byref(f),
                         side.ctypes.data_as(_U8P))
    if rc:
        raise ValueError(f"orc_dinic failed ({rc})")
    return f.value, side


VARIANTS = {"data": 0, "topo": 1, "pp": 2}


class PushRelabel:
    """CPU restatement of the reference push-relabel with its update repair."""

    def __init__(self, g, s: int, t: int, threads: int = 1):
        self.g = g
        self._keep = [_i32(g.row), _i32(g.col), _i32(g.cap), _i32(g.rev)]
        row, col, cap, rev = self._keep
        self.h = lib().orc_pr_create(g.n, g.m, _p(row), _p(col), _p(cap), _p(rev), s, t)
        if not self.h:
            raise ValueError("orc_pr_create failed")
        lib().orc_pr_set_threads(self.h, threads)

    def close(self):
        if self.h:
            lib().orc_pr_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_max_rounds(self, r: int):
        lib().orc_pr_set_max_rounds(self.h, r)

    def reset(self):
        lib().orc_pr_reset(self.h)

    def static(self, variant="data", kc=1) -> int:
        rc = lib().orc_pr_static(self.h, VARIANTS[variant], kc)
        if rc:
            raise RuntimeError(f"orc_pr_static rc={rc}")
        return self.flow()

    def apply(self, src, dst, w, variant="data", kc=1) -> int:
        src, dst, w = _i32(src), _i32(dst), _i32(w)
        rc = lib().orc_pr_apply(self.h, len(src), _p(src), _p(dst), _p(w), VARIANTS[variant], kc)
        if rc:
            raise RuntimeError(f"orc_pr_apply rc={rc}")
        return self.flow()

    def set_caps(self, src, dst, w):
        src, dst, w = _i32(src), _i32(dst), _i32(w)
        rc = lib().orc_pr_set_caps(self.h, len(src), _p(src), _p(dst), _p(w))
        if rc:
            raise RuntimeError(f"orc_pr_set_caps rc={rc}")

    def flow(self) -> int:
        return int(lib().orc_pr_flow(self.h))

    def excess_t(self) -> int:
        return int(lib().orc_pr_excess_t(self.h))

    def min_cut(self) -> np.ndarray:
        side = np.zeros(self.g.n, dtype=np.uint8)
        lib().orc_pr_min_cut(self.h, side.ctypes.data_as(_U8P))
        return side

    def state(self):
        g = self.g
        res, rres, cap = (np.zeros(g.m, np.int32) for _ in range(3))
        ex = np.zeros(g.n, np.int64)
        h = np.zeros(g.n, np.int32)
        lib().orc_pr_read_state(self.h, _p(res), _p(rres), _p(cap), ex.ctypes.data_as(_I64P), _p(h))
        return res, rres, cap, ex, h

    def counters(self) -> dict:
        out = np.zeros(6, np.int64)
        lib().orc_pr_counters(self.h, out.ctypes.data_as(_I64P))
        return dict(zip(["rounds", "pushes", "relabels", "repairs", "bfs_levels", "arc_visits"],
                        out.tolist()))


def max_threads() -> int:
    return int(lib().orc_max_threads())


def check_state(g, res, cap, excess, s=None):
    """Residual-state invariants: 0 <= res, pair consistency res[e]+res[rev]=cap[e]+cap[rev],
    excess[v] = net inflow = sum over out-arcs of (res - cap) (exact int64, as differences of
    a running sum over the CSR).  Returns a list of violations."""
    bad = []
    if (res < 0).any():
        bad.append(f"negative residual on {(res < 0).sum()} arcs")
    # res[e] + res[rev] == cap[e] + cap[rev]  <=>  d[e] == -d[rev] with d = res - cap
    d = res.astype(np.int64) - cap.astype(np.int64)
    if not np.array_equal(d, -d[g.rev]):
        bad.append("arc pair residuals do not sum to the pair capacity")
    csum = np.concatenate([[0], np.cumsum(d)])
    row = g.row.astype(np.int64)
    if not np.array_equal(csum[row[1:]] - csum[row[:-1]], excess.astype(np.int64)):
        bad.append("excess differs from net inflow")
    return bad


def check_flow(g, res, cap, excess, sources, sinks, flow):
    """A flow, not a preflow: check_state, plus conservation (excess 0 at every
    vertex that is neither a source nor a sink) and excess of the sinks summing
    to the flow value, of the sources to minus it.  Returns a list of violations."""
    bad = check_state(g, res, cap, excess)
    ex = excess.astype(np.int64)
    term = np.zeros(g.n, bool)
    term[np.asarray(sources)] = True
    term[np.asarray(sinks)] = True
    off = np.flatnonzero(~term & (ex != 0))
    if len(off):
        bad.append(f"conservation fails at {len(off)} vertices (first {off[:5].tolist()})")
    if int(ex[np.asarray(sinks)].sum()) != flow:
        bad.append(f"sink excess {int(ex[np.asarray(sinks)].sum())} != flow {flow}")
    if int(ex[np.asarray(sources)].sum()) != -flow:
        bad.append(f"source excess {int(ex[np.asarray(sources)].sum())} != -flow {-flow}")
    return bad


def cut_capacity(g, cap, side):
    """Capacity of the cut (S, V \\ S) with S = {v : side[v] != 0}: sum of cap over
    arcs leaving S.  Equal to the value of a feasible flow, it certifies that
    flow maximum and the cut minimum (max-flow = min-cut)."""
    inS = np.asarray(side).astype(bool)
    tails = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.row))
    leave = inS[tails] & ~inS[g.col]
    return int(cap.astype(np.int64)[leave].sum())
