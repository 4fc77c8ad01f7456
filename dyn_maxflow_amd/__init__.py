"""Python mirror of the dynamic max-flow engine's C-ABI (include/dmf.h).

The product is libdmf.so (host C++ + HIP kernels for gfx950).  This module only
loads it with ctypes and converts numpy arrays; it never computes a flow itself.
Loading fails loudly if the library has not been built.

Reference interface this mirrors (dyn_maxflow has no library API of its own):
  graph.hpp:202-223 loadGraph            -> read_graph_bin
  update.hpp:52-89 loadUpdateFile        -> read_updates_bin
  create_bin.cpp:18-124                  -> build_csr / write_graph_bin
  create_update_bin.cpp:11-52            -> write_updates_bin
  dyn_*_bench.cu recalculate_max_flow    -> recalculate_max_flow / Engine
  dyn_*_bench.cu static_recalculate      -> static_recalculate
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdmf.so")

OK = 0
EINVAL = -1
EIO = -2
ENOARC = -3
ENOMEM = -4
EDEVICE = -5
ENOCONV = -6
EFORMAT = -7
EADMIT = -8
EAGAIN = -9
ESIZE = -10

VARIANT_DATA = 0
VARIANT_TOPO = 1
VARIANT_PP = 2
VARIANTS = {"data": VARIANT_DATA, "topo": VARIANT_TOPO, "pp": VARIANT_PP}

SCHED_FIXED = 0
SCHED_LEARNED = 1


class DmfError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = _lib().dmf_strerror(status).decode() if _LIB is not None else str(status)
        super().__init__(f"{what}: {msg} ({status})" if what else f"{msg} ({status})")


class _Csr(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("m", C.c_int32),
        ("row", C.POINTER(C.c_int32)),
        ("col", C.POINTER(C.c_int32)),
        ("cap", C.POINTER(C.c_int32)),
        ("rev", C.POINTER(C.c_int32)),
    ]


class Params(C.Structure):
    """dmf_params: variant, kc, relabel_interval, scheduler, max_rounds, mode, lazy_start, cascade."""

    _fields_ = [
        ("variant", C.c_int32),
        ("kc", C.c_int32),
        ("relabel_interval", C.c_int32),
        ("scheduler", C.c_int32),
        ("max_rounds", C.c_int32),
        ("mode", C.c_int32),
        ("lazy_start", C.c_int32),
        ("cascade", C.c_int32),
    ]


class Counters(C.Structure):
    _fields_ = [
        ("arc_visits", C.c_uint64),
        ("vertex_visits", C.c_uint64),
        ("pushes", C.c_uint64),
        ("relabels", C.c_uint64),
        ("repairs", C.c_uint64),
        ("bfs_levels", C.c_uint64),
        ("global_relabels", C.c_uint64),
        ("rounds", C.c_uint64),
        ("updates", C.c_uint64),
        ("atomics", C.c_uint64),
        ("launches", C.c_uint64),
        ("device_ms", C.c_double),
        ("bfs_arc_visits", C.c_uint64),
        ("bfs_vertex_visits", C.c_uint64),
        ("bfs_level_launches", C.c_uint64),
        ("bfs_level_ms", C.c_double),
        ("variant", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerConfig(C.Structure):
    """dmf_server_config."""

    _fields_ = [
        ("hbm_budget", C.c_int64),
        ("staging_bytes", C.c_int64),
        ("max_sessions", C.c_int32),
        ("pad", C.c_int32),
        ("params", Params),
    ]


class Result(C.Structure):
    """dmf_result: one served batch."""

    _fields_ = [
        ("ticket", C.c_uint64),
        ("session", C.c_int32),
        ("status", C.c_int32),
        ("flow", C.c_int64),
        ("k", C.c_int32),
        ("pad", C.c_int32),
        ("submit_ms", C.c_double),
        ("admit_ms", C.c_double),
        ("start_ms", C.c_double),
        ("done_ms", C.c_double),
        ("device_ms", C.c_double),
        ("arc_visits", C.c_uint64),
        ("rounds", C.c_uint64),
        ("global_relabels", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad"}


class ServerStats(C.Structure):
    _fields_ = [
        ("sessions", C.c_int32),
        ("in_flight", C.c_int32),
        ("queued", C.c_int32),
        ("pad", C.c_int32),
        ("completed", C.c_uint64),
        ("hbm_budget", C.c_int64),
        ("hbm_held", C.c_int64),
        ("staged_bytes", C.c_int64),
        ("peak_staged_bytes", C.c_int64),
        ("clock_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad"}


_LIB = None
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)


def _lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m dyn_maxflow_amd.build` "
            "(or __graft_entry__.build()); there is no fallback path")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "dmf_strerror": (C.c_char_p, [C.c_int]),
        "dmf_free": (None, [C.c_void_p]),
        "dmf_csr_free": (None, [C.POINTER(_Csr)]),
        "dmf_csr_check": (C.c_int, [C.POINTER(_Csr)]),
        "dmf_read_graph_bin": (C.c_int, [C.c_char_p, C.POINTER(_Csr)]),
        "dmf_write_graph_bin": (C.c_int, [C.c_char_p, C.POINTER(_Csr)]),
        "dmf_build_csr": (C.c_int, [C.c_int64, _I32P, _I32P, _I32P, C.POINTER(_Csr)]),
        "dmf_read_updates_bin": (C.c_int, [C.c_char_p, _I32P, C.POINTER(_I32P), C.POINTER(_I32P),
                                           C.POINTER(_I32P)]),
        "dmf_write_updates_bin": (C.c_int, [C.c_char_p, C.c_int32, _I32P, _I32P, _I32P]),
        "dmf_gen_rmat": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                   C.c_uint64, C.c_int32, C.c_int32, _I64P, C.POINTER(_I32P),
                                   C.POINTER(_I32P), C.POINTER(_I32P)]),
        "dmf_gen_rmat_n": (C.c_int, [C.c_int32, C.c_int64, C.c_double, C.c_double, C.c_double,
                                     C.c_uint64, C.c_int32, C.c_int32, _I64P, C.POINTER(_I32P),
                                     C.POINTER(_I32P), C.POINTER(_I32P)]),
        "dmf_gen_grid": (C.c_int, [C.c_int32, C.c_int32, C.c_uint64, C.c_int32, C.c_int32, _I64P,
                                   C.POINTER(_I32P), C.POINTER(_I32P), C.POINTER(_I32P)]),
        "dmf_pick_source_sink": (C.c_int, [C.c_int32, C.c_int64, _I32P, _I32P, _I32P, _I32P]),
        "dmf_gen_updates": (C.c_int, [C.c_int64, _I32P, _I32P, _I32P, C.c_int64, C.c_double,
                                      C.c_int32, C.c_int32, C.c_uint64, C.c_int, _I32P,
                                      C.POINTER(_I32P), C.POINTER(_I32P), C.POINTER(_I32P)]),
        "dmf_params_default": (None, [C.POINTER(_Csr), C.POINTER(Params)]),
        "dmf_engine_create": (C.c_int, [C.POINTER(_Csr), C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
        "dmf_engine_destroy": (C.c_int, [C.c_void_p]),
        "dmf_engine_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
        "dmf_engine_reset": (C.c_int, [C.c_void_p]),
        "dmf_engine_static_solve": (C.c_int, [C.c_void_p, C.POINTER(Params), _I64P]),
        "dmf_engine_apply_batch": (C.c_int, [C.c_void_p, C.c_int32, _I32P, _I32P, _I32P,
                                             C.POINTER(Params), _I64P]),
        "dmf_engine_apply_batch_device": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                                    C.c_void_p, C.POINTER(Params), _I64P]),
        "dmf_engine_flow": (C.c_int, [C.c_void_p, _I64P]),
        "dmf_engine_min_cut": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint8)]),
        "dmf_engine_complete_flow": (C.c_int, [C.c_void_p, _I64P]),
        "dmf_engine_read_state": (C.c_int, [C.c_void_p, _I32P, _I32P, _I64P, _I32P]),
        "dmf_engine_counters": (C.c_int, [C.c_void_p, C.POINTER(Counters)]),
        "dmf_engine_device_bytes": (C.c_int64, [C.c_void_p]),
        "dmf_recalculate_max_flow": (C.c_int, [C.POINTER(_Csr), C.c_int32, _I32P, _I32P, _I32P,
                                               C.c_int32, C.c_int32, C.c_int32, C.c_int32, _I64P,
                                               C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "dmf_static_recalculate": (C.c_int, [C.POINTER(_Csr), C.c_int32, _I32P, _I32P, _I32P,
                                             C.c_int32, C.c_int32, C.c_int32, C.c_int32, _I64P,
                                             C.POINTER(C.c_double)]),
        "dmf_static_recalculate_noalloc": (C.c_int, [C.POINTER(_Csr), C.c_int32, _I32P, _I32P, _I32P,
                                                     C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                                     _I64P]),
        "dmf_device_count": (C.c_int, []),
        "dmf_engine_autotune": (C.c_int, [C.c_void_p, C.POINTER(Params), _I32P, _I32P]),
        "dmf_engine_autotune_schedule": (C.c_int, [C.c_void_p, C.POINTER(Params), _I32P, _I32P]),
        "dmf_batch_create": (C.c_int, [C.c_int32, C.POINTER(_Csr), _I32P, _I32P, C.POINTER(C.c_void_p)]),
        "dmf_engine_instances": (C.c_int32, [C.c_void_p]),
        "dmf_batch_static_solve": (C.c_int, [C.c_void_p, C.POINTER(Params), _I64P]),
        "dmf_batch_apply": (C.c_int, [C.c_void_p, _I32P, _I32P, _I32P, _I32P, C.POINTER(Params), _I64P]),
        "dmf_batch_flows": (C.c_int, [C.c_void_p, _I64P]),
        "dmf_engine_footprint": (C.c_int64, [C.c_int32, C.c_int32]),
        "dmf_engine_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
        "dmf_engine_reserve_updates": (C.c_int, [C.c_void_p, C.c_int32]),
        "dmf_server_config_default": (None, [C.POINTER(ServerConfig)]),
        "dmf_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
        "dmf_server_open": (C.c_int, [C.c_void_p, C.POINTER(_Csr), C.c_int32, C.c_int32, C.c_int32,
                                      _I32P, _I64P]),
        "dmf_server_submit": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _I32P, _I32P, _I32P,
                                        C.POINTER(C.c_uint64)]),
        "dmf_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, C.POINTER(Result)]),
        "dmf_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
        "dmf_server_close": (C.c_int, [C.c_void_p, C.c_int32]),
        "dmf_server_destroy": (C.c_int, [C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def exported_symbols() -> list[str]:
    """Every dmf_* function declared in include/dmf.h (checked by the tests)."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "dmf.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"\b(dmf_[a-z0-9_]+)\s*\(", text)))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_I32P)


def _same_length(*arrays):
    """int32 copies of parallel record arrays, refused (EINVAL) unless all are
    one-dimensional and of one length: the library reads len(first) entries
    of each."""
    out = tuple(_i32(a) for a in arrays)
    if any(a.ndim != 1 or len(a) != len(out[0]) for a in out):
        raise DmfError(EINVAL, "record arrays differ in length or are not one-dimensional")
    return out


def _check(rc: int, what: str) -> None:
    if rc != OK:
        raise DmfError(rc, what)


def _take_i32(p, count: int) -> np.ndarray:
    """Copy a library-allocated int32 array into numpy and free it."""
    lib = _lib()
    if count > 0:
        out = np.ctypeslib.as_array(p, shape=(count,)).copy()
    else:
        out = np.zeros(0, dtype=np.int32)
    lib.dmf_free(C.cast(p, C.c_void_p))
    return out


@dataclass
class Graph:
    """Bidirectional CSR as the reference stores it (code/graph.hpp:21-30)."""

    n: int
    m: int
    row: np.ndarray
    col: np.ndarray
    cap: np.ndarray
    rev: np.ndarray

    def _csr(self) -> _Csr:
        self.row = _i32(self.row)
        self.col = _i32(self.col)
        self.cap = _i32(self.cap)
        self.rev = _i32(self.rev)
        if (self.n < 0 or self.m < 0 or self.row.shape != (self.n + 1,)
                or any(a.shape != (self.m,) for a in (self.col, self.cap, self.rev))):
            raise DmfError(EINVAL, f"CSR arrays do not match n={self.n}, m={self.m}")
        return _Csr(self.n, self.m, _ptr(self.row), _ptr(self.col), _ptr(self.cap), _ptr(self.rev))

    @classmethod
    def _from_csr(cls, c: _Csr) -> "Graph":
        n, m = c.n, c.m
        row = np.ctypeslib.as_array(c.row, shape=(n + 1,)).copy()
        col = np.ctypeslib.as_array(c.col, shape=(m,)).copy() if m else np.zeros(0, np.int32)
        cap = np.ctypeslib.as_array(c.cap, shape=(m,)).copy() if m else np.zeros(0, np.int32)
        rev = np.ctypeslib.as_array(c.rev, shape=(m,)).copy() if m else np.zeros(0, np.int32)
        _lib().dmf_csr_free(C.byref(c))
        return cls(n, m, row, col, cap, rev)

    def tails(self) -> np.ndarray:
        return np.repeat(np.arange(self.n, dtype=np.int32), np.diff(self.row))


def csr_check(g: Graph) -> None:
    """Structural check of a CSR (dmf_csr_check); raises DmfError(EFORMAT) on a
    malformed graph.  Engine creation runs the same check itself."""
    c = g._csr()
    _check(_lib().dmf_csr_check(C.byref(c)), "csr_check")


def read_graph_bin(path: str) -> Graph:
    c = _Csr()
    _check(_lib().dmf_read_graph_bin(path.encode(), C.byref(c)), f"read {path}")
    return Graph._from_csr(c)


def write_graph_bin(path: str, g: Graph) -> None:
    c = g._csr()
    _check(_lib().dmf_write_graph_bin(path.encode(), C.byref(c)), f"write {path}")


def build_csr(u, v, w) -> Graph:
    u, v, w = _same_length(u, v, w)
    c = _Csr()
    _check(_lib().dmf_build_csr(len(u), _ptr(u), _ptr(v), _ptr(w), C.byref(c)), "build_csr")
    return Graph._from_csr(c)


def read_updates_bin(path: str):
    k = C.c_int32()
    a, b, c = _I32P(), _I32P(), _I32P()
    _check(_lib().dmf_read_updates_bin(path.encode(), C.byref(k), C.byref(a), C.byref(b), C.byref(c)),
           f"read {path}")
    return _take_i32(a, k.value), _take_i32(b, k.value), _take_i32(c, k.value)


def write_updates_bin(path: str, src, dst, w) -> None:
    src, dst, w = _same_length(src, dst, w)
    _check(_lib().dmf_write_updates_bin(path.encode(), len(src), _ptr(src), _ptr(dst), _ptr(w)),
           f"write {path}")


def gen_rmat(scale: int, edge_factor: int, seed: int, a=0.57, b=0.19, c=0.19, cap_lo=1, cap_hi=100):
    k = C.c_int64()
    u, v, w = _I32P(), _I32P(), _I32P()
    _check(_lib().dmf_gen_rmat(scale, edge_factor, a, b, c, seed, cap_lo, cap_hi, C.byref(k),
                               C.byref(u), C.byref(v), C.byref(w)), "gen_rmat")
    return _take_i32(u, k.value), _take_i32(v, k.value), _take_i32(w, k.value)


def gen_rmat_n(n: int, samples: int, seed: int, a=0.57, b=0.19, c=0.19, cap_lo=1, cap_hi=100):
    """R-MAT over any vertex count n (dmf_gen_rmat_n)."""
    k = C.c_int64()
    u, v, w = _I32P(), _I32P(), _I32P()
    _check(_lib().dmf_gen_rmat_n(n, samples, a, b, c, seed, cap_lo, cap_hi, C.byref(k),
                                 C.byref(u), C.byref(v), C.byref(w)), "gen_rmat_n")
    return _take_i32(u, k.value), _take_i32(v, k.value), _take_i32(w, k.value)


def gen_grid(rows: int, cols: int, seed: int, cap_lo=1, cap_hi=1000):
    k = C.c_int64()
    u, v, w = _I32P(), _I32P(), _I32P()
    _check(_lib().dmf_gen_grid(rows, cols, seed, cap_lo, cap_hi, C.byref(k), C.byref(u), C.byref(v),
                               C.byref(w)), "gen_grid")
    return _take_i32(u, k.value), _take_i32(v, k.value), _take_i32(w, k.value)


def pick_source_sink(n: int, u, v):
    u, v = _same_length(u, v)
    s, t = C.c_int32(), C.c_int32()
    _check(_lib().dmf_pick_source_sink(n, len(u), _ptr(u), _ptr(v), C.byref(s), C.byref(t)),
           "pick_source_sink")
    return s.value, t.value


def gen_updates(u, v, cur_w: np.ndarray, count: int, seed: int, inc_frac=0.5, cap_lo=1, cap_hi=100,
                apply=True):
    """One batch of distinct-edge capacity updates; cur_w (int32) is updated in place if apply."""
    u, v = _same_length(u, v)
    if cur_w.dtype != np.int32 or not cur_w.flags.c_contiguous or cur_w.shape != u.shape:
        raise DmfError(EINVAL, "cur_w must be a contiguous int32 array with one entry per edge")
    k = C.c_int32()
    a, b, c = _I32P(), _I32P(), _I32P()
    _check(_lib().dmf_gen_updates(len(u), _ptr(u), _ptr(v), _ptr(cur_w), count, inc_frac, cap_lo,
                                  cap_hi, seed, 1 if apply else 0, C.byref(k), C.byref(a),
                                  C.byref(b), C.byref(c)), "gen_updates")
    return _take_i32(a, k.value), _take_i32(b, k.value), _take_i32(c, k.value)


def device_count() -> int:
    return int(_lib().dmf_device_count())


def default_params(g: Graph, variant="data", kc=None, relabel_interval=1, mode=None,
                   max_rounds=0, scheduler=SCHED_FIXED, lazy_start=0, cascade=0) -> Params:
    """dmf_params_default, then the given fields (mode None keeps the engine's
    default driver: stepped with graph-replayed BFS levels)."""
    p = Params()
    c = g._csr()
    _lib().dmf_params_default(C.byref(c), C.byref(p))
    p.variant = VARIANTS[variant] if isinstance(variant, str) else int(variant)
    if kc is not None:
        p.kc = kc
    p.relabel_interval = relabel_interval
    if mode is not None:
        p.mode = mode
    p.max_rounds = max_rounds
    p.scheduler = scheduler
    p.lazy_start = lazy_start
    p.cascade = cascade
    return p


class Engine:
    """Device-resident residual graph (dmf_engine)."""

    def __init__(self, g: Graph, s: int, t: int):
        self.g = g
        self.s, self.t = s, t
        self._h = C.c_void_p()
        c = g._csr()
        _check(_lib().dmf_engine_create(C.byref(c), s, t, C.byref(self._h)), "engine_create")

    def close(self) -> None:
        if self._h:
            _lib().dmf_engine_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def reset(self) -> None:
        _check(_lib().dmf_engine_reset(self._h), "engine_reset")

    def static_solve(self, params: Params) -> int:
        f = C.c_int64()
        _check(_lib().dmf_engine_static_solve(self._h, C.byref(params), C.byref(f)), "static_solve")
        return f.value

    def apply_batch(self, src, dst, w, params: Params) -> int:
        src, dst, w = _same_length(src, dst, w)
        f = C.c_int64()
        _check(_lib().dmf_engine_apply_batch(self._h, len(src), _ptr(src), _ptr(dst), _ptr(w),
                                             C.byref(params), C.byref(f)), "apply_batch")
        return f.value

    def apply_batch_status(self, src, dst, w, params: Params) -> int:
        src, dst, w = _same_length(src, dst, w)
        f = C.c_int64()
        return _lib().dmf_engine_apply_batch(self._h, len(src), _ptr(src), _ptr(dst), _ptr(w),
                                             C.byref(params), C.byref(f))

    def apply_batch_device(self, k: int, d_src: int, d_dst: int, d_w: int, params: Params) -> int:
        f = C.c_int64()
        _check(_lib().dmf_engine_apply_batch_device(self._h, k, C.c_void_p(d_src), C.c_void_p(d_dst),
                                                    C.c_void_p(d_w), C.byref(params), C.byref(f)),
               "apply_batch_device")
        return f.value

    def flow(self) -> int:
        f = C.c_int64()
        _check(_lib().dmf_engine_flow(self._h, C.byref(f)), "flow")
        return f.value

    def min_cut(self) -> np.ndarray:
        side = np.zeros(self.g.n, dtype=np.uint8)
        _check(_lib().dmf_engine_min_cut(self._h, side.ctypes.data_as(C.POINTER(C.c_uint8))), "min_cut")
        return side

    def complete_flow(self) -> int:
        """Maximum preflow -> maximum flow of the same value (dmf_engine_complete_flow)."""
        f = C.c_int64()
        _check(_lib().dmf_engine_complete_flow(self._h, C.byref(f)), "complete_flow")
        return f.value

    def state(self):
        res = np.zeros(self.g.m, np.int32)
        cap = np.zeros(self.g.m, np.int32)
        ex = np.zeros(self.g.n, np.int64)
        h = np.zeros(self.g.n, np.int32)
        _check(_lib().dmf_engine_read_state(self._h, _ptr(res), _ptr(cap), ex.ctypes.data_as(_I64P),
                                            _ptr(h)), "read_state")
        return res, cap, ex, h

    def counters(self) -> dict:
        c = Counters()
        _check(_lib().dmf_engine_counters(self._h, C.byref(c)), "counters")
        return c.as_dict()

    def device_bytes(self) -> int:
        return int(_lib().dmf_engine_device_bytes(self._h))

    def set_stream(self, stream_ptr: int) -> None:
        _check(_lib().dmf_engine_set_stream(self._h, C.c_void_p(stream_ptr)), "set_stream")

    def set_profiling(self, on: bool = True) -> None:
        """Times every BFS level launch with HIP events (counters bfs_level_ms/_launches)."""
        _check(_lib().dmf_engine_set_profiling(self._h, 1 if on else 0), "set_profiling")

    def reserve_updates(self, k: int) -> None:
        _check(_lib().dmf_engine_reserve_updates(self._h, k), "reserve_updates")

    def autotune(self, params: Params):
        """Picks the fastest (lane_max, wide_t) degree-bucket thresholds; resets the engine."""
        a, b = C.c_int32(), C.c_int32()
        _check(_lib().dmf_engine_autotune(self._h, C.byref(params), C.byref(a), C.byref(b)), "autotune")
        return a.value, b.value

    def autotune_schedule(self, params: Params):
        """autotune, then (fixed schedule) the fastest global-relabel interval, written into
        params.relabel_interval; returns (lane_max, wide_t, relabel_interval) and resets the engine."""
        a, b = C.c_int32(), C.c_int32()
        _check(_lib().dmf_engine_autotune_schedule(self._h, C.byref(params), C.byref(a), C.byref(b)),
               "autotune_schedule")
        return a.value, b.value, params.relabel_interval


def engine_footprint(n: int, m: int) -> int:
    return int(_lib().dmf_engine_footprint(n, m))


def server_config(hbm_budget=0, staging_bytes=0, max_sessions=0, params: Params | None = None,
                  **param_fields) -> ServerConfig:
    c = ServerConfig()
    _lib().dmf_server_config_default(C.byref(c))
    c.hbm_budget = hbm_budget
    c.staging_bytes = staging_bytes
    if max_sessions:
        c.max_sessions = max_sessions
    if params is not None:
        c.params = params
    for k, v in param_fields.items():
        setattr(c.params, k, VARIANTS[v] if k == "variant" and isinstance(v, str) else v)
    return c


class Server:
    """Streaming update server (dmf_server_*): resident sessions, batches
    admitted by device memory, solved as they arrive."""

    def __init__(self, config: ServerConfig | None = None):
        self._h = C.c_void_p()
        self._graphs = {}
        _check(_lib().dmf_server_create(C.byref(config) if config is not None else None, C.byref(self._h)),
               "server_create")

    def open(self, g: Graph, s: int, t: int, max_batch: int):
        """Returns (session id, static flow)."""
        sid = C.c_int32()
        f = C.c_int64()
        c = g._csr()
        _check(_lib().dmf_server_open(self._h, C.byref(c), s, t, max_batch, C.byref(sid), C.byref(f)),
               "server_open")
        self._graphs[sid.value] = g
        return sid.value, f.value

    def submit(self, session: int, src, dst, w) -> int:
        src, dst, w = _same_length(src, dst, w)
        tk = C.c_uint64()
        _check(_lib().dmf_server_submit(self._h, session, len(src), _ptr(src), _ptr(dst), _ptr(w),
                                        C.byref(tk)), "server_submit")
        return tk.value

    def submit_status(self, session: int, src, dst, w) -> int:
        src, dst, w = _same_length(src, dst, w)
        tk = C.c_uint64()
        return _lib().dmf_server_submit(self._h, session, len(src), _ptr(src), _ptr(dst), _ptr(w), C.byref(tk))

    def wait(self, ticket: int, timeout_ms: float = -1.0) -> Result:
        r = Result()
        _check(_lib().dmf_server_wait(self._h, ticket, timeout_ms, C.byref(r)), "server_wait")
        return r

    def stats(self) -> dict:
        st = ServerStats()
        _check(_lib().dmf_server_get_stats(self._h, C.byref(st)), "server_stats")
        return st.as_dict()

    def close(self, session: int) -> None:
        _check(_lib().dmf_server_close(self._h, session), "server_close")
        self._graphs.pop(session, None)

    def destroy(self) -> None:
        if self._h:
            _lib().dmf_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class BatchEngine(Engine):
    """Many independent (graph, s, t) instances resident as one disjoint union
    (dmf_batch_create); every solve covers all of them at once."""

    def __init__(self, graphs, sources, sinks):
        self._h = C.c_void_p()
        self.graphs = list(graphs)
        k = len(self.graphs)
        s, t = _same_length(sources, sinks)
        if len(s) != k:
            raise DmfError(EINVAL, f"batch_create takes one source and one sink per graph ({k})")
        self._csrs = (_Csr * k)(*[g._csr() for g in self.graphs])
        _check(_lib().dmf_batch_create(k, self._csrs, _ptr(s), _ptr(t), C.byref(self._h)), "batch_create")
        self.k = k
        self.offsets = np.concatenate([[0], np.cumsum([g.n for g in self.graphs])]).astype(np.int64)
        n = int(self.offsets[-1])
        self.g = Graph(n, sum(g.m for g in self.graphs), np.zeros(n + 1, np.int32), np.zeros(0, np.int32),
                       np.zeros(0, np.int32), np.zeros(0, np.int32))

    def static_solve_all(self, params: Params) -> np.ndarray:
        f = np.zeros(self.k, np.int64)
        _check(_lib().dmf_batch_static_solve(self._h, C.byref(params), f.ctypes.data_as(_I64P)),
               "batch_static_solve")
        return f

    def apply_all(self, updates, params: Params) -> np.ndarray:
        """updates: one (src, dst, w) per instance, instance-local ids."""
        if len(updates) != self.k:
            raise DmfError(EINVAL, f"apply_all takes one (src, dst, w) per instance ({self.k})")
        recs = [_same_length(*u) for u in updates]
        counts = _i32([len(r[0]) for r in recs])
        src = _i32(np.concatenate([r[0] for r in recs]))
        dst = _i32(np.concatenate([r[1] for r in recs]))
        w = _i32(np.concatenate([r[2] for r in recs]))
        f = np.zeros(self.k, np.int64)
        _check(_lib().dmf_batch_apply(self._h, _ptr(counts), _ptr(src), _ptr(dst), _ptr(w),
                                      C.byref(params), f.ctypes.data_as(_I64P)), "batch_apply")
        return f

    def instances(self) -> int:
        return int(_lib().dmf_engine_instances(self._h))

    def flows(self) -> np.ndarray:
        f = np.zeros(self.k, np.int64)
        _check(_lib().dmf_batch_flows(self._h, f.ctypes.data_as(_I64P)), "batch_flows")
        return f

    def min_cut_all(self):
        side = np.zeros(int(self.offsets[-1]), dtype=np.uint8)
        _check(_lib().dmf_engine_min_cut(self._h, side.ctypes.data_as(C.POINTER(C.c_uint8))), "min_cut")
        return [side[self.offsets[i]:self.offsets[i + 1]] for i in range(self.k)]


def recalculate_max_flow(g: Graph, src, dst, w, s: int, t: int, kc: int, variant="data"):
    src, dst, w = _same_length(src, dst, w)
    c = g._csr()
    f = C.c_int64()
    ms_s, ms_b = C.c_double(), C.c_double()
    var = VARIANTS[variant] if isinstance(variant, str) else variant
    _check(_lib().dmf_recalculate_max_flow(C.byref(c), len(src), _ptr(src), _ptr(dst), _ptr(w), s, t, kc,
                                           var, C.byref(f), C.byref(ms_s), C.byref(ms_b)),
           "recalculate_max_flow")
    return f.value, ms_s.value, ms_b.value


def static_recalculate(g: Graph, src, dst, w, s: int, t: int, kc: int, variant="data"):
    src, dst, w = _same_length(src, dst, w)
    c = g._csr()
    f = C.c_int64()
    ms = C.c_double()
    var = VARIANTS[variant] if isinstance(variant, str) else variant
    _check(_lib().dmf_static_recalculate(C.byref(c), len(src), _ptr(src), _ptr(dst), _ptr(w), s, t, kc,
                                         var, C.byref(f), C.byref(ms)), "static_recalculate")
    return f.value, ms.value


def static_recalculate_noalloc_status(g: Graph, src, dst, w, s: int, t: int, kc: int, variant="data",
                                      stream_ptr: int = 0):
    """dmf_static_recalculate_noalloc: (status, flow).  The library allocates
    nothing for it (fixed workspace; DMF_ESIZE beyond its limits)."""
    src, dst, w = _same_length(src, dst, w)
    c = g._csr()
    f = C.c_int64()
    var = VARIANTS[variant] if isinstance(variant, str) else variant
    rc = _lib().dmf_static_recalculate_noalloc(C.byref(c), len(src), _ptr(src), _ptr(dst), _ptr(w), s, t, kc, var,
                                              C.c_void_p(stream_ptr) if stream_ptr else None, C.byref(f))
    return rc, f.value


def static_recalculate_noalloc(g: Graph, src, dst, w, s: int, t: int, kc: int, variant="data") -> int:
    rc, f = static_recalculate_noalloc_status(g, src, dst, w, s, t, kc, variant)
    _check(rc, "static_recalculate_noalloc")
    return f
