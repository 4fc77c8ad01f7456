"""The allocation-free entry point (dmf_static_recalculate_noalloc) against
dmf_static_recalculate and the oracle's Dinic, through the C-ABI.

Two calls in a row share the fixed workspace; the second has a smaller input,
so a stale entry of the first would show.  Every variant; the update records
include a duplicate arc (last wins) and a self-loop (skipped).  After a warm
call, further calls leave the device's free memory exactly where it was.
Inputs beyond the workspace are refused with DMF_ESIZE before any device work."""
import ctypes

import numpy as np
import pytest

import oracle
from util import caps_after, rmat_case

pytestmark = pytest.mark.gpu


def free_device_bytes():
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def batch_for(dmf, g, u, v, w, seed):
    cur = w.copy()
    src, dst, nw = dmf.gen_updates(u, v, cur, len(u) // 20, seed, apply=False)
    # a duplicate arc (the last record wins) and a self-loop (skipped)
    src = np.concatenate([src, [u[0], u[0], 3]]).astype(np.int32)
    dst = np.concatenate([dst, [v[0], v[0], 3]]).astype(np.int32)
    nw = np.concatenate([nw, [77, 5, 9]]).astype(np.int32)
    return src, dst, nw


@pytest.mark.parametrize("variant", ["data", "pp", "topo"])
def test_noalloc_matches_static_recalculate(gpu, variant):
    dmf = gpu
    for scale, ef, seed in ((16, 8, 23), (11, 8, 22)):  # the second input is smaller
        g, s, t, (u, v, w) = rmat_case(scale, ef, seed)
        assert g.n <= 1 << 16 and g.m <= 1 << 20
        kc = max(1, g.m // g.n)
        src, dst, nw = batch_for(dmf, g, u, v, w, seed * 7)
        print(f"[noalloc] variant={variant} rmat-{scale} n={g.n} m={g.m} k={len(src)}", flush=True)
        f_na = dmf.static_recalculate_noalloc(g, src, dst, nw, s, t, kc, variant)
        f_ref, _ = dmf.static_recalculate(g, src, dst, nw, s, t, kc, variant)
        fd, _ = oracle.dinic(g, s, t, cap=caps_after(g, [(src, dst, nw)]))
        assert f_na == f_ref == fd, (scale, f_na, f_ref, fd)
        # no update records at all
        f0 = dmf.static_recalculate_noalloc(g, [], [], [], s, t, kc, variant)
        assert f0 == oracle.dinic(g, s, t)[0]


def test_noalloc_allocates_nothing_and_refuses(gpu):
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(12, 16, 5)
    kc = max(1, g.m // g.n)
    src, dst, nw = batch_for(dmf, g, u, v, w, 91)
    want = oracle.dinic(g, s, t, cap=caps_after(g, [(src, dst, nw)]))[0]
    assert dmf.static_recalculate_noalloc(g, src, dst, nw, s, t, kc) == want  # warm call
    before = free_device_bytes()
    for variant in ("data", "pp", "topo"):
        assert dmf.static_recalculate_noalloc(g, src, dst, nw, s, t, kc, variant) == want
    assert free_device_bytes() == before
    # errors: absent arc, negative capacity, out-of-range id, oversize inputs
    tails = g.tails()
    present = set(zip(tails.tolist(), g.col.tolist()))
    a = next(x for x in range(g.n) if (s, x) not in present and x != s)
    assert dmf.static_recalculate_noalloc_status(g, [s], [a], [5], s, t, kc)[0] == dmf.ENOARC
    assert dmf.static_recalculate_noalloc_status(g, [u[0]], [v[0]], [-1], s, t, kc)[0] == dmf.EINVAL
    assert dmf.static_recalculate_noalloc_status(g, [g.n], [0], [1], s, t, kc)[0] == dmf.EINVAL
    big, bs, bt, _ = rmat_case(17, 4, 1)
    assert big.n > 1 << 16
    assert dmf.static_recalculate_noalloc_status(big, [], [], [], bs, bt, 4)[0] == dmf.ESIZE
    many = np.zeros((1 << 20) + 1, np.int32)
    assert dmf.static_recalculate_noalloc_status(g, many, many, many, s, t, kc)[0] == dmf.ESIZE
    # the workspace still gives the right answer after the refusals
    assert dmf.static_recalculate_noalloc(g, src, dst, nw, s, t, kc) == want
