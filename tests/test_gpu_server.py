"""Streaming server (dmf_server_*): every result is the same alone and in the
mix, admission by device memory, per-session order.  Results are compared
with a solo engine replaying the same session's batches and with Dinic."""
import pytest

import oracle
from util import caps_after, rmat_case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _sessions(dmf):
    out = []
    for scale, ef, seed in ((9, 16, 61), (11, 8, 62), (12, 16, 63), (10, 32, 64)):
        g, s, t, (u, v, w) = rmat_case(scale, ef, seed)
        out.append((g, s, t, u, v, w))
    u, v, w = dmf.gen_grid(24, 24, 65, 1, 1000)
    g = dmf.build_csr(u, v, w)
    out.append((g, 0, g.n - 1, u, v, w))
    return out


def test_results_same_alone_and_in_mix(gpu):
    dmf = gpu
    cases = _sessions(dmf)
    nb = 6
    batches = []
    for i, (g, s, t, u, v, w) in enumerate(cases):
        cur = w.copy()
        bl = []
        for b in range(nb):
            k = max(1, len(u) * (1 + b % 3) // 40)
            bl.append(dmf.gen_updates(u, v, cur, k, 900 + 10 * i + b, inc_frac=0.4,
                                      cap_hi=1000 if i == 4 else 100))
        batches.append(bl)
    # a staging arena smaller than all queued batches together: admission waits
    total = sum(12 * len(b[0]) for bl in batches for b in bl)
    cfg = dmf.server_config(staging_bytes=max(1 << 16, total // 6))
    say(f"[server] mix of {len(cases)} sessions, staged bytes needed {total}, arena {cfg.staging_bytes}")
    with dmf.Server(cfg) as srv:
        sids = []
        for g, s, t, u, v, w in cases:
            sid, f0 = srv.open(g, s, t, max_batch=max(len(b[0]) for bl in batches for b in bl))
            assert f0 == oracle.dinic(g, s, t)[0]
            sids.append(sid)
        tickets = {}
        for b in range(nb):  # interleaved arrivals
            for i, sid in enumerate(sids):
                tickets[(i, b)] = srv.submit(sid, *batches[i][b])
        res = {key: srv.wait(tk) for key, tk in tickets.items()}
        st = srv.stats()
        say(f"[server] stats {st}")
        assert st["completed"] == len(cases) * nb
        assert 0 < st["peak_staged_bytes"] <= cfg.staging_bytes
    for i, (g, s, t, u, v, w) in enumerate(cases):
        with dmf.Engine(g, s, t) as solo:
            p = dmf.default_params(g, mode=1, scheduler=dmf.SCHED_LEARNED)
            solo.static_solve(dmf.default_params(g, mode=1))
            applied = []
            for b in range(nb):
                r = res[(i, b)]
                assert r.status == 0 and r.session == sids[i] and r.k == len(batches[i][b][0])
                assert r.submit_ms <= r.admit_ms <= r.start_ms <= r.done_ms
                applied.append(batches[i][b])
                f_solo = solo.apply_batch(*batches[i][b], p)
                f_dinic = oracle.dinic(g, s, t, cap=caps_after(g, applied))[0]
                assert r.flow == f_solo == f_dinic, (i, b, r.flow, f_solo, f_dinic)
        # per-session order: each batch started after the previous one finished
        for b in range(1, nb):
            assert res[(i, b)].start_ms >= res[(i, b - 1)].done_ms


def test_admission_by_memory(gpu):
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(10, 16, 66)
    need = dmf.engine_footprint(g.n, g.m)
    assert need > 28 * g.m
    # a budget that holds the arena but not the session
    cfg = dmf.server_config(hbm_budget=need // 2, staging_bytes=1 << 16)
    with dmf.Server(cfg) as srv:
        with pytest.raises(dmf.DmfError) as ex:
            srv.open(g, s, t, max_batch=1000)
        assert ex.value.status == dmf.EADMIT
    cfg = dmf.server_config(staging_bytes=1 << 16)
    with dmf.Server(cfg) as srv:
        sid, _ = srv.open(g, s, t, max_batch=100)
        cur = w.copy()
        src, dst, nw = dmf.gen_updates(u, v, cur, 200, 5)
        assert srv.submit_status(sid, src, dst, nw) == dmf.EADMIT  # above max_batch
        tk = srv.submit(sid, src[:50], dst[:50], nw[:50])
        r = srv.wait(tk)
        assert r.status == 0
        with pytest.raises(dmf.DmfError) as ex:
            srv.wait(tk)  # returned once only
        assert ex.value.status == dmf.EINVAL
        srv.close(sid)
        assert srv.submit_status(sid, src[:5], dst[:5], nw[:5]) == dmf.EINVAL
    cfg = dmf.server_config(staging_bytes=4096)
    with dmf.Server(cfg) as srv:
        sid, _ = srv.open(g, s, t, max_batch=10000)
        big = dmf.gen_updates(u, v, w.copy(), 2000, 6)
        assert srv.submit_status(sid, *big) == dmf.EADMIT  # alone larger than the arena


def test_single_session_queue_learned(gpu):
    """Batches queued back to back on one session (the next one staged while
    the current one solves) under the learned scheduler."""
    dmf = gpu
    g, s, t, (u, v, w) = rmat_case(13, 16, 67)
    cur = w.copy()
    bl = [dmf.gen_updates(u, v, cur, int(k), 700 + i, inc_frac=0.4)
          for i, k in enumerate((100, 3000, 20000, 800, 12000, 5000, 40000, 2000))]
    with dmf.Server(dmf.server_config()) as srv:
        sid, f0 = srv.open(g, s, t, max_batch=40000)
        tks = [srv.submit(sid, *b) for b in bl]
        res = [srv.wait(tk) for tk in tks]
    applied = []
    for b, r in zip(bl, res):
        applied.append(b)
        assert r.status == 0
        assert r.flow == oracle.dinic(g, s, t, cap=caps_after(g, applied))[0]
