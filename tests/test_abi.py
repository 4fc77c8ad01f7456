"""The C-ABI library loads and exports every entry point include/dmf.h declares.
No compute call is made here (this suite also runs on machines without a GPU)."""
import ctypes
import os
import re

import pytest

import dyn_maxflow_amd as dmf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "dmf.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dmf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(dmf.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared in dmf.h but not exported: {missing}"


def test_python_mirror_binds_every_symbol():
    lib = dmf._lib()
    for n in declared_functions():
        fn = getattr(lib, n)
        assert fn.argtypes is not None, f"{n} has no ctypes signature"


def test_status_strings():
    lib = dmf._lib()
    for code in (dmf.OK, dmf.EINVAL, dmf.EIO, dmf.ENOARC, dmf.ENOMEM, dmf.EDEVICE, dmf.ENOCONV,
                 dmf.EFORMAT, dmf.EADMIT):
        assert lib.dmf_strerror(code)
    assert lib.dmf_strerror(-99) == b"unknown status"


def test_engine_create_rejects_bad_arguments_without_device_work():
    u = [0, 1, 2]
    v = [1, 2, 0]
    w = [3, 4, 5]
    g = dmf.build_csr(u, v, w)
    for s, t in ((0, 0), (-1, 2), (0, 7)):
        try:
            dmf.Engine(g, s, t)
        except dmf.DmfError as e:
            assert e.status == dmf.EINVAL
        else:
            raise AssertionError("bad s/t accepted")


def test_default_params_follow_reference_schedule():
    u, v, w = dmf.gen_rmat(8, 16, 3)
    g = dmf.build_csr(u, v, w)
    p = dmf.default_params(g)
    # kc = m/n (code/dyn_data_bench.cu:529), relabel every round
    assert p.kc == g.m // g.n
    assert p.relabel_interval == 1
    assert p.variant == dmf.VARIANT_DATA


def test_engine_footprint_bounds_the_arrays():
    for n, m in ((1000, 16000), (1 << 20, 31401920), (8387651, 258678280)):
        b = dmf.engine_footprint(n, m)
        assert b >= 28 * m + 31 * n  # the arc and vertex arrays alone
        assert b <= (28 * m + 40 * n + 12 * 4 * (2 * m + 16) + 8 * (m // 64 + 1) + 48 * min(n + 1, m // 17 + 1)
                     + (1 << 21))
    assert dmf.engine_footprint(2000, 32000) > dmf.engine_footprint(1000, 16000)


def test_server_refuses_without_device():
    if dmf.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(dmf.DmfError) as ex:
        dmf.Server(dmf.server_config())
    assert ex.value.status == dmf.EDEVICE


def test_mirror_refuses_mismatched_record_arrays():
    """The library reads len(src) entries of every record array; the mirror
    refuses arrays that disagree before anything reaches it."""
    with pytest.raises(dmf.DmfError) as e:
        dmf.build_csr([0, 1, 2], [1, 2], [5, 5, 5])
    assert e.value.status == dmf.EINVAL
    with pytest.raises(dmf.DmfError) as e:
        dmf.pick_source_sink(3, [0, 1], [1])
    assert e.value.status == dmf.EINVAL
    u, v, w = dmf.gen_rmat(6, 4, 1)
    with pytest.raises(dmf.DmfError) as e:
        dmf.gen_updates(u, v, w[:-1].copy(), 5, 1)
    assert e.value.status == dmf.EINVAL
    g = dmf.build_csr(u, v, w)
    with pytest.raises(dmf.DmfError) as e:
        dmf.recalculate_max_flow(g, [0, 1], [1], [3, 3], 0, 1, kc=1)
    assert e.value.status == dmf.EINVAL
    with pytest.raises(dmf.DmfError) as e:
        dmf.BatchEngine([g, g], [0, 0], [1])
    assert e.value.status == dmf.EINVAL
