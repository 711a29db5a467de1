"""smp_exec_device_noalloc (the entry point that allocates nothing) against the
engine's smp_create_device + smp_exec on the same device-resident lists, through
the C ABI: the same man -> woman array (bit-exact) and the same proposal count,
on consecutive calls whose second input is smaller than the first (the shared
scratch keeps the first call's contents), plus the host-side size limit and a
malformed list rejected as smp_exec rejects it."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_N = 4096


def noalloc(smp, dm, dw, n):
    from pactae_amd import _abi
    lib = _abi.load()
    out = smp.DeviceArray(np.full(n, -7, np.int32))
    props = ctypes.c_int64(-1)
    st = lib.smp_exec_device_noalloc(dm.ptr, dw.ptr, n, 0, out.ptr, ctypes.byref(props))
    res = out.to_host()
    out.free()
    return st, res, props.value, lib.smp_last_error().decode()


def engine(smp, dm, dw, n):
    with smp.SmpEngineBamboo.from_device(dm, dw, n) as e:
        e.exec()
        return e.result(), e.profile()["proposals"]


@pytest.mark.parametrize("kinds", [(0, 0), (0, 1), (3, 7), (1, 3)])
def test_noalloc_matches_engine_twice_then_smaller(gpu, oracle, kinds):
    smp = gpu
    sizes = (MAX_N, 1500) if kinds == (0, 0) else (2048, 333)
    for kind, n in zip(kinds, sizes):
        print(f"noalloc: kind {kind} n {n}", flush=True)
        pm, pw = smp.generate(kind, n, seed=70 + n)
        dm, dw = smp.DeviceArray(pm), smp.DeviceArray(pw)
        st, got, props, err = noalloc(smp, dm, dw, n)
        assert st == 0, err
        want, want_props = engine(smp, dm, dw, n)
        assert np.array_equal(got, want), (kind, n)
        assert props == want_props, (kind, n)
        if n <= 2048:
            assert np.array_equal(got, oracle.gs_seq(n, pm, pw)[0]), (kind, n)
        dm.free()
        dw.free()


def test_noalloc_size_limit_and_bad_lists(gpu):
    smp = gpu
    n = MAX_N + 1
    pm = np.zeros(4, np.int32)
    dm, dw = smp.DeviceArray(pm), smp.DeviceArray(pm)
    from pactae_amd import _abi
    lib = _abi.load()
    out = smp.DeviceArray(np.zeros(4, np.int32))
    st = lib.smp_exec_device_noalloc(dm.ptr, dw.ptr, n, 0, out.ptr, None)
    assert st == 1
    assert "4096" in lib.smp_last_error().decode()
    dm.free()
    dw.free()
    out.free()
    n = 200
    pm, pw = smp.generate(0, n, seed=3)
    bad = pw.copy()
    bad[7 * n] = bad[7 * n + 1]  # woman 7's row is not a permutation
    dm, dw = smp.DeviceArray(pm), smp.DeviceArray(bad)
    st, _, _, err = noalloc(smp, dm, dw, n)
    assert st == 1 and "permutation" in err
    # and a good solve right after the failed one
    dw.free()
    dw = smp.DeviceArray(pw)
    st, got, _, err = noalloc(smp, dm, dw, n)
    assert st == 0, err
    want, _ = engine(smp, dm, dw, n)
    assert np.array_equal(got, want)
    dm.free()
    dw.free()
