"""Batched persistent launch: every instance bit-exact against the oracle, equal to
the looped single-instance launches, and stable across graph replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_batch(smp, count, n_min, n_max, seed):
    sizes, kinds = smp.batch_plan(count, n_min, n_max, seed)
    pms, pws = [], []
    for i, (n, k) in enumerate(zip(sizes, kinds)):
        pm, pw = smp.generate(int(k), int(n), seed=seed * 100003 + i)
        pms.append(pm)
        pws.append(pw)
    return sizes, kinds, pms, pws


def test_batch_matches_oracle_and_looped(gpu, oracle):
    smp = gpu
    sizes, kinds, pms, pws = make_batch(smp, 96, 1, 700, seed=11)
    want = [oracle.gs_seq(int(n), pm, pw)[0] for n, pm, pw in zip(sizes, pms, pws)]
    with smp.SmpBatch(sizes, pms, pws) as b:
        b.exec()
        for i, got in enumerate(b.result()):
            assert np.array_equal(got, want[i]), (i, sizes[i], kinds[i])
        p = b.profile()
        b.exec_looped()
        for i, got in enumerate(b.result()):
            assert np.array_equal(got, want[i]), ("looped", i)
        assert b.profile()["proposals"] == p["proposals"]
        # graph replays with eager work and synchronisation in between
        for _ in range(3):
            b.exec()
            res = b.result()
            assert all(np.array_equal(g, w) for g, w in zip(res, want))
            b.exec_looped()


@pytest.mark.parametrize("block", [256, 512, 1024])
def test_batch_fixed_configs(gpu, oracle, block):
    smp = gpu
    sizes, kinds, pms, pws = make_batch(smp, 40, 64, 2048, seed=block)
    want = [oracle.mw_seq(int(n), pm, pw)[0] for n, pm, pw in zip(sizes, pms, pws)]
    with smp.SmpBatch(sizes, pms, pws, block_threads=block, wgs_per_cu=1, use_graph=0) as b:
        b.exec()
        for i, got in enumerate(b.result()):
            assert np.array_equal(got, want[i]), (i, sizes[i], kinds[i])
