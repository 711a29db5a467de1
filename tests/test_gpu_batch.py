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


@pytest.mark.parametrize("mode", ["auto", "fused", "pool"])
def test_batch_matches_oracle_and_looped(gpu, oracle, mode):
    smp = gpu
    sizes, kinds, pms, pws = make_batch(smp, 96, 1, 700, seed=11)
    want = [oracle.gs_seq(int(n), pm, pw)[0] for n, pm, pw in zip(sizes, pms, pws)]
    md = {"auto": smp.SmpBatch.AUTO, "fused": smp.SmpBatch.FUSED, "pool": smp.SmpBatch.POOL}[mode]
    with smp.SmpBatch(sizes, pms, pws, mode=md) as b:
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
def test_batch_fused_configs(gpu, oracle, block):
    smp = gpu
    sizes, kinds, pms, pws = make_batch(smp, 40, 64, 2048, seed=block)
    want = [oracle.mw_seq(int(n), pm, pw)[0] for n, pm, pw in zip(sizes, pms, pws)]
    with smp.SmpBatch(sizes, pms, pws, block_threads=block, wgs_per_cu=1, use_graph=0,
                      mode=smp.SmpBatch.FUSED) as b:
        b.exec()
        assert b.profile()["mode"] == smp.SmpBatch.FUSED
        for i, got in enumerate(b.result()):
            assert np.array_equal(got, want[i]), (i, sizes[i], kinds[i])


@pytest.mark.parametrize("layout", ["women_rows", "men_rows"])
@pytest.mark.parametrize("group", [8, 16, 64])
@pytest.mark.parametrize("slots", [1, 2, 4])
def test_batch_pool_configs(gpu, oracle, slots, group, layout):
    """Every pool shape, on a batch with more instances than slots in the grid."""
    smp = gpu
    lay = {"women_rows": smp.SmpBatch.WOMEN_ROWS, "men_rows": smp.SmpBatch.MEN_ROWS}[layout]
    sizes, kinds, pms, pws = make_batch(smp, 300, 1, 400, seed=slots * 31 + group)
    want = [oracle.gs_seq(int(n), pm, pw)[0] for n, pm, pw in zip(sizes, pms, pws)]
    with smp.SmpBatch(sizes, pms, pws, mode=smp.SmpBatch.POOL, block_threads=512, wgs_per_cu=1,
                      group_lanes=group, pool_slots=slots, rank_layout=lay, use_graph=0) as b:
        for _ in range(2):
            b.exec()
            p = b.profile()
            assert (p["mode"], p["pool_slots"], p["group_lanes"], p["rank_layout"]) == (
                smp.SmpBatch.POOL, slots, group, lay)
            for i, got in enumerate(b.result()):
                assert np.array_equal(got, want[i]), (i, sizes[i], kinds[i])


def test_batch_pool_large_instances(gpu, oracle):
    """Pool mode on instances up to n = 2048 of every kind, both layouts."""
    smp = gpu
    sizes, kinds, pms, pws = make_batch(smp, 24, 900, 2048, seed=77)
    want = [oracle.mw_seq(int(n), pm, pw)[0] for n, pm, pw in zip(sizes, pms, pws)]
    for lay in (smp.SmpBatch.WOMEN_ROWS, smp.SmpBatch.MEN_ROWS):
        with smp.SmpBatch(sizes, pms, pws, mode=smp.SmpBatch.POOL, rank_layout=lay) as b:
            b.exec()
            for i, got in enumerate(b.result()):
                assert np.array_equal(got, want[i]), (lay, i, sizes[i], kinds[i])


@pytest.mark.parametrize("config", ["fused", "pool_women_rows", "pool_men_rows"])
def test_batch_master_list_instances(gpu, oracle, config):
    """Half-congested batch (every man of an instance holds the same list): both
    modes solve these from the women's side; results and the men's proposal
    counts n(n+1)/2 must equal the oracle's, beside uniform instances in the same
    batch."""
    smp = gpu
    opts = {"fused": {"mode": smp.SmpBatch.FUSED},
            "pool_women_rows": {"mode": smp.SmpBatch.POOL, "rank_layout": smp.SmpBatch.WOMEN_ROWS},
            "pool_men_rows": {"mode": smp.SmpBatch.POOL, "rank_layout": smp.SmpBatch.MEN_ROWS}}[config]
    sizes, kinds = smp.batch_plan(40, 1, 600, seed=21, mix=(0.5, 0.5, 0.0))
    pms, pws, want, props = [], [], [], 0
    for i, (n, k) in enumerate(zip(sizes, kinds)):
        pm, pw = smp.generate(int(k), int(n), seed=5000 + i)
        m, t = oracle.gs_seq(int(n), pm, pw)
        pms.append(pm)
        pws.append(pw)
        want.append(m)
        props += t["proposals"]
    with smp.SmpBatch(sizes, pms, pws, **opts) as b:
        for _ in range(2):
            b.exec()
            for i, got in enumerate(b.result()):
                assert np.array_equal(got, want[i]), (i, sizes[i], kinds[i])
            assert b.profile()["proposals"] == props


@pytest.mark.parametrize("mode", [1, 2])
def test_batch_invalid_lists_raise(gpu, mode):
    smp = gpu
    sizes, kinds, pms, pws = make_batch(smp, 12, 16, 300, seed=4)
    bad = [pw.copy() for pw in pws]
    n = int(sizes[5])
    bad[5][2 * n] = bad[5][2 * n + 1]  # a woman's row that is not a permutation
    with smp.SmpBatch(sizes, pms, bad, mode=mode, use_graph=0) as b:
        with pytest.raises(smp.SmpError):
            b.exec()


@pytest.mark.parametrize("mode", [1, 2])
def test_batch_invalid_shared_mens_list_raises(gpu, mode):
    """Every man of one instance holds the same list, but that list repeats a
    woman: the master-list path must reject it (its Linv finds the repeat)."""
    smp = gpu
    sizes, kinds, pms, pws = make_batch(smp, 12, 16, 300, seed=6)
    n = int(sizes[3])
    row = np.arange(n, dtype=np.int32)
    row[1] = row[0]
    pms = [pm.copy() for pm in pms]
    pms[3] = np.tile(row, n)
    with smp.SmpBatch(sizes, pms, pws, mode=mode, use_graph=0) as b:
        with pytest.raises(smp.SmpError):
            b.exec()
