"""GPU parity of the single-instance engine, through the C ABI, against the oracle.

Bar: bit-exact man -> woman arrays (integer work), plus zero blocking pairs.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))

TAIL_CONFIGS = [
    dict(tail_mode=1),                          # GPU finishes every chain
    dict(tail_mode=2, tail_threshold=1),        # host takes the last free man
    dict(tail_mode=2, tail_threshold=37),       # host takes the last 37
    dict(tail_mode=1, prefix_cols=1),           # almost everything via the direct path
    dict(tail_mode=2, tail_threshold=8, prefix_cols=64),
    dict(tail_mode=1, prefix_cols=100000),      # full n x n PR matrix (reference layout)
    dict(tail_mode=2, tail_threshold=37, prefix_cols=1),  # every host chain is handed back
    dict(tail_mode=2, tail_threshold=5, poll_interval=1, coherent_hint=1),
    dict(tail_mode=2, tail_threshold=512, prefix_cols=32),  # a long host tail over a small prefix
]


# Init modes: 1 = full rank matrix then gather; 2 = sparse prefix pass (rank
# matrix only if a chain walks past the prefix or a bucket overflows).
INIT_CONFIGS = [
    dict(init_mode=1),
    dict(init_mode=2),
    dict(init_mode=2, prefix_cols=8, tail_mode=1),  # most chains walk past the prefix
    dict(init_mode=2, prefix_cols=8, tail_mode=2, tail_threshold=16),
    dict(init_mode=2, prefix_cols=100000),          # K = n: every man is in every bucket
]


def solve(smp, n, pm, pw, **opts):
    obj = smp.SmpObj.create_from_flat_lists(pm, pw, n)
    with smp.SmpEngineBamboo(obj, n, **opts) as e:
        m = e.find_stable_matching()
        return m, e.profile()


def test_golden_matchings(gpu, oracle):
    smp = gpu
    with open(os.path.join(HERE, "golden", "matchings.json")) as f:
        rows = json.load(f)
    for r in rows:
        pm, pw = smp.generate(r["kind"], r["n"], seed=r["seed"])
        for cfg in (TAIL_CONFIGS[0], TAIL_CONFIGS[2]):
            got, prof = solve(smp, r["n"], pm, pw, **cfg)
            assert got.tolist() == r["match"], (r["kind"], r["n"], r["seed"], cfg)
            assert prof["proposals"] == r["proposals"]


@pytest.mark.parametrize("cfg", TAIL_CONFIGS)
@pytest.mark.parametrize("kind", range(8))
def test_parity_mid_sizes(gpu, oracle, kind, cfg):
    smp = gpu
    for n in (1, 2, 3, 5, 63, 64, 65, 257, 1000):
        if kind == 7 and n < 4:
            continue
        pm, pw = smp.generate(kind, n, seed=1000 + n)
        want, t = oracle.gs_seq(n, pm, pw)
        got, prof = solve(smp, n, pm, pw, **cfg)
        assert np.array_equal(got, want), (kind, n, cfg)
        assert prof["proposals"] == t["proposals"]


@pytest.mark.parametrize("cfg", [dict(), dict(tail_mode=1), dict(tail_mode=2, tail_threshold=37)])
@pytest.mark.parametrize("n", [1024, 1500, 3001])
def test_prefix_widened_after_overflow(gpu, oracle, n, cfg):
    """Correlated lists overflow the sparse pass's buckets; with the automatic
    prefix width the engine then gathers 256 columns (n >= 1024) and keeps them
    for its later solves of the same lists.  Every solve is the oracle's
    matching with the oracle's proposal count."""
    smp = gpu
    pm, pw = smp.generate(smp.Workload.CORRELATED, n, seed=600 + n)
    want, t = oracle.gs_seq(n, pm, pw)
    obj = smp.SmpObj.create_from_flat_lists(pm, pw, n)
    with smp.SmpEngineBamboo(obj, n, **cfg) as e:
        for rep in range(3):
            got = e.find_stable_matching()
            prof = e.profile()
            assert np.array_equal(got, want), (n, cfg, rep)
            assert prof["proposals"] == t["proposals"], (n, cfg, rep)
            # The first solve runs the sparse pass, which overflows; later ones
            # go straight to the full init.
            assert prof["rank_matrix"] == 1 and prof["sparse_init"] == (1 if rep == 0 else 0), (n, cfg, rep)
            assert prof["prefix_cols"] == 256, (n, cfg, rep)
    # A width the caller sets is kept.
    with smp.SmpEngineBamboo(obj, n, prefix_cols=96, **cfg) as e:
        got = e.find_stable_matching()
        assert np.array_equal(got, want)
        assert e.profile()["prefix_cols"] == 96


@pytest.mark.parametrize("cfg", INIT_CONFIGS)
@pytest.mark.parametrize("kind", range(8))
def test_init_modes(gpu, oracle, kind, cfg):
    """Both init modes, including chains that walk past a sparse prefix (the rank
    matrix is then built before they resume), give the oracle's matching."""
    smp = gpu
    for n in (1, 2, 3, 5, 63, 64, 65, 257, 1000):
        if kind == 7 and n < 4:
            continue
        pm, pw = smp.generate(kind, n, seed=5000 + n)
        want, t = oracle.gs_seq(n, pm, pw)
        got, prof = solve(smp, n, pm, pw, **cfg)
        assert np.array_equal(got, want), (kind, n, cfg)
        assert prof["proposals"] == t["proposals"]
        if prof["is_perfect"] or prof["master_list"]:
            continue
        if cfg["init_mode"] == 1:
            assert prof["sparse_init"] == 0 and prof["rank_matrix"] == 1
        else:
            assert prof["sparse_init"] == 1


@pytest.mark.parametrize("n", [7, 64, 333, 1024, 4001])
def test_sparse_prefix_matrices(gpu, n):
    """Kernel-level parity of the sparse prefix pass: the PR prefix it writes
    against a numpy gather, and the rank matrix it builds on demand against
    numpy argsort."""
    smp = gpu
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=77 + n)
    M, W = pm.reshape(n, n), pw.reshape(n, n)
    rank_w = np.argsort(W, axis=1)
    for cols in (1, 64, n):
        obj = smp.SmpObj.create_from_flat_lists(pm, pw, n)
        with smp.SmpEngineBamboo(obj, n, prefix_cols=cols, init_mode=2, precheck=0, tail_mode=1) as e:
            e.find_stable_matching()
            prof = e.profile()
            K = min(cols, n)
            # Sparse buffers cover prefixes up to 256 columns (any width for
            # n <= 2048); a wider prefix takes the full init.
            assert prof["sparse_init"] == (1 if K <= 256 or n <= 2048 else 0)
            pr = e.pr_matrix()
            assert pr.shape == (n, K, 2)
            assert np.array_equal(pr[:, :, 0], M[:, :K])
            assert np.array_equal(pr[:, :, 1], rank_w[M[:, :K], np.arange(n)[:, None]])
            assert np.array_equal(e.rank_matrix(), rank_w)


def test_sparse_overflow_falls_back(gpu, oracle):
    """All men share one list and the precheck is off: each of the first K women
    of that list is named by all n men, far more than a bucket holds, so the
    sparse pass writes the full rank matrix, the proposal kernel stands down, and
    the engine gathers the prefix and runs the proposal phase again (with the
    GPU finishing alone, and with a host tail).  Result: the serial-dictatorship
    matching with n(n+1)/2 proposals."""
    smp = gpu
    n = 1000
    pm, pw = smp.generate(smp.Workload.CONGESTED, n, seed=31)
    want = oracle.serial_dictatorship(n, pm[:n], pw)
    for cfg in (dict(tail_mode=1), dict(tail_mode=2, tail_threshold=16)):
        got, prof = solve(smp, n, pm, pw, init_mode=2, prefix_cols=16, precheck=0, **cfg)
        assert prof["sparse_init"] == 1 and prof["rank_matrix"] == 1
        assert np.array_equal(got, want), cfg
        assert prof["proposals"] == n * (n + 1) // 2


def test_sparse_overflow_is_remembered(gpu, oracle):
    """Correlated lists: every man's prefix names the same few women, so the
    sparse pass overflows.  In auto mode the engine then keeps the full init for
    these lists; forced sparse mode keeps trying.  Every run is bit-exact."""
    smp = gpu
    n = 3000
    pm, pw = smp.generate(smp.Workload.CORRELATED, n, seed=8)
    want, t = oracle.gs_seq(n, pm, pw)
    obj = smp.SmpObj.create_from_flat_lists(pm, pw, n)
    for mode, second_sparse in ((0, 0), (2, 1)):
        with smp.SmpEngineBamboo(obj, n, init_mode=mode) as e:
            assert np.array_equal(e.find_stable_matching(), want)
            p1 = e.profile()
            assert p1["sparse_init"] == 1 and p1["rank_matrix"] == 1
            assert np.array_equal(e.find_stable_matching(), want)
            p2 = e.profile()
            assert p2["sparse_init"] == second_sparse and p2["rank_matrix"] == 1
            assert p2["proposals"] == t["proposals"]


@pytest.mark.parametrize("kind", range(8))
def test_mw_cas_baseline_schedule(gpu, oracle, kind):
    """The reference's GPU baseline (one thread per man, 32-bit CAS) through the
    same engine: bit-exact, same proposal count, no fast paths taken."""
    smp = gpu
    for n in (1, 2, 5, 64, 65, 300, 1000):
        if kind == 7 and n < 4:
            continue
        pm, pw = smp.generate(kind, n, seed=2000 + n)
        want, t = oracle.gs_seq(n, pm, pw)
        got, prof = solve(smp, n, pm, pw, schedule=smp.Schedule.MW_CAS)
        assert prof["schedule"] == smp.Schedule.MW_CAS
        assert prof["is_perfect"] == 0 and prof["master_list"] == 0
        assert np.array_equal(got, want), (kind, n)
        assert prof["proposals"] == t["proposals"], (kind, n)


@pytest.mark.parametrize("schedule", [2, 3, 4, 5])
@pytest.mark.parametrize("kind", range(8))
def test_alternative_schedules(gpu, oracle, kind, schedule):
    """The reference's PR-entry CAS walk (kernel.cu:122-165), its atomicMin walks
    with and without the read-and-skip (kernel.cu:468-512, 515-552) and the
    frontier rounds with LDS pre-reduction: bit-exact, same proposal count, on
    every generator, with a 64-column prefix (the LA walks and the frontier read
    past it through the rank matrix)."""
    smp = gpu
    for n in (1, 2, 5, 64, 65, 300, 1000):
        if kind == 7 and n < 4:
            continue
        pm, pw = smp.generate(kind, n, seed=2500 + n)
        want, t = oracle.gs_seq(n, pm, pw)
        for cfg in (dict(), dict(prefix_cols=64)):
            got, prof = solve(smp, n, pm, pw, schedule=schedule, **cfg)
            assert prof["schedule"] == schedule
            assert np.array_equal(got, want), (kind, n, schedule, cfg)
            assert prof["proposals"] == t["proposals"], (kind, n, schedule, cfg)


def test_tuner_picks_a_schedule_on_contention(gpu, oracle):
    """All-shared lists with the precheck off at n = 8,000 take more than 5 ms per
    solve, so the tuner measures the schedules and keeps the fastest; later
    solves use it and stay bit-exact.  A fast instance keeps the default."""
    smp = gpu
    n = 8000
    pm, pw = smp.generate(smp.Workload.CONGESTED, n, seed=41)
    want, t = oracle.gs_seq(n, pm, pw)
    obj = smp.SmpObj.create_from_flat_lists(pm, pw, n)
    with smp.SmpEngineBamboo(obj, n, precheck=0) as e:
        r = e.tune(2)
        assert 0 <= r["schedule"] < 6 and r["candidates"] == 0, r
        got = e.find_stable_matching()
        assert e.profile()["schedule"] == r["schedule"]
        assert np.array_equal(got, want)
        assert e.profile()["proposals"] == t["proposals"]
    pm, pw = smp.generate(smp.Workload.UNIFORM, 2000, seed=42)
    with smp.SmpEngineBamboo(smp.SmpObj.create_from_flat_lists(pm, pw, 2000), 2000) as e:
        r = e.tune(2)
        assert r["schedule"] == 0 and r["candidates"] > 0
        assert np.array_equal(e.find_stable_matching(), oracle.gs_seq(2000, pm, pw)[0])


def test_frontier_schedule_20000(gpu, oracle):
    """The frontier schedule at n = 20,000 on uniform and correlated lists, and on
    all-shared lists with the precheck off (no fast path)."""
    smp = gpu
    n = 20000
    for kind, opts in ((0, {}), (3, {}), (1, dict(precheck=0))):
        pm, pw = smp.generate(kind, n, seed=0x5EED0003)
        want, t = oracle.gs_seq(n, pm, pw)
        got, prof = solve(smp, n, pm, pw, schedule=5, **opts)
        assert np.array_equal(got, want), kind
        assert prof["proposals"] == t["proposals"], kind


def host_ranks(pw, n):
    """The women's inverse rank matrix derived on the host, the ranked engine's
    input: rank[w, pref_w[w, r]] = r."""
    rows = np.asarray(pw, dtype=np.int32).reshape(n, n)
    rank = np.empty_like(rows)
    np.put_along_axis(rank, rows, np.broadcast_to(np.arange(n, dtype=np.int32), (n, n)), axis=1)
    return rank


@pytest.mark.parametrize("cfg", [TAIL_CONFIGS[0], TAIL_CONFIGS[2], TAIL_CONFIGS[3], TAIL_CONFIGS[6]])
@pytest.mark.parametrize("kind", range(8))
def test_ranked_input(gpu, oracle, kind, cfg):
    """The women's ranks given instead of their lists: the same matching and
    proposal count as the oracle on every generator, with the GPU finishing
    alone and with a host tail, and with chains walking past the prefix
    (prefix_cols 1, where every proposal past position 0 reads the given ranks).
    All-shared men's lists take the general walk (no women's lists to walk)."""
    smp = gpu
    for n in (1, 2, 5, 64, 65, 300, 1000):
        if kind == 7 and n < 4:
            continue
        pm, pw = smp.generate(kind, n, seed=3000 + n)
        want, t = oracle.gs_seq(n, pm, pw)
        with smp.SmpEngineBamboo.from_ranked(pm, host_ranks(pw, n), n, **cfg) as e:
            got = e.find_stable_matching()
            prof = e.profile()
        assert np.array_equal(got, want), (kind, n, cfg)
        assert prof["proposals"] == t["proposals"], (kind, n, cfg)
        assert prof["sparse_init"] == 0 and prof["master_list"] == 0


def test_config1_ranks_given(gpu, oracle):
    """Config 1 as stated: n = 2,000 uniform, the women's inverse rank matrix
    derived on the host; lists and ranks resident in HBM."""
    smp = gpu
    n = 2000
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=0x5EED0001)
    want, t = oracle.gs_seq(n, pm, pw)
    rank = host_ranks(pw, n)
    dm, dr = smp.DeviceArray(pm), smp.DeviceArray(rank)
    for cfg in (dict(tail_mode=1), dict(tail_mode=2, tail_threshold=37), dict(tail_mode=0)):
        with smp.SmpEngineBamboo.from_device_ranked(dm, dr, n, **cfg) as e:
            for _ in range(2):
                assert np.array_equal(e.find_stable_matching(), want), cfg
            assert e.profile()["proposals"] == t["proposals"]
            assert np.array_equal(e.rank_matrix().reshape(n, n), rank)
    assert oracle.count_blocking_pairs(n, pm, pw, want) == 0


def test_ranked_input_errors(gpu, oracle):
    """A rank outside [0, n) is reported (where a proposal or the gather reads
    it), not used; the MW-CAS schedule and the init-only run need the women's
    lists and are refused."""
    smp = gpu
    n = 64
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=5)
    rank = host_ranks(pw, n)
    first = np.asarray(pm, dtype=np.int32).reshape(n, n)[:, 0]
    bad = rank.copy()
    bad[first[0], 0] = n  # read by man 0's first proposal, inside the prefix
    with smp.SmpEngineBamboo.from_ranked(pm, bad, n, tail_mode=1, precheck=0) as e:
        with pytest.raises(smp.SmpError) as ei:
            e.exec()
        assert "outside" in str(ei.value)
    with pytest.raises(smp.SmpError) as ei:
        smp.SmpEngineBamboo.from_ranked(pm, rank, n, schedule=smp.Schedule.MW_CAS)
    assert ei.value.status == 1
    want, _ = oracle.gs_seq(n, pm, pw)
    with smp.SmpEngineBamboo.from_ranked(pm, rank, n) as e:
        with pytest.raises(smp.SmpError):
            e.init_only()
        assert np.array_equal(e.find_stable_matching(), want)  # still usable after the refusal


def test_config1_uniform_2000(gpu, oracle):
    smp = gpu
    n = 2000
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=0x5EED0001)
    want, _ = oracle.gs_seq(n, pm, pw)
    for cfg in TAIL_CONFIGS:
        got, _ = solve(smp, n, pm, pw, **cfg)
        assert np.array_equal(got, want), cfg
    assert oracle.count_blocking_pairs(n, pm, pw, want) == 0


def test_repeated_exec_and_device_resident(gpu, oracle):
    smp = gpu
    n = 1500
    pm, pw = smp.generate(smp.Workload.CORRELATED, n, seed=3)
    want, _ = oracle.mw_seq(n, pm, pw)
    dm, dw = smp.DeviceArray(pm), smp.DeviceArray(pw)
    with smp.SmpEngineBamboo.from_device(dm, dw, n, tail_mode=2, tail_threshold=16) as e:
        for _ in range(4):
            assert np.array_equal(e.find_stable_matching(), want)


def test_config2_uniform_20000(gpu, oracle):
    smp = gpu
    n = 20000
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=0x5EED0002)
    want, t = oracle.gs_seq(n, pm, pw)
    for cfg in (dict(tail_mode=1), dict(tail_mode=2, tail_threshold=16), dict(tail_mode=0),
                dict(tail_mode=2, tail_threshold=256), dict(init_mode=1),
                dict(init_mode=2, prefix_cols=32)):
        got, prof = solve(smp, n, pm, pw, **cfg)
        assert np.array_equal(got, want), cfg
        assert prof["proposals"] == t["proposals"]
    assert oracle.count_blocking_pairs(n, pm, pw, want) == 0


def test_tail_extension_windows_20000(gpu, oracle):
    """Seed 0x5EED0007 of the n = 20,000 uniform workload has two men whose final
    list positions are 128 and 131, past a 128-column prefix.  The host tail asks
    the GPU for extension windows of their lists and finishes every chain itself
    (nothing handed back); with the GPU finishing alone, the parked chains take
    the same route.  Both bit-exact against the oracle."""
    smp = gpu
    n = 20000
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=0x5EED0007)
    want, t = oracle.gs_seq(n, pm, pw)
    pos = np.argmax(pm.reshape(n, n) == want[:, None], axis=1)
    assert int(pos.max()) >= 128
    for cfg in (dict(prefix_cols=128, tail_mode=2, tail_threshold=24),
                dict(prefix_cols=96, tail_mode=2, tail_threshold=64),
                dict(prefix_cols=128, tail_mode=1)):
        got, prof = solve(smp, n, pm, pw, **cfg)
        assert np.array_equal(got, want), cfg
        assert prof["proposals"] == t["proposals"], cfg
        assert prof["tail_windows"] > 0 and prof["bailed"] == 0, (cfg, prof["tail_windows"], prof["bailed"])
        assert prof["rank_matrix"] == 0, cfg


@pytest.mark.parametrize("entries", [4096, 0])
def test_tail_windows_run_out_falls_back(gpu, oracle, entries):
    """A 1-column prefix with 5,000 men handed to the host: every chain leaves the
    prefix at once.  With window storage capped at 4,096 entries (64 per first
    window) the chains go back to the GPU, which resumes them with the full rank
    matrix; with the automatic storage the host finishes them from windows.
    Bit-exact either way."""
    smp = gpu
    n = 6000
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=77)
    want, t = oracle.gs_seq(n, pm, pw)
    got, prof = solve(smp, n, pm, pw, prefix_cols=1, tail_mode=2, tail_threshold=5000, init_mode=2,
                      tail_window_entries=entries)
    assert np.array_equal(got, want)
    assert prof["proposals"] == t["proposals"]
    if entries:
        assert prof["bailed"] > 0 and prof["rank_matrix"] == 1
    else:
        assert prof["bailed"] == 0 and prof["tail_windows"] >= 5000 and prof["tail_rounds"] > 0


def test_tail_band_correlated_20000(gpu, oracle):
    """Correlated lists, n = 20,000, seed 0x5EED0004: 594 men end past position 256
    (the deepest at 441), and the host tail's chains reach them one at a time.
    After a few rounds of windows the host takes a band of every man's list
    beyond the prefix and finishes from it.  Bit-exact against the oracle."""
    smp = gpu
    n = 20000
    pm, pw = smp.generate(smp.Workload.CORRELATED, n, seed=0x5EED0004)
    want, t = oracle.gs_seq(n, pm, pw)
    pos = np.argmax(pm.reshape(n, n) == want[:, None], axis=1)
    assert int((pos >= 256).sum()) > 500
    got, prof = solve(smp, n, pm, pw, tail_mode=2, tail_threshold=64)
    assert np.array_equal(got, want)
    assert prof["proposals"] == t["proposals"]
    assert prof["tail_bands"] >= 1 and prof["tail_full_matrix"] == 0 and prof["bailed"] == 0, prof


@pytest.mark.parametrize("kind,param", [(4, 5), (7, 0)])
def test_fig9_workloads_4000(gpu, oracle, kind, param):
    """The reference's own Fig 9 inputs (benchmark/figure_9.cu:40-57) at n = 4,000:
    RANDOM with groups of 5 (about n^2 / 2 proposals under heavy contention) and
    SOLO (one displacement chain of n^2 - n + 1 proposals).  SOLO's chain walks
    every list to its end, so the host tail takes the whole PR matrix."""
    smp = gpu
    n = 4000
    pm, pw = smp.generate(kind, n, seed=0x5EED0009, param=param)
    want, t = oracle.gs_seq(n, pm, pw)
    got, prof = solve(smp, n, pm, pw)
    assert np.array_equal(got, want)
    assert prof["proposals"] == t["proposals"]
    if kind == 7:
        assert t["proposals"] == n * n - n + 1
        assert prof["tail_full_matrix"] == 1 and prof["bailed"] == 0


@pytest.mark.parametrize("tail", [dict(), dict(tail_mode=1)])
def test_deep_lists_take_the_whole_pr_matrix(gpu, oracle, tail):
    """The reference's RANDOM with groups of 5 (about n^2 / 2 proposals): the
    first solve's chains walk far past the prefix, so the proposal kernel parks
    them, the whole PR matrix is gathered and they resume over it; later solves
    of the same lists start with the whole matrix.  Uniform and correlated lists
    keep the automatic prefix.  Every solve bit-exact, with and without the
    host tail."""
    smp = gpu
    n = 4000
    pm, pw = smp.generate(smp.Workload.REF_RANDOM, n, seed=0x5EED0009, param=5)
    want, t = oracle.gs_seq(n, pm, pw)
    dm, dw = smp.DeviceArray(pm), smp.DeviceArray(pw)
    with smp.SmpEngineBamboo.from_device(dm, dw, n, **tail) as e:
        seen = []
        for _ in range(3):
            e.exec()
            p = e.profile()
            assert np.array_equal(e.result(), want)
            assert p["proposals"] == t["proposals"]
            seen.append((p["prefix_cols"], p["deep_restart"]))
        assert seen == [(n, 1), (n, 0), (n, 0)], seen
    dm.free()
    dw.free()
    for kind in (smp.Workload.UNIFORM, smp.Workload.CORRELATED):
        pmu, pwu = smp.generate(kind, n, seed=3)
        wantu, _ = oracle.gs_seq(n, pmu, pwu)
        got, prof = solve(smp, n, pmu, pwu, **tail)
        assert np.array_equal(got, wantu) and prof["prefix_cols"] < n and prof["deep_restart"] == 0, kind


@pytest.mark.parametrize("kind", [0, 1, 3, 4, 6, 7])
def test_start_order_either_way(gpu, oracle, kind):
    """Men started in the order of woman 0's list (the default) or by index:
    the same matching and proposal count, with and without the host tail."""
    smp = gpu
    for n in (64, 700, 2500):
        pm, pw = smp.generate(kind, n, seed=4100 + n, param=5 if kind == 4 else 0)
        want, t = oracle.gs_seq(n, pm, pw)
        for order in (0, 1):
            for tail in (dict(tail_mode=1), dict(tail_mode=2, tail_threshold=32)):
                got, prof = solve(smp, n, pm, pw, precheck=0, start_order=order, **tail)
                assert np.array_equal(got, want), (kind, n, order, tail)
                assert prof["proposals"] == t["proposals"], (kind, n, order, tail)


def test_config3_congested_and_shifted_20000(gpu, oracle):
    smp = gpu
    n = 20000
    pm, pw = smp.generate(smp.Workload.CONGESTED, n, seed=0x5EED0003)
    want = oracle.serial_dictatorship(n, pm[:n], pw)
    got, prof = solve(smp, n, pm, pw, tail_mode=1)
    assert prof["master_list"] == 1 and prof["master_both"] == 0
    assert np.array_equal(got, want)
    assert prof["proposals"] == n * (n + 1) // 2
    # the man-proposing walk on the same input (precheck off: no fast paths)
    got, prof = solve(smp, n, pm, pw, tail_mode=1, precheck=0)
    assert prof["master_list"] == 0
    assert np.array_equal(got, want)
    assert prof["proposals"] == n * (n + 1) // 2
    pm, pw = smp.generate(smp.Workload.SHIFTED, n, seed=0x5EED0003)
    got, prof = solve(smp, n, pm, pw)
    assert prof["is_perfect"] == 1
    assert np.array_equal(got, pm.reshape(n, n)[:, 0])


def test_reference_congested_30000(gpu):
    """The reference's CONGESTED lists at n = 30,000 (men share L, women share
    P): the stable matching is match[P[k]] = L[k] with n(n+1)/2 man-proposing
    proposals, answered without a walk."""
    smp = gpu
    n = 30000
    pm, pw = smp.generate(smp.Workload.REF_CONGESTED, n, seed=0x5EED0003)
    L, P = pm[:n], pw[:n]
    want = np.empty(n, np.int32)
    want[P] = L
    obj = smp.SmpObj.create_from_flat_lists(pm, pw, n)
    with smp.SmpEngineBamboo(obj, n) as e:
        for _ in range(2):  # the second solve starts from the first one's state
            got = e.find_stable_matching()
            prof = e.profile()
            assert prof["master_list"] == 1 and prof["master_both"] == 1
            assert np.array_equal(got, want)
            assert prof["proposals"] == n * (n + 1) // 2


@pytest.mark.parametrize("n", [2, 3, 64, 65, 1000, 2048])
def test_master_list_path(gpu, oracle, n):
    """Every man holds the same list: the women-proposing walk answers, bit-exact,
    with the man-proposing proposal count n(n+1)/2; a list that differs from the
    others in its last two entries (or only in a later row) keeps the general path."""
    smp = gpu
    for kind in (smp.Workload.CONGESTED, smp.Workload.REF_CONGESTED):
        pm, pw = smp.generate(kind, n, seed=4242 + n)
        want, t = oracle.gs_seq(n, pm, pw)
        assert t["proposals"] == n * (n + 1) // 2
        got, prof = solve(smp, n, pm, pw)
        assert prof["master_list"] == 1 and prof["is_perfect"] == 0
        # the reference's CONGESTED lists: the women share a list too (for
        # tiny n the random women's lists of CONGESTED can coincide as well)
        women_share = bool((pw.reshape(n, n) == pw[:n]).all())
        assert women_share or kind == smp.Workload.CONGESTED
        assert prof["master_both"] == women_share, kind
        assert np.array_equal(got, want), kind
        assert prof["proposals"] == t["proposals"]
        assert oracle.count_blocking_pairs(n, pm, pw, got) == 0
        if kind == smp.Workload.REF_CONGESTED:
            # one woman's list differs from woman 0's (her last two entries,
            # or for n = 2 woman 0's own): the women-proposing walk answers
            for row in sorted({0, n - 1}):
                wnear = pw.copy().reshape(n, n)
                wnear[row, -2], wnear[row, -1] = wnear[row, -1], wnear[row, -2]
                wnear = wnear.reshape(-1)
                want3, t3 = oracle.gs_seq(n, pm, wnear)
                got3, prof3 = solve(smp, n, pm, wnear)
                assert prof3["master_list"] == 1 and prof3["master_both"] == 0, row
                assert np.array_equal(got3, want3), row
                assert prof3["proposals"] == t3["proposals"], row
        near = pm.copy().reshape(n, n)
        row = n - 1
        near[row, -2], near[row, -1] = near[row, -1], near[row, -2]
        near = near.reshape(-1)
        want2, t2 = oracle.gs_seq(n, near, pw)
        got2, prof2 = solve(smp, n, near, pw)
        assert prof2["master_list"] == 0
        assert np.array_equal(got2, want2)
        assert prof2["proposals"] == t2["proposals"]


def test_master_list_invalid_and_state(gpu):
    smp = gpu
    n = 300
    pm, pw = smp.generate(smp.Workload.CONGESTED, n, seed=9)
    bad_w = pw.copy()
    bad_w[11 * n + 5] = bad_w[11 * n + 6]  # woman 11's row is not a permutation
    with pytest.raises(smp.SmpError) as ei:
        solve(smp, n, pm, bad_w)
    assert ei.value.status == 1
    obj = smp.SmpObj.create_from_flat_lists(pm, pw, n)
    with smp.SmpEngineBamboo(obj, n) as e:
        e.find_stable_matching()
        assert e.profile()["master_list"] == 1
        with pytest.raises(smp.SmpError):
            e.rank_matrix()  # no rank matrix was built on this path


def test_invalid_lists_are_rejected(gpu):
    smp = gpu
    n = 50
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=1)
    bad_w = pw.copy()
    bad_w[7 * n + 3] = bad_w[7 * n + 4]  # woman 7's row is not a permutation
    with pytest.raises(smp.SmpError) as ei:
        solve(smp, n, pm, bad_w)
    assert ei.value.status == 1
    bad_m = pm.copy()
    bad_m[5 * n + 2] = n + 3  # out of range entry
    with pytest.raises(smp.SmpError) as ei:
        solve(smp, n, bad_m, pw, precheck=0)
    assert ei.value.status == 1


@pytest.mark.parametrize("n", [1, 7, 64, 333, 1024, 4001])
def test_init_only_rank_and_pr_matrices(gpu, n):
    """Kernel-level parity: k_invert_rows against numpy argsort; k_gather_pr against a
    numpy gather, for prefix and full widths (SmpInitEngine / figure 5 path)."""
    smp = gpu
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=n)
    M, W = pm.reshape(n, n), pw.reshape(n, n)
    rank_w = np.argsort(W, axis=1)
    for cols in (1, 64, n):
        with smp.SmpEngineBamboo(smp.SmpObj.create_from_flat_lists(pm, pw, n), n, prefix_cols=cols) as e:
            t = e.init_only(init_pr=True)
            assert t["pr_cols"] == min(cols, n)
            assert np.array_equal(e.rank_matrix(), rank_w)
            pr = e.pr_matrix()
            K = min(cols, n)
            assert pr.shape == (n, K, 2)
            assert np.array_equal(pr[:, :, 0], M[:, :K])
            assert np.array_equal(pr[:, :, 1], rank_w[M[:, :K], np.arange(n)[:, None]])


def test_handed_back_chains_finish_on_deep_lists(gpu, oracle):
    """The reference's RANDOM lists at n = 2,500 with the host taking
    every chain at the proposal kernel's first poll: the host hands hundreds of
    deep chains back to the GPU, whose resumed walk passes the deep-list step
    count; it must finish them (no restart) rather than park them with nothing
    left to resume them."""
    smp = gpu
    n = 2500
    pm, pw = smp.generate(smp.Workload.REF_RANDOM, n, seed=3000 + n)
    want, t = oracle.gs_seq(n, pm, pw)
    got, prof = solve(smp, n, pm, pw, tail_mode=2, tail_threshold=100000, poll_interval=1)
    assert np.array_equal(got, want)
    assert prof["proposals"] == t["proposals"]
    assert prof["bailed"] > 0
