"""The batched and serving workloads at their stated sizes, every result against
the oracle (GS-Seq, bit-exact, tolerance 0 for this integer work):

  batched: 4,096 independent instances, n log-uniform in [64, 2048], 60% uniform /
           20% congested / 20% correlated, one auto-tuned persistent launch captured
           in a hipGraph, then a replay of the graph;
  serving: a Poisson stream of requests, n log-uniform in [128, 8192] over the same
           generator mix plus one request of exactly n = 8192 of each kind, once with
           the default split (large requests on the single-instance pipeline) and
           once with every request on the lanes (large_n above max_n, so the lane
           kernel runs n = 8192 with its partner words in LDS).
"""
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_all(oracle, sizes, pm_all, pw_all, offsets):
    def one(i):
        n = int(sizes[i])
        o = int(offsets[i])
        return oracle.gs_seq(n, pm_all[o:o + n * n], pw_all[o:o + n * n])[0]

    with ThreadPoolExecutor(8) as ex:
        return list(ex.map(one, range(len(sizes))))


def test_batched_4096_mixed_instances_graph(gpu, oracle):
    smp = gpu
    seed = 0x5EED0004
    sizes, kinds = smp.batch_plan(4096, 64, 2048, seed=seed)
    assert sizes.min() >= 64 and sizes.max() <= 2048
    frac = [(kinds == k).mean() for k in (0, 1, 3)]
    assert abs(frac[0] - 0.6) < 0.05 and abs(frac[1] - 0.2) < 0.05 and abs(frac[2] - 0.2) < 0.05
    sq = sizes.astype(np.int64) ** 2
    offsets = np.zeros(len(sizes), np.int64)
    offsets[1:] = np.cumsum(sq)[:-1]
    pm_all = np.empty(int(sq.sum()), np.int32)
    pw_all = np.empty(int(sq.sum()), np.int32)
    t0 = time.perf_counter()
    for i, (n, k) in enumerate(zip(sizes, kinds)):
        n = int(n)
        pm, pw = smp.generate(int(k), n, seed=seed * 1000003 + i, threads=1)
        pm_all[offsets[i]:offsets[i] + n * n] = pm
        pw_all[offsets[i]:offsets[i] + n * n] = pw
    print(f"generated 4096 instances in {time.perf_counter() - t0:.1f}s", flush=True)
    want = _oracle_all(oracle, sizes, pm_all, pw_all, offsets)
    print(f"oracle done at {time.perf_counter() - t0:.1f}s", flush=True)
    with smp.SmpBatch(sizes, pm_all, pw_all) as b:  # AUTO mode, graph on (defaults)
        print(f"batch created (tuned) at {time.perf_counter() - t0:.1f}s", flush=True)
        for replay in range(2):
            b.exec()
            got = b.result()
            bad = [i for i in range(len(sizes)) if not np.array_equal(got[i], want[i])]
            assert not bad, (replay, bad[:5], [int(sizes[i]) for i in bad[:5]],
                             [int(kinds[i]) for i in bad[:5]])
        prof = b.profile()
    print(f"batch mode {prof['mode']} {prof['exec_ms']:.2f} ms", flush=True)


def _serve_stream(smp, pool, picks, arrivals, **opts):
    results = [None] * len(picks)
    infos = [None] * len(picks)
    with smp.SmpServer(lanes=3, max_batch=64, max_n=8192, **opts) as srv:
        tickets = [None] * len(picks)
        start = time.perf_counter()

        def producer():
            for i, p in enumerate(picks):
                delay = start + arrivals[i] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                n, dm, dw = pool[p]
                tickets[i] = srv.submit(dm, dw, n)

        th = threading.Thread(target=producer)
        th.start()
        for i in range(len(picks)):
            while tickets[i] is None:
                time.sleep(0.0002)
            results[i], infos[i] = srv.wait(tickets[i], timeout_ms=60000)
        th.join()
        st = srv.stats()
    return results, infos, st


@pytest.mark.parametrize("split", ["default", "all_lanes"])
def test_serving_poisson_stream_to_8192(gpu, oracle, split):
    smp = gpu
    seed = 0x5EED0005
    sizes, kinds = smp.batch_plan(30, 128, 8192, seed=seed)
    sizes = np.concatenate([sizes, [8192, 8192, 8192]]).astype(np.int32)
    kinds = np.concatenate([kinds, [0, 1, 3]]).astype(np.int32)
    pool, want = [], []
    for i, (n, k) in enumerate(zip(sizes, kinds)):
        n = int(n)
        pm, pw = smp.generate(int(k), n, seed=seed * 7 + i, threads=16)
        pool.append((n, smp.DeviceArray(pm), smp.DeviceArray(pw)))
        want.append(oracle.gs_seq(n, pm, pw)[0])
    rng = np.random.default_rng(seed)
    count = 400
    # every pool entry at least once, the rest drawn uniformly; Poisson arrivals at 1,000/s
    picks = np.concatenate([np.arange(len(pool)), rng.integers(0, len(pool), count - len(pool))])
    rng.shuffle(picks)
    arrivals = np.cumsum(rng.exponential(1.0 / 1000.0, size=count))
    # all_lanes: no request goes to the single-instance workers, so the lane
    # kernels see every size up to 8,192
    opts = {} if split == "default" else {"large_n": 100000, "small_to_idle_worker": 0}
    print(f"serving {count} requests, split {split}", flush=True)
    got, infos, st = _serve_stream(smp, pool, picks, arrivals, **opts)
    for i, p in enumerate(picks):
        assert np.array_equal(got[i], want[p]), (i, int(sizes[p]), int(kinds[p]))
    big = [infos[i] for i, p in enumerate(picks) if sizes[p] == 8192]
    assert big
    if split == "all_lanes":
        assert st["large_solved"] == 0
        assert all(x["lane"] >= 0 for x in big)
    else:
        # every large request, and small ones that found a worker idle
        assert st["large_solved"] >= sum(int(sizes[p]) >= 512 for p in picks)
        assert all(x["lane"] == -1 for x in big)
    assert st["completed"] == count and st["arena_bytes"] == 0
    for _, dm, dw in pool:
        dm.free()
        dw.free()
