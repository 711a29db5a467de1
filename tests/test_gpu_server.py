"""Serving loop: each request's result is bit-exact against the oracle, and the
same alone as in a concurrent mix (tolerance 0); the plain baseline agrees."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def requests(smp, count, seed, n_min=16, n_max=1500):
    sizes, kinds = smp.batch_plan(count, n_min, n_max, seed)
    return [(int(n), smp.generate(int(k), int(n), seed=seed * 7919 + i)) for i, (n, k) in
            enumerate(zip(sizes, kinds))]


def test_alone_equals_mix_equals_oracle(gpu, oracle):
    smp = gpu
    reqs = requests(smp, 48, seed=3)
    want = [oracle.gs_seq(n, pm, pw)[0] for n, (pm, pw) in reqs]
    with smp.SmpServer(lanes=3, max_batch=8) as srv:
        alone = []
        for n, (pm, pw) in reqs:
            m, info = srv.wait(srv.submit(pm, pw, n), timeout_ms=60000)
            assert info["batch_size"] == 1
            alone.append(m)
        tickets = [srv.submit(pm, pw, n) for n, (pm, pw) in reqs]
        mix, infos = [], []
        for t in tickets:
            m, info = srv.wait(t, timeout_ms=60000)
            mix.append(m)
            infos.append(info)
        st = srv.stats()
    for i in range(len(reqs)):
        assert np.array_equal(alone[i], want[i]), i
        assert np.array_equal(mix[i], want[i]), i
    assert max(i["batch_size"] for i in infos) > 1  # requests really were batched
    assert st["completed"] == 2 * len(reqs)
    assert st["arena_bytes"] == 0


def test_device_resident_requests_and_admission(gpu, oracle):
    smp = gpu
    reqs = requests(smp, 24, seed=5, n_min=256, n_max=2048)
    want = [oracle.mw_seq(n, pm, pw)[0] for n, (pm, pw) in reqs]
    dev = [(smp.DeviceArray(pm), smp.DeviceArray(pw)) for _, (pm, pw) in reqs]
    # A budget that holds only a few of these requests at once forces admission to wait
    # for memory to free up.
    budget = 3 * 2048 * 2048 * 2 + (1 << 20)
    with smp.SmpServer(lanes=2, max_batch=4, hbm_budget_bytes=budget) as srv:
        tickets = [srv.submit(dm, dw, n) for (n, _), (dm, dw) in zip(reqs, dev)]
        got = [srv.wait(t, timeout_ms=60000)[0] for t in tickets]
        st = srv.stats()
    for i, g in enumerate(got):
        assert np.array_equal(g, want[i]), i
    assert st["peak_arena_bytes"] <= budget


@pytest.mark.parametrize("large_n", [300, 100000])
@pytest.mark.parametrize("idle", [0, 1])
def test_large_request_worker(gpu, oracle, large_n, idle):
    """Requests at or above large_n go to the single-instance pipeline workers
    (lane -1).  The rest go to the lanes, or with small_to_idle_worker to a
    worker that is idle while none waits for the workers.  A large_n above
    max_n means no workers at all.  Every result matches the oracle."""
    smp = gpu
    reqs = requests(smp, 30, seed=13, n_min=64, n_max=900)
    want = [oracle.gs_seq(n, pm, pw)[0] for n, (pm, pw) in reqs]
    dev = [(smp.DeviceArray(pm), smp.DeviceArray(pw)) for _, (pm, pw) in reqs]
    with smp.SmpServer(lanes=2, max_batch=8, max_n=1024, large_n=large_n, small_to_idle_worker=idle) as srv:
        tickets = [srv.submit(pm, pw, n) for n, (pm, pw) in reqs]
        out = [srv.wait(t, timeout_ms=60000) for t in tickets]
        st = srv.stats()
        # One at a time, lists in HBM: with small_to_idle_worker every request
        # finds an idle worker; without it only the large ones go there.
        alone = [srv.wait(srv.submit(dm, dw, n), timeout_ms=60000) for (n, _), (dm, dw) in zip(reqs, dev)]
    for (n, _), (got, info), (got1, info1), w in zip(reqs, out, alone, want):
        assert np.array_equal(got, w), n
        assert np.array_equal(got1, w), n
        if n >= large_n:
            assert info["lane"] == -1 and info1["lane"] == -1, (n, info["lane"])
        elif not idle or large_n > 1024:
            assert info["lane"] >= 0 and info1["lane"] >= 0, (n, info["lane"])
        else:
            assert info1["lane"] == -1, (n, info1["lane"])
    if not idle or large_n > 1024:
        assert st["large_solved"] == sum(n >= large_n for n, _ in reqs)
    for dm, dw in dev:
        dm.free()
        dw.free()


def test_plain_mode_matches(gpu, oracle):
    """Plain mode, host-submitted and device-resident requests.  Its pipeline
    has buffers of its own, so the arena holds only the host requests' list
    copies (no rank block): device-resident requests take no arena room."""
    smp = gpu
    reqs = requests(smp, 12, seed=9)
    with smp.SmpServer(mode=smp.SmpServer.PLAIN) as srv:
        tickets = [srv.submit(pm, pw, n) for n, (pm, pw) in reqs]
        got = [srv.wait(t, timeout_ms=60000)[0] for t in tickets]
        peak_host = srv.stats()["peak_arena_bytes"]
    for (n, (pm, pw)), g in zip(reqs, got):
        assert np.array_equal(g, oracle.gs_seq(n, pm, pw)[0])
    assert peak_host <= sum(2 * (((n * n * 4) + 255) // 256 * 256) for n, _ in reqs)
    dev = [(n, smp.DeviceArray(pm), smp.DeviceArray(pw)) for n, (pm, pw) in reqs]
    with smp.SmpServer(mode=smp.SmpServer.PLAIN) as srv:
        tickets = [srv.submit(dm, dw, n) for n, dm, dw in dev]
        got = [srv.wait(t, timeout_ms=60000)[0] for t in tickets]
        st = srv.stats()
    for (n, (pm, pw)), g in zip(reqs, got):
        assert np.array_equal(g, oracle.gs_seq(n, pm, pw)[0])
    assert st["peak_arena_bytes"] <= 256 * len(reqs), st["peak_arena_bytes"]
    for _, dm, dw in dev:
        dm.free()
        dw.free()


def test_invalid_request_fails_alone(gpu, oracle):
    smp = gpu
    n = 64
    pm, pw = smp.generate(0, n, seed=1)
    bad = pw.copy()
    bad[3 * n] = bad[3 * n + 1]
    with smp.SmpServer(lanes=2, max_batch=4) as srv:
        t_bad = srv.submit(pm, bad, n)
        t_ok = srv.submit(pm, pw, n)
        with pytest.raises(smp.SmpError):
            srv.wait(t_bad, timeout_ms=60000)
        got, _ = srv.wait(t_ok, timeout_ms=60000)
    assert np.array_equal(got, oracle.gs_seq(n, pm, pw)[0])


@pytest.mark.parametrize("max_batch", [1, 8])
@pytest.mark.parametrize("fused", [1, 0])
def test_lane_kernels_at_the_small_request_bound(gpu, oracle, max_batch, fused):
    """Lane launches whose requests all have n <= 512 run the small one-kernel
    solve; a launch whose largest request has n <= 2,048 runs the fused
    per-instance kernel (fused_lanes 1) or the rank pass and the chain kernel;
    a larger one the rank pass and the chain kernel.  Sizes on both sides of
    both bounds, every kind (congested lists share one men's list), all on the
    lanes."""
    smp = gpu
    sizes = [1, 2, 3, 64, 511, 512, 513, 700, 512, 40, 2047, 2048, 1500, 2049, 3000]
    kinds = [0, 1, 3, 0, 3, 1, 0, 3, 0, 1, 0, 3, 1, 1, 0]
    reqs = [(n, smp.generate(k, n, seed=4242 + i)) for i, (n, k) in enumerate(zip(sizes, kinds))]
    with smp.SmpServer(lanes=2, max_batch=max_batch, max_n=4096, large_n=100000, small_to_idle_worker=0,
                       fused_lanes=fused) as srv:
        tickets = [srv.submit(pm, pw, n) for n, (pm, pw) in reqs]
        out = [srv.wait(t, timeout_ms=60000) for t in tickets]
    for (n, (pm, pw)), (got, info) in zip(reqs, out):
        want, t = oracle.gs_seq(n, pm, pw)
        assert np.array_equal(got, want), n
        assert info["proposals"] == t["proposals"], n
        assert info["lane"] >= 0
