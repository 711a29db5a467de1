"""GPU parity of the one-workgroup late phase (k_late: the chains the proposal
kernel parks at late_threshold, walked with the women's state in LDS until
tail_threshold are left for the host), through the C ABI, against the oracle.

Bar: bit-exact man -> woman arrays and proposal counts.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LATE_CONFIGS = [
    dict(tail_mode=2, tail_threshold=16, late_threshold=256),
    dict(tail_mode=2, tail_threshold=1, late_threshold=64),
    # Everything goes to the late phase at the proposal kernel's first poll.
    dict(tail_mode=2, tail_threshold=1, late_threshold=100000, poll_interval=1),
    # Chains park at the end of a one- or eight-column prefix.
    dict(tail_mode=2, tail_threshold=4, late_threshold=100000, prefix_cols=1),
    dict(tail_mode=2, tail_threshold=4, late_threshold=512, prefix_cols=8),
    # Full rank matrix, and the whole PR matrix (the late phase walks to n).
    dict(tail_mode=2, tail_threshold=8, late_threshold=512, init_mode=1),
    dict(tail_mode=2, tail_threshold=2, late_threshold=100000, prefix_cols=100000),
]


def solve(smp, n, pm, pw, **opts):
    obj = smp.SmpObj.create_from_flat_lists(pm, pw, n)
    with smp.SmpEngineBamboo(obj, n, **opts) as e:
        m = e.find_stable_matching()
        return m, e.profile()


@pytest.mark.parametrize("cfg", LATE_CONFIGS)
@pytest.mark.parametrize("kind", range(8))
def test_late_phase_parity(gpu, oracle, kind, cfg):
    smp = gpu
    ran = 0
    for n in (2, 3, 5, 63, 64, 65, 257, 1000, 2500):
        if kind == 7 and n < 4:
            continue
        pm, pw = smp.generate(kind, n, seed=3000 + n)
        want, t = oracle.gs_seq(n, pm, pw)
        got, prof = solve(smp, n, pm, pw, **cfg)
        assert np.array_equal(got, want), (kind, n, cfg)
        assert prof["proposals"] == t["proposals"], (kind, n, cfg)
        ran += prof["late_chains"] > 0
    if kind in (0, 3):  # uniform and correlated lists: the late phase runs at some n
        assert ran > 0, cfg


def test_late_phase_uniform_20000(gpu, oracle):
    smp = gpu
    n = 20000
    for seed in (0x5EED0002, 0x5EED0007):
        pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=seed)
        want, t = oracle.gs_seq(n, pm, pw)
        for cfg in (dict(tail_mode=2, tail_threshold=16, late_threshold=256),
                    dict(tail_mode=2, tail_threshold=48, late_threshold=1024, prefix_cols=128),
                    dict(tail_mode=2, tail_threshold=1, late_threshold=4096, prefix_cols=96)):
            got, prof = solve(smp, n, pm, pw, **cfg)
            assert np.array_equal(got, want), (seed, cfg)
            assert prof["proposals"] == t["proposals"], (seed, cfg)
            # With a threshold above the queue's 10,896 chains the late phase may
            # find more parked chains than it holds and leave them to the host.
            if cfg["late_threshold"] <= 1024:
                assert prof["late_chains"] > cfg["tail_threshold"], (seed, cfg, prof["late_chains"])


def test_late_phase_off_above_its_lds(gpu):
    """n = 27,300 does not fit one workgroup's LDS (6 bytes per woman and man):
    the option is ignored, and the result is the one without it."""
    smp = gpu
    n = 27300
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=11, threads=16)
    base, p0 = solve(smp, n, pm, pw, tail_mode=2, tail_threshold=16)
    got, prof = solve(smp, n, pm, pw, tail_mode=2, tail_threshold=16, late_threshold=256)
    assert np.array_equal(got, base)
    assert prof["proposals"] == p0["proposals"]
    assert prof["late_chains"] == 0
