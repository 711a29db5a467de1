"""Seeded generators: byte-exact fixtures, determinism and structure of every kind."""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def test_golden_workload_digests(smp):
    with open(os.path.join(HERE, "golden", "workloads.json")) as f:
        rows = json.load(f)
    for r in rows:
        pm, pw = smp.generate(r["kind"], r["n"], seed=r["seed"], param=r["param"])
        assert digest(pm) == r["sha256_pref_m"], r
        assert digest(pw) == r["sha256_pref_w"], r


@pytest.mark.parametrize("kind", range(8))
def test_thread_count_does_not_change_bytes(smp, kind):
    n = 97
    a = smp.generate(kind, n, seed=5, threads=1)
    b = smp.generate(kind, n, seed=5, threads=7)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    c = smp.generate(kind, n, seed=6, threads=3)
    assert not (np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1]))


@pytest.mark.parametrize("kind", range(8))
def test_rows_are_permutations(smp, kind):
    n = 131
    pm, pw = smp.generate(kind, n, seed=2)
    ident = np.arange(n)
    for M in (pm.reshape(n, n), pw.reshape(n, n)):
        assert np.array_equal(np.sort(M, axis=1), np.broadcast_to(ident, (n, n)))


def test_structure_of_each_kind(smp):
    n = 100
    pm, pw = smp.generate(smp.Workload.CONGESTED, n, seed=1)
    M = pm.reshape(n, n)
    assert (M == M[0]).all()
    assert len({tuple(r) for r in pw.reshape(n, n)}) == n  # women independent
    pm, _ = smp.generate(smp.Workload.SHIFTED, n, seed=1)
    M = pm.reshape(n, n)
    for i in range(n):
        assert np.array_equal(M[i], np.roll(M[0], -i))
    pm, _ = smp.generate(smp.Workload.CORRELATED, n, seed=1)
    M = pm.reshape(n, n)
    base_pos = np.argsort(M[0])
    for i in range(n):
        # each row is the shared base with n/10 = 10 transpositions applied: it
        # differs from any other row in at most 4 * 10 positions
        assert (M[i] != M[(i + 1) % n]).sum() <= 4 * (n // 10)
    pm, pw = smp.generate(smp.Workload.REF_CONGESTED, n, seed=1)
    assert (pm.reshape(n, n) == pm.reshape(n, n)[0]).all()
    assert (pw.reshape(n, n) == pw.reshape(n, n)[0]).all()
    pm, _ = smp.generate(smp.Workload.REF_PERFECT, n, seed=1)
    assert len(np.unique(pm.reshape(n, n)[:, 0])) == n
    pm, pw = smp.generate(smp.Workload.REF_RANDOM, n, seed=1, param=5)
    M = pm.reshape(n, n)
    for g in range(0, n, 5):  # every row holds the same id set inside each group of 5
        s0 = set(M[0, g:g + 5])
        assert all(set(M[i, g:g + 5]) == s0 for i in range(n))
    pm, pw = smp.generate(smp.Workload.REF_SOLO, n, seed=1)
    M, W = pm.reshape(n, n), pw.reshape(n, n)
    assert M[0, 0] == 0 and M[n - 1, n - 1] == n - 1
    assert W[0, 0] == 1 and W[0, 1] == 0


def test_generator_rejects_bad_args(smp):
    with pytest.raises(smp.SmpError):
        smp.generate(99, 10)
    with pytest.raises(smp.SmpError):
        smp.generate(0, 0)
    with pytest.raises(smp.SmpError):
        smp.generate(0, 70000)


def test_batch_plan_is_deterministic_and_log_uniform(smp):
    s1, k1 = smp.batch_plan(4096, 64, 2048, seed=7)
    s2, k2 = smp.batch_plan(4096, 64, 2048, seed=7)
    assert np.array_equal(s1, s2) and np.array_equal(k1, k2)
    assert s1.min() >= 64 and s1.max() <= 2048
    # log-uniform: median near sqrt(64 * 2048) = 362
    assert 300 < np.median(s1) < 440
    frac = [(k1 == k).mean() for k in (0, 1, 3)]
    assert abs(frac[0] - 0.6) < 0.04 and abs(frac[1] - 0.2) < 0.03 and abs(frac[2] - 0.2) < 0.03


def test_workload_text_cache_round_trip(smp, tmp_path):
    # Format of SaveSmpWorkload (generate_workloads_flat.h:302-323): "n\n", n rows of
    # M then n rows of W, every value followed by a space.
    n = 5
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=1)
    path = str(tmp_path / smp.workload_filename(smp.WorkloadType.RANDOM, n, 5))
    smp.save_workload(path, n, pm, pw)
    text = open(path).read().split("\n")
    assert text[0] == "5"
    M, W = pm.reshape(n, n), pw.reshape(n, n)
    for r in range(n):
        assert text[1 + r] == "".join(f"{v} " for v in M[r])
        assert text[1 + n + r] == "".join(f"{v} " for v in W[r])
    n2, pm2, pw2 = smp.load_workload(path)
    assert n2 == n and np.array_equal(pm, pm2) and np.array_equal(pw, pw2)


def test_workload_filenames_and_cache(smp, tmp_path):
    WT = smp.WorkloadType
    assert smp.workload_filename(WT.RANDOM, 30000, 5) == "smp_workload_random_30000_5.txt"
    assert smp.workload_filename(WT.SOLO, 30000) == "smp_workload_solo_30000.txt"
    assert smp.workload_filename(WT.CONGESTED, 7) == "smp_workload_congested_7.txt"
    assert smp.workload_filename(WT.PERFECT, 9) == "smp_workload_perfect_9.txt"
    d = str(tmp_path / "wl")
    a = smp.generate_workload_cached(WT.SOLO, 300, seed=4, directory=d)
    b = smp.generate_workload_cached(WT.SOLO, 300, seed=99, directory=d)  # cache hit: seed unused
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    import shutil
    shutil.copy(os.path.join(d, smp.workload_filename(WT.SOLO, 300)),
                os.path.join(d, smp.workload_filename(WT.SOLO, 301)))
    with pytest.raises(ValueError):
        smp.generate_workload_cached(WT.SOLO, 301, directory=d)
    with pytest.raises(smp.SmpError):
        smp.load_workload(os.path.join(d, "missing.txt"))


def test_workload_load_rejects_bad_files(smp, tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("3\n0 1 2 \n1 2 \n")
    with pytest.raises(smp.SmpError):
        smp.load_workload(str(p))
    p.write_text("x\n")
    with pytest.raises(smp.SmpError):
        smp.load_workload(str(p))
