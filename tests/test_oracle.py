"""Pins the CPU oracle (oracle/smp_oracle.cpp) against checks that do not restate it.

The reference ships no tests, fixtures or golden vectors for this path, so the
oracle is pinned by: hand-worked cases, brute-force enumeration (n <= 8), closed
forms (distinct first choices; all men sharing one list), the mutual agreement of
four engines with different schedules, the definition of stability, and numpy's
argsort for the rank inversion.  Each planted mistake below is one a plausible
oracle bug would make, and a check here must reject it.
"""
import itertools

import numpy as np
import pytest

ALL_KINDS = range(8)


def flat(rows):
    return np.asarray(rows, dtype=np.int32).reshape(-1)


# ---------------------------------------------------------------- hand-worked
def test_brute_force_hand_cases(oracle):
    # n = 1
    m, cnt = oracle.brute_force(1, flat([[0]]), flat([[0]]))
    assert list(m) == [0] and cnt == 1
    # n = 2, both men want w0 first; w0 prefers m1 -> m1-w0, m0-w1.
    pm = flat([[0, 1], [0, 1]])
    pw = flat([[1, 0], [1, 0]])
    m, cnt = oracle.brute_force(2, pm, pw)
    assert list(m) == [1, 0] and cnt == 1
    # n = 2 with two stable matchings: men-first {m0-w0, m1-w1}, women-first {m0-w1, m1-w0}.
    pm = flat([[0, 1], [1, 0]])
    pw = flat([[1, 0], [0, 1]])
    m, cnt = oracle.brute_force(2, pm, pw)
    assert list(m) == [0, 1] and cnt == 2
    # n = 3 worked example: men m0:[0,1,2] m1:[1,0,2] m2:[0,1,2];
    # women w0:[1,0,2] w1:[0,1,2] w2:[0,1,2].  GS: m0->w0, m1->w1, m2->w0 (rejected, w0 has m0 ranked 1 < 2),
    # m2->w1 (rejected), m2->w2 -> {m0:w0, m1:w1, m2:w2}.
    pm = flat([[0, 1, 2], [1, 0, 2], [0, 1, 2]])
    pw = flat([[1, 0, 2], [0, 1, 2], [0, 1, 2]])
    m, _ = oracle.brute_force(3, pm, pw)
    assert list(m) == [0, 1, 2]


def test_blocking_pair_counts_hand_cases(oracle):
    pm = flat([[0, 1], [0, 1]])
    pw = flat([[0, 1], [0, 1]])
    assert oracle.count_blocking_pairs(2, pm, pw, np.array([0, 1], np.int32)) == 0
    # m0 and w0 prefer each other to their partners: exactly one blocking pair.
    assert oracle.count_blocking_pairs(2, pm, pw, np.array([1, 0], np.int32)) == 1
    # not a perfect matching
    assert oracle.count_blocking_pairs(2, pm, pw, np.array([0, 0], np.int32)) == -1
    assert oracle.count_blocking_pairs(2, pm, pw, np.array([0, 2], np.int32)) == -1
    # n = 3, everyone shares identity lists: the identity matching is the only stable
    # one; the reversed matching {m0:w2, m1:w1, m2:w0} has blocking pairs
    # (m0,w0),(m0,w1),(m1,w0) = 3.
    pm = flat([[0, 1, 2]] * 3)
    pw = flat([[0, 1, 2]] * 3)
    assert oracle.count_blocking_pairs(3, pm, pw, np.array([2, 1, 0], np.int32)) == 3


def test_invert_rows_matches_argsort(oracle, smp):
    n = 37
    _, pw = smp.generate(smp.Workload.UNIFORM, n, seed=3)
    got = oracle.invert_rows(n, pw).reshape(n, n)
    want = np.argsort(pw.reshape(n, n), axis=1, kind="stable")
    assert np.array_equal(got, want)


# -------------------------------------------------------- brute force, n <= 8
def _all_engines(oracle, n, pm, pw):
    out = {}
    out["gs_seq"], t_gs = oracle.gs_seq(n, pm, pw)
    out["mw_seq"], t_mw = oracle.mw_seq(n, pm, pw)
    out["gs_par"], _ = oracle.gs_par(n, pm, pw, threads=3)
    out["mw_par"], _ = oracle.mw_par(n, pm, pw, threads=3)
    return out, t_gs, t_mw


@pytest.mark.parametrize("kind", ALL_KINDS)
def test_engines_equal_brute_force(oracle, smp, kind):
    for n in range(1, 9):
        if kind == smp.Workload.REF_SOLO and n < 4:
            continue
        for seed in range(6):
            pm, pw = smp.generate(kind, n, seed=seed, param=2 if kind == smp.Workload.REF_RANDOM else 0)
            want, n_stable = oracle.brute_force(n, pm, pw)
            assert n_stable >= 1
            got, t_gs, t_mw = _all_engines(oracle, n, pm, pw)
            for name, m in got.items():
                assert np.array_equal(m, want), (name, kind, n, seed)
            # Every Gale-Shapley schedule makes the same proposals: each man down to
            # his final partner once.
            rank_m = np.argsort(pm.reshape(n, n), axis=1)
            expected_props = int(sum(rank_m[i, want[i]] + 1 for i in range(n)))
            assert t_gs["proposals"] == expected_props
            assert t_mw["proposals"] == expected_props


@pytest.mark.parametrize("kind", [0, 1, 3, 4, 7])
def test_ranked_engines_equal_brute_force_and_unranked(oracle, smp, kind):
    """The engines with the women's rank matrix given (numpy argsort of their
    lists, derived here) answer what brute force answers for n <= 8, and what the
    same engine answers from the lists at n = 400."""
    for n in (4, 6, 8, 400):
        pm, pw = smp.generate(kind, n, seed=n + 17, param=2 if kind == smp.Workload.REF_RANDOM else 0)
        rank = np.argsort(pw.reshape(n, n), axis=1).astype(np.int32)
        want = oracle.brute_force(n, pm, pw)[0] if n <= 8 else oracle.gs_seq(n, pm, pw)[0]
        for fn in (oracle.gs_seq, oracle.mw_seq):
            got, t = fn(n, pm, pw, rank=rank)
            assert np.array_equal(got, want), (fn.__name__, kind, n)
            assert t["proposals"] == fn(n, pm, pw)[1]["proposals"]
        for fn in (oracle.gs_par, oracle.mw_par):
            got, _ = fn(n, pm, pw, threads=3, rank=rank)
            assert np.array_equal(got, want), (fn.__name__, kind, n)
    # A transposed rank matrix (the men's ranks where the women's belong) must not pass.
    n = 200
    pm, pw = smp.generate(0, n, seed=5)
    rank = np.argsort(pw.reshape(n, n), axis=1).astype(np.int32)
    bad, _ = oracle.gs_seq(n, pm, pw, rank=np.ascontiguousarray(rank.T))
    assert oracle.count_blocking_pairs(n, pm, pw, bad) != 0


def test_brute_force_on_exhaustive_n3(oracle):
    # Every pair of preference profiles for n = 3 with the men's side fixed to a
    # handful of profiles: engines == brute force on all of them.
    perms = list(itertools.permutations(range(3)))
    men_profiles = [(perms[0], perms[0], perms[0]), (perms[0], perms[3], perms[5]),
                    (perms[1], perms[2], perms[4])]
    for men in men_profiles:
        for women in itertools.product(perms, repeat=3):
            pm, pw = flat(men), flat(women)
            want, _ = oracle.brute_force(3, pm, pw)
            got, _ = oracle.gs_seq(3, pm, pw)
            got2, _ = oracle.mw_seq(3, pm, pw)
            assert np.array_equal(got, want) and np.array_equal(got2, want)


# ---------------------------------------------------------------- closed forms
@pytest.mark.parametrize("kind", [2, 6])  # SHIFTED, REF_PERFECT
def test_first_choice_closed_form(oracle, smp, kind):
    n = 301
    pm, pw = smp.generate(kind, n, seed=11)
    assert oracle.is_perfect(n, pm)
    firsts = pm.reshape(n, n)[:, 0]
    assert len(np.unique(firsts)) == n
    got, _ = oracle.gs_seq(n, pm, pw)
    assert np.array_equal(got, firsts)


def test_is_perfect_false_branch(oracle, smp):
    """IsPerfect (include/smp/smp_engine_bamboo.h:25-33) must say no whenever two
    men share a first choice.  Checked on uniform and congested lists (first
    choices collide with certainty at these sizes, counted here with numpy), and
    on lists built so that row 0 is a permutation while column 0 repeats one
    woman: an oracle that read row 0 (pref_m[m]) instead of column 0
    (pref_m[m * n]) would answer yes on exactly that input."""
    for kind, n in ((smp.Workload.UNIFORM, 300), (smp.Workload.CONGESTED, 300),
                    (smp.Workload.CORRELATED, 257), (smp.Workload.REF_RANDOM, 200)):
        pm, _ = smp.generate(kind, n, seed=31)
        firsts = pm.reshape(n, n)[:, 0]
        assert len(np.unique(firsts)) < n
        assert not oracle.is_perfect(n, pm), kind
    n = 6
    rows = np.empty((n, n), np.int32)
    rows[0] = np.arange(n)                       # row 0 is the identity permutation
    for m in range(1, n):                        # every other man starts with woman 0
        rows[m] = np.roll(np.arange(n), -((m * 5) % n) if m % 2 else 0)
        rows[m, [0, list(rows[m]).index(0)]] = rows[m, [list(rows[m]).index(0), 0]]
    assert (rows[:, 0] == 0).all() and all(sorted(r) == list(range(n)) for r in rows)
    row0_misread = len(set(rows[0, :n].tolist())) == n   # what the planted bug would compute
    assert row0_misread
    assert not oracle.is_perfect(n, flat(rows))
    # One collision among otherwise distinct first choices is enough.
    rows = np.array([np.roll(np.arange(n), -m) for m in range(n)], np.int32)
    assert oracle.is_perfect(n, flat(rows))
    rows[4] = np.roll(np.arange(n), -1)          # man 4 now starts with woman 1, as man 1 does
    assert not oracle.is_perfect(n, flat(rows))


def test_serial_dictatorship_closed_form(oracle, smp):
    for n in (1, 2, 5, 64, 257):
        pm, pw = smp.generate(smp.Workload.CONGESTED, n, seed=n)
        L = pm.reshape(n, n)[0]
        want = oracle.serial_dictatorship(n, L, pw)
        got, t = oracle.gs_seq(n, pm, pw)
        assert np.array_equal(got, want)
        # all men share L: exactly n(n+1)/2 proposals (man matched to L[k] proposed k+1 times)
        assert t["proposals"] == n * (n + 1) // 2
    for n in range(1, 8):
        pm, pw = smp.generate(smp.Workload.CONGESTED, n, seed=100 + n)
        bf, _ = oracle.brute_force(n, pm, pw)
        assert np.array_equal(oracle.serial_dictatorship(n, pm.reshape(n, n)[0], pw), bf)


def test_reference_congested_closed_form(oracle, smp):
    # Reference CONGESTED (men share L, women share P): woman L[k] gets man P[k].
    n = 200
    pm, pw = smp.generate(smp.Workload.REF_CONGESTED, n, seed=5)
    L, P = pm.reshape(n, n)[0], pw.reshape(n, n)[0]
    want = np.empty(n, np.int32)
    want[P] = L
    got, _ = oracle.mw_seq(n, pm, pw)
    assert np.array_equal(got, want)


# ------------------------------------------------------- larger, agreement
@pytest.mark.parametrize("kind,n", [(0, 700), (1, 300), (3, 500), (4, 600), (7, 400)])
def test_engines_agree_and_stable(oracle, smp, kind, n):
    pm, pw = smp.generate(kind, n, seed=42)
    got, t_gs, t_mw = _all_engines(oracle, n, pm, pw)
    ref = got["gs_seq"]
    for name, m in got.items():
        assert np.array_equal(m, ref), name
    assert oracle.count_blocking_pairs(n, pm, pw, ref) == 0
    assert t_gs["proposals"] == t_mw["proposals"]
    for variant in ("tuned", "cxx"):  # the other two builds of the same source
        other, _ = oracle.gs_seq(n, pm, pw, variant=variant)
        assert np.array_equal(other, ref), variant
        other, _ = oracle.mw_par(n, pm, pw, variant=variant, threads=4)
        assert np.array_equal(other, ref), variant


# ------------------------------------------------------------ planted mistakes
def test_planted_swap_is_caught(oracle, smp):
    n = 120
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=9)
    good, _ = oracle.gs_seq(n, pm, pw)
    bad = good.copy()
    bad[[3, 77]] = bad[[77, 3]]  # two men trade partners
    assert oracle.count_blocking_pairs(n, pm, pw, bad) > 0


def test_planted_woman_optimal_is_caught(oracle, smp):
    # Running the proposals from the women's side (a swapped-roles bug) yields a
    # stable matching, so the stability check alone cannot catch it; brute force
    # and the man-optimality criterion do, on instances with several stable matchings.
    caught = 0
    for seed in range(60):
        n = 6
        pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=seed)
        want, n_stable = oracle.brute_force(n, pm, pw)
        w2m, _ = oracle.gs_seq(n, pw, pm)  # women propose
        bad = np.empty(n, np.int32)
        bad[w2m] = np.arange(n, dtype=np.int32)
        assert oracle.count_blocking_pairs(n, pm, pw, bad) == 0
        if n_stable > 1:
            assert not np.array_equal(bad, want)
            caught += 1
    assert caught > 0


def test_planted_transposed_ranks_is_caught(oracle, smp):
    # Using the men's lists where the women's belong (a transposed operand).
    n = 150
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=21)
    bad, _ = oracle.gs_seq(n, pm, pm)
    good, _ = oracle.gs_seq(n, pm, pw)
    assert not np.array_equal(bad, good)
    assert oracle.count_blocking_pairs(n, pm, pw, bad) > 0
