"""Parity sweep over the host tail's and the proposal kernel's less common routes,
every result bit-exact against the oracle's GS-Seq (tolerance 0, integer work):

  - extension windows (narrow prefixes, so that many chains leave them),
  - windows handed out ahead of time (a rank matrix present: full init),
  - bands of columns and the whole PR matrix (correlated and SOLO lists),
  - window storage that runs out (chains handed back to the GPU),
  - the restart over the whole PR matrix (the reference's RANDOM lists),

on every generator kind at n from 300 to 3,000, with several hand-off
thresholds.  The configurations are fixed (seeded), so a failure names its case.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KINDS = [0, 1, 3, 4, 5, 6, 7]  # uniform, congested, correlated, RANDOM(5), CONGESTED, PERFECT, SOLO

CASES = []
_rng = np.random.default_rng(0x7A11)
for kind in KINDS:
    for n in (300, 1200, 3000):
        for cfg in (
            dict(prefix_cols=32, tail_mode=2, tail_threshold=int(_rng.integers(8, 200))),
            dict(prefix_cols=64, init_mode=1, tail_mode=2, tail_threshold=int(_rng.integers(8, 200))),
            dict(prefix_cols=32, tail_mode=2, tail_threshold=500, tail_window_entries=int(_rng.integers(64, 5000))),
            dict(tail_mode=2, tail_threshold=int(_rng.integers(16, 128))),
            dict(tail_mode=1),
        ):
            CASES.append((kind, n, cfg))


def _solve(smp, n, pm, pw, opts, repeats):
    obj = smp.SmpObj.create_from_flat_lists(pm, pw, n)
    out = []
    with smp.SmpEngineBamboo(obj, n, precheck=0, **opts) as e:
        for _ in range(repeats):
            e.exec()
            out.append((e.result().copy(), e.profile()))
    return out


@pytest.mark.parametrize("kind,n,cfg", CASES, ids=[f"k{k}-n{n}-{i % 5}" for i, (k, n, _) in enumerate(CASES)])
def test_tail_routes(gpu, oracle, kind, n, cfg):
    smp = gpu
    seed = 0x5EED1000 + 97 * kind + n
    pm, pw = smp.generate(kind, n, seed=seed, param=5 if kind == 4 else 0)
    want, t = oracle.gs_seq(n, pm, pw)
    # Two solves on one engine: the second one starts from what the first
    # learned about these lists (prefix width, whole PR matrix).
    for i, (got, prof) in enumerate(_solve(smp, n, pm, pw, cfg, 2)):
        assert np.array_equal(got, want), (kind, n, seed, cfg, i)
        assert prof["proposals"] == t["proposals"], (kind, n, seed, cfg, i)
