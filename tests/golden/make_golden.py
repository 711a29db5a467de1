"""Writes the golden fixtures under tests/golden/.

  workloads.json   sha256 of the generator output for a grid of (kind, n, seed, param):
                   pins the seeded generators byte for byte across hosts and versions.
  matchings.json   small instances (n = 4..64, every generator) with their man-optimal
                   matching from the oracle's GS-Seq, cross-checked here against
                   MW-Seq, zero blocking pairs, and brute force for n <= 8.

The reference repository has no fixtures or golden vectors for this path (it has
no tests at all), so these are produced by this repository's own generator and
oracle; the cross-checks above are what make them trustworthy.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import pactae_amd.smp as smp  # noqa: E402
from oracle import oracle  # noqa: E402

WORKLOAD_GRID = [(k, n, s, p) for k in range(8) for n in (4, 9, 64, 300) for s in (0, 1)
                 for p in ((0, 3) if k in (3, 4) else (0,))]
MATCH_GRID = [(k, n, s) for k in range(8) for n in (4, 7, 8, 16, 33, 64) for s in (0, 1, 2)]


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def main() -> None:
    wl = []
    for kind, n, seed, param in WORKLOAD_GRID:
        pm, pw = smp.generate(kind, n, seed=seed, param=param)
        wl.append({"kind": kind, "n": n, "seed": seed, "param": param,
                   "sha256_pref_m": digest(pm), "sha256_pref_w": digest(pw)})
    with open(os.path.join(HERE, "workloads.json"), "w") as f:
        json.dump(wl, f, indent=1)

    ms = []
    for kind, n, seed in MATCH_GRID:
        pm, pw = smp.generate(kind, n, seed=seed)
        m_gs, t = oracle.gs_seq(n, pm, pw)
        m_mw, _ = oracle.mw_seq(n, pm, pw)
        assert np.array_equal(m_gs, m_mw)
        assert oracle.count_blocking_pairs(n, pm, pw, m_gs) == 0
        if n <= 8:
            bf, _ = oracle.brute_force(n, pm, pw)
            assert np.array_equal(bf, m_gs)
        ms.append({"kind": kind, "n": n, "seed": seed,
                   "pref_m": pm.tolist() if n <= 16 else None,
                   "pref_w": pw.tolist() if n <= 16 else None,
                   "sha256_pref_m": digest(pm), "sha256_pref_w": digest(pw),
                   "match": m_gs.tolist(), "proposals": int(t["proposals"])})
    with open(os.path.join(HERE, "matchings.json"), "w") as f:
        json.dump(ms, f, indent=0)
    print(f"wrote {len(wl)} workload digests and {len(ms)} matchings")


if __name__ == "__main__":
    main()
