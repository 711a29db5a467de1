"""ctypes wrapper of the CPU oracle (oracle/smp_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and by the
cpu_baseline leg of bench.py, always as the checker.  The product package never
imports it.

Every function takes numpy int32 arrays laid out like the reference's flat lists
(``pref[row * n + k]``) and returns a man -> woman int32 array.  See
smp_oracle.cpp for the reference functions each one restates (file:line).

Parity status of each function (what pins it, not what restates it):
  gs_seq, mw_seq, gs_par, mw_par  pinned: brute force (n <= 8, every kind),
      closed forms (first choices, serial dictatorship), mutual agreement,
      zero blocking pairs, schedule-independent proposal count.  With rank=
      (ranks given): equal to the same engine without it, and to brute force.
  count_blocking_pairs            pinned: hand-built matchings with known counts.
  brute_force                     pinned: hand-worked n <= 3 cases.
  serial_dictatorship             pinned: equals brute force on congested n <= 8.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS: dict[str, ctypes.CDLL] = {}


class OracleTimes(ctypes.Structure):
    _fields_ = [
        ("precheck_ms", ctypes.c_double),
        ("init_ms", ctypes.c_double),
        ("core_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("proposals", ctypes.c_longlong),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def build() -> None:
    """Compile the oracle variants (g++ only; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(variant: str = "ref") -> ctypes.CDLL:
    name = {"ref": "liboracle.so", "tuned": "liboracle_tuned.so", "cxx": "liboracle_cxx.so"}[variant]
    if variant in _LIBS:
        return _LIBS[variant]
    path = os.path.join(_HERE, "_build", name)
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    i32p = ctypes.POINTER(ctypes.c_int32)
    tp = ctypes.POINTER(OracleTimes)
    for fn in ("oracle_gs_seq", "oracle_mw_seq"):
        getattr(lib, fn).argtypes = [ctypes.c_int, i32p, i32p, i32p, tp]
        getattr(lib, fn).restype = ctypes.c_int
    for fn in ("oracle_gs_par", "oracle_mw_par"):
        getattr(lib, fn).argtypes = [ctypes.c_int, i32p, i32p, ctypes.c_int, i32p, tp]
        getattr(lib, fn).restype = ctypes.c_int
    for fn in ("oracle_gs_seq_ranked", "oracle_mw_seq_ranked"):
        getattr(lib, fn).argtypes = [ctypes.c_int, i32p, i32p, i32p, i32p, tp]
        getattr(lib, fn).restype = ctypes.c_int
    for fn in ("oracle_gs_par_ranked", "oracle_mw_par_ranked"):
        getattr(lib, fn).argtypes = [ctypes.c_int, i32p, i32p, i32p, ctypes.c_int, i32p, tp]
        getattr(lib, fn).restype = ctypes.c_int
    lib.oracle_is_perfect.argtypes = [ctypes.c_int, i32p]
    lib.oracle_is_perfect.restype = ctypes.c_int
    lib.oracle_invert_rows.argtypes = [ctypes.c_int, i32p, i32p]
    lib.oracle_invert_rows.restype = None
    lib.oracle_count_blocking_pairs.argtypes = [ctypes.c_int, i32p, i32p, i32p]
    lib.oracle_count_blocking_pairs.restype = ctypes.c_longlong
    lib.oracle_brute_force.argtypes = [ctypes.c_int, i32p, i32p, i32p]
    lib.oracle_brute_force.restype = ctypes.c_int
    lib.oracle_serial_dictatorship.argtypes = [ctypes.c_int, i32p, i32p, i32p]
    lib.oracle_serial_dictatorship.restype = ctypes.c_int
    _LIBS[variant] = lib
    return lib


def _p(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _check(n: int, pm: np.ndarray, pw: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    pm = np.ascontiguousarray(pm, dtype=np.int32).reshape(-1)
    pw = np.ascontiguousarray(pw, dtype=np.int32).reshape(-1)
    if pm.size != n * n or pw.size != n * n:
        raise ValueError(f"expected {n}x{n} lists")
    return pm, pw


def _rank(n, rank):
    rank = np.ascontiguousarray(rank, dtype=np.int32).reshape(-1)
    if rank.size != n * n:
        raise ValueError(f"expected an {n}x{n} rank matrix")
    return rank


def _run_seq(fn: str, n, pm, pw, variant, rank=None):
    pm, pw = _check(n, pm, pw)
    out = np.full(n, -1, dtype=np.int32)
    t = OracleTimes()
    if rank is None:
        st = getattr(_lib(variant), fn)(n, _p(pm), _p(pw), _p(out), ctypes.byref(t))
    else:
        st = getattr(_lib(variant), fn + "_ranked")(n, _p(pm), _p(pw), _p(_rank(n, rank)), _p(out),
                                                    ctypes.byref(t))
    if st:
        raise RuntimeError(f"{fn} failed ({st})")
    return out, t.as_dict()


def _run_par(fn: str, n, pm, pw, threads, variant, rank=None):
    pm, pw = _check(n, pm, pw)
    out = np.full(n, -1, dtype=np.int32)
    t = OracleTimes()
    if rank is None:
        st = getattr(_lib(variant), fn)(n, _p(pm), _p(pw), int(threads), _p(out), ctypes.byref(t))
    else:
        st = getattr(_lib(variant), fn + "_ranked")(n, _p(pm), _p(pw), _p(_rank(n, rank)), int(threads),
                                                    _p(out), ctypes.byref(t))
    if st:
        raise RuntimeError(f"{fn} failed ({st})")
    return out, t.as_dict()


# rank: the women's inverse rank matrix given as input (rank[w*n + m] = position of
# man m in woman w's list); init then skips the inversion.  None = built in init.
def gs_seq(n, pm, pw, variant="ref", rank=None):
    """GS-Seq (SmpEngineGsF9). Returns (match, times)."""
    return _run_seq("oracle_gs_seq", n, pm, pw, variant, rank)


def mw_seq(n, pm, pw, variant="ref", rank=None):
    """MW-Seq (SmpEngineMwF9). Returns (match, times)."""
    return _run_seq("oracle_mw_seq", n, pm, pw, variant, rank)


def gs_par(n, pm, pw, threads=0, variant="ref", rank=None):
    """GS-Par (SmpEngineGs2F9). Returns (match, times)."""
    return _run_par("oracle_gs_par", n, pm, pw, threads, variant, rank)


def mw_par(n, pm, pw, threads=0, variant="ref", rank=None):
    """MW-Par (SmpEngineMw2F9). Returns (match, times)."""
    return _run_par("oracle_mw_par", n, pm, pw, threads, variant, rank)


def is_perfect(n, pm) -> bool:
    pm = np.ascontiguousarray(pm, dtype=np.int32).reshape(-1)
    return bool(_lib().oracle_is_perfect(n, _p(pm)))


def invert_rows(n, pref) -> np.ndarray:
    pref = np.ascontiguousarray(pref, dtype=np.int32).reshape(-1)
    out = np.empty(n * n, dtype=np.int32)
    _lib().oracle_invert_rows(n, _p(pref), _p(out))
    return out


def count_blocking_pairs(n, pm, pw, match) -> int:
    """-1 if `match` is not a perfect matching, else the number of blocking pairs."""
    pm, pw = _check(n, pm, pw)
    match = np.ascontiguousarray(match, dtype=np.int32)
    return int(_lib().oracle_count_blocking_pairs(n, _p(pm), _p(pw), _p(match)))


def brute_force(n, pm, pw) -> tuple[np.ndarray, int]:
    """Man-optimal stable matching by enumeration (n <= 9): (match, #stable matchings)."""
    pm, pw = _check(n, pm, pw)
    out = np.full(n, -1, dtype=np.int32)
    st = _lib().oracle_brute_force(n, _p(pm), _p(pw), _p(out))
    if st < 0:
        raise RuntimeError(f"brute force failed ({st})")
    return out, st


def serial_dictatorship(n, L, pw) -> np.ndarray:
    """Closed form when every man's list is L."""
    L = np.ascontiguousarray(L, dtype=np.int32).reshape(-1)
    pw = np.ascontiguousarray(pw, dtype=np.int32).reshape(-1)
    out = np.full(n, -1, dtype=np.int32)
    _lib().oracle_serial_dictatorship(n, _p(L), _p(pw), _p(out))
    return out
