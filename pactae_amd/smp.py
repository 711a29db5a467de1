"""Python mirror of the reference's SMP engine interface over the C ABI.

Reference shapes (paths relative to the reference repository root):
  WorkloadType                 include/utils/types.h:85-90
  GenerateWorkload             include/utils/generate_workloads_flat.h:250-280
  SmpObj::CreateFromFlatLists  include/smp/smp.h:25-29
  SmpEngineBamboo(smp, n)      include/smp/smp_engine_bamboo.h:13-119
  FindStableMatching()         include/smp/smp_engine_abs.h:10-36
  PrintProfilingInfo()         include/smp/smp_engine_abs.h:38-52
  SmpEngineExecutor<E>(...)    benchmark/smp_engine_exec.h:6-32

Errors from the library raise SmpError carrying the C status and message.
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass

import numpy as np

from . import _abi

_lib = _abi.load()


class SmpError(RuntimeError):
    def __init__(self, status: int, where: str):
        msg = _lib.smp_last_error().decode(errors="replace")
        super().__init__(f"{where}: status {status}: {msg}")
        self.status = status


def _check(st: int, where: str) -> None:
    if st != _abi.SMP_OK:
        raise SmpError(st, where)


def _i32p(a: np.ndarray):
    return a.ctypes.data_as(_abi.i32p)


class Workload(enum.IntEnum):
    """Generators of include/smp_hip.h (enum smp_workload)."""
    UNIFORM = 0
    CONGESTED = 1
    SHIFTED = 2
    CORRELATED = 3
    REF_RANDOM = 4
    REF_CONGESTED = 5
    REF_PERFECT = 6
    REF_SOLO = 7


class WorkloadType(enum.IntEnum):
    """The reference's WorkloadType (types.h:85-90), mapped to generators below."""
    PERFECT = 0
    RANDOM = 1
    CONGESTED = 2
    SOLO = 3


_REF_KIND = {
    WorkloadType.PERFECT: Workload.REF_PERFECT,
    WorkloadType.RANDOM: Workload.REF_RANDOM,
    WorkloadType.CONGESTED: Workload.REF_CONGESTED,
    WorkloadType.SOLO: Workload.REF_SOLO,
}


def generate(kind: int, n: int, seed: int = 0, param: int = 0, threads: int = 0):
    """Seeded generator -> (pref_m, pref_w), each int32[n*n] (flat, row-major)."""
    pm = np.empty(n * n, dtype=np.int32)
    pw = np.empty(n * n, dtype=np.int32)
    _check(_lib.smp_generate(int(kind), n, seed & (2**64 - 1), param, _i32p(pm), _i32p(pw), threads),
           "smp_generate")
    return pm, pw


def generate_workload(workload_type: WorkloadType, n: int, group_size: int = 5, seed: int = 0):
    """GenerateWorkload(type, n, M, W, group) with the reference's four workload types."""
    return generate(_REF_KIND[WorkloadType(workload_type)], n, seed, group_size)


def workload_filename(workload_type: WorkloadType, n: int, group_size: int = 5) -> str:
    """GetWorkloadFilename (generate_workloads_flat.h:285-300)."""
    buf = ctypes.create_string_buffer(128)
    _check(_lib.smp_workload_filename(int(workload_type), n, group_size, buf, len(buf)),
           "smp_workload_filename")
    return buf.value.decode()


def save_workload(path: str, n: int, pm, pw) -> None:
    """SaveSmpWorkload (generate_workloads_flat.h:302-323)."""
    pm = np.ascontiguousarray(pm, dtype=np.int32).reshape(-1)
    pw = np.ascontiguousarray(pw, dtype=np.int32).reshape(-1)
    _check(_lib.smp_workload_save(path.encode(), n, _i32p(pm), _i32p(pw)), "smp_workload_save")


def load_workload(path: str):
    """LoadSmpWorkload (generate_workloads_flat.h:325-352) -> (n, pref_m, pref_w)."""
    n = ctypes.c_int32(0)
    _check(_lib.smp_workload_load(path.encode(), ctypes.byref(n), None, None, 0), "smp_workload_load")
    nn = n.value * n.value
    pm = np.empty(nn, dtype=np.int32)
    pw = np.empty(nn, dtype=np.int32)
    _check(_lib.smp_workload_load(path.encode(), ctypes.byref(n), _i32p(pm), _i32p(pw), nn),
           "smp_workload_load")
    return n.value, pm, pw


def generate_workload_cached(workload_type: WorkloadType, n: int, group_size: int = 5, seed: int = 0,
                             directory: str = "data/workloads"):
    """GenerateWorkloadCached (generate_workloads_flat.h:355-376): load the cached file if
    present, else generate, save and return.  Returns (pref_m, pref_w)."""
    import os
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, workload_filename(workload_type, n, group_size))
    if os.path.exists(path):
        n_file, pm, pw = load_workload(path)
        if n_file != n:
            raise ValueError(f"cached workload {path} holds n={n_file}, asked for n={n}")
        return pm, pw
    pm, pw = generate_workload(workload_type, n, group_size, seed)
    save_workload(path, n, pm, pw)
    return pm, pw


def batch_plan(count: int, n_min: int, n_max: int, seed: int, mix=(0.6, 0.2, 0.2)):
    sizes = np.empty(count, dtype=np.int32)
    kinds = np.empty(count, dtype=np.int32)
    mix_arr = (ctypes.c_double * 3)(*mix)
    _check(_lib.smp_batch_plan(count, n_min, n_max, seed & (2**64 - 1), mix_arr, _i32p(sizes),
                               _i32p(kinds)), "smp_batch_plan")
    return sizes, kinds


@dataclass
class SmpObj:
    """Owns the flat lists (SmpObj, smp.h:8-52)."""
    flatten_pref_lists_m: np.ndarray
    flatten_pref_lists_w: np.ndarray
    n: int

    @staticmethod
    def create_from_flat_lists(flat_m, flat_w, n: int) -> "SmpObj":
        m = np.ascontiguousarray(flat_m, dtype=np.int32).reshape(-1).copy()
        w = np.ascontiguousarray(flat_w, dtype=np.int32).reshape(-1).copy()
        if m.size != n * n or w.size != n * n:
            raise ValueError(f"flat lists must hold n*n = {n * n} entries")
        return SmpObj(m, w, n)


class DeviceArray:
    """int32 buffer in HBM allocated through the library (no other GPU runtime)."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=np.int32)
        self.nbytes = host.nbytes
        self.ptr = ctypes.c_void_p()
        _check(_lib.smp_device_alloc(self.nbytes, ctypes.byref(self.ptr)), "smp_device_alloc")
        _check(_lib.smp_copy_h2d(self.ptr, host.ctypes.data_as(ctypes.c_void_p), self.nbytes),
               "smp_copy_h2d")

    def to_host(self) -> np.ndarray:
        out = np.empty(self.nbytes // 4, dtype=np.int32)
        _check(_lib.smp_copy_d2h(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes),
               "smp_copy_d2h")
        return out

    def free(self) -> None:
        if self.ptr:
            _lib.smp_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def default_opts(**overrides) -> _abi.SmpOpts:
    o = _abi.SmpOpts()
    _lib.smp_default_opts(ctypes.byref(o))
    for k, v in overrides.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, int(v))
    return o


class Schedule(enum.IntEnum):
    """enum smp_schedule: this build's pipeline; the reference's thread-per-man
    GPU walks (MW-Par-GPU-CAS, SmpEngineMw3F9, src/smp/smp_engine_mw_3_f9.cu:37-97,
    and the CAS/MIN variants of src/smp/kernel.cu:5-55, 122-165, 468-512,
    515-552); or the frontier rounds with LDS pre-reduction per woman."""
    BAMBOO = _abi.SMP_SCHEDULE_BAMBOO
    MW_CAS = _abi.SMP_SCHEDULE_MW_CAS
    LA_CAS = _abi.SMP_SCHEDULE_LA_CAS
    MW_MIN = _abi.SMP_SCHEDULE_MW_MIN
    LA_MIN = _abi.SMP_SCHEDULE_LA_MIN
    FRONTIER = _abi.SMP_SCHEDULE_FRONTIER


class InitMode(enum.IntEnum):
    """enum smp_init_mode: how init builds the men's PR prefix.  FULL is the
    reference's init (n x n rank matrix, then the gather, kernel.cu:672-709);
    SPARSE looks up only the prefix pairs in the women's rows and builds the rank
    matrix only when a chain needs it.  Results are identical in every mode."""
    AUTO = _abi.SMP_INIT_AUTO
    FULL = _abi.SMP_INIT_FULL
    SPARSE = _abi.SMP_INIT_SPARSE


class SmpEngineBamboo:
    """MI355X BambooSMP engine (SmpEngineBamboo, smp_engine_bamboo.h:13-119).

    ``SmpEngineBamboo(smp, n)`` copies the host lists into HBM like the reference
    constructor.  ``SmpEngineBamboo.from_device(d_m, d_w, n)`` borrows lists that
    are already resident.  ``from_ranked`` and ``from_device_ranked`` take the
    women's inverse rank matrix (``rank_w[w, m]`` = position of man m in woman w's
    list) instead of the women's lists; init is then the prefix gather alone.
    """

    name = "BambooSMP-MI355X"

    def __init__(self, smp: SmpObj | None, n: int, **opts):
        self._h = ctypes.c_void_p()
        self.n = n
        self._keep = None
        if smp is None:
            return
        o = default_opts(**opts)
        _check(_lib.smp_create(_i32p(smp.flatten_pref_lists_m), _i32p(smp.flatten_pref_lists_w), n,
                               ctypes.byref(o), ctypes.byref(self._h)), "smp_create")

    @classmethod
    def from_device(cls, d_pref_m: DeviceArray, d_pref_w: DeviceArray, n: int, **opts):
        e = cls(None, n)
        o = default_opts(**opts)
        _check(_lib.smp_create_device(d_pref_m.ptr, d_pref_w.ptr, n, ctypes.byref(o),
                                      ctypes.byref(e._h)), "smp_create_device")
        e._keep = (d_pref_m, d_pref_w)
        return e

    @classmethod
    def from_ranked(cls, pref_m: np.ndarray, rank_w: np.ndarray, n: int, **opts):
        e = cls(None, n)
        o = default_opts(**opts)
        pm = np.ascontiguousarray(pref_m, dtype=np.int32).reshape(-1)
        rw = np.ascontiguousarray(rank_w, dtype=np.int32).reshape(-1)
        if pm.size != n * n or rw.size != n * n:
            raise ValueError("pref_m and rank_w must hold n*n entries")
        _check(_lib.smp_create_ranked(_i32p(pm), _i32p(rw), n, ctypes.byref(o), ctypes.byref(e._h)),
               "smp_create_ranked")
        return e

    @classmethod
    def from_device_ranked(cls, d_pref_m: DeviceArray, d_rank_w: DeviceArray, n: int, **opts):
        e = cls(None, n)
        o = default_opts(**opts)
        _check(_lib.smp_create_device_ranked(d_pref_m.ptr, d_rank_w.ptr, n, ctypes.byref(o),
                                             ctypes.byref(e._h)), "smp_create_device_ranked")
        e._keep = (d_pref_m, d_rank_w)
        return e

    def exec(self) -> None:
        _check(_lib.smp_exec(self._h), "smp_exec")

    def result(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        _check(_lib.smp_result(self._h, _i32p(out)), "smp_result")
        return out

    def find_stable_matching(self) -> np.ndarray:
        """FindStableMatching(): one full solve, returns the man -> woman Matching."""
        self.exec()
        return self.result()

    def profile(self) -> dict:
        t = _abi.SmpTimes()
        _check(_lib.smp_profile(self._h, ctypes.byref(t)), "smp_profile")
        return t.as_dict()

    def print_profiling_info(self) -> None:
        """Same lines as AbsSmpEngine::PrintProfilingInfo (smp_engine_abs.h:38-44)."""
        p = self.profile()
        print(f"{self.name} Profiling Info:")
        print(f"{self.name} Precheck Time: {p['precheck_ms']} ms")
        print(f"{self.name} Init Time: {p['init_ms']} ms")
        print(f"{self.name} Core Time: {p['core_ms']} ms")
        print(f"{self.name} Total Time: {p['precheck_ms'] + p['init_ms'] + p['core_ms']} ms")
        print("=" * 66)

    def init_only(self, init_pr: bool = True) -> dict:
        """SmpInitEngine::Init (smp_init_abs.h:16-28): rank matrix [+ PR matrix] + D2H."""
        t = _abi.SmpInitTimes()
        _check(_lib.smp_init_exec(self._h, int(init_pr), ctypes.byref(t)), "smp_init_exec")
        return t.as_dict()

    def tune(self, trials: int = 3) -> dict:
        """Measure the hand-off threshold and prefix width (smp_tune); AUTO mode uses them."""
        r = _abi.SmpTuneResult()
        _check(_lib.smp_tune(self._h, trials, ctypes.byref(r)), "smp_tune")
        return r.as_dict()

    def rank_matrix(self) -> np.ndarray:
        out = np.empty(self.n * self.n, dtype=np.int32)
        _check(_lib.smp_rank_matrix(self._h, _i32p(out)), "smp_rank_matrix")
        return out.reshape(self.n, self.n)

    def pr_matrix(self) -> np.ndarray:
        """PR matrix as PRNode pairs: array [n, cols, 2] of (woman, rank)."""
        c = ctypes.c_int32(0)
        _check(_lib.smp_pr_matrix(self._h, None, 0, ctypes.byref(c)), "smp_pr_matrix")
        out = np.empty(self.n * c.value * 2, dtype=np.int32)
        _check(_lib.smp_pr_matrix(self._h, _i32p(out), c.value, ctypes.byref(c)), "smp_pr_matrix")
        return out.reshape(self.n, c.value, 2)

    def close(self) -> None:
        if self._h:
            _lib.smp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def smp_engine_executor(workload_type: WorkloadType, workload_size: int, group_size: int = 5,
                        seed: int = 0, **opts) -> tuple[np.ndarray, dict]:
    """SmpEngineExecutor<SmpEngineBamboo> (smp_engine_exec.h:6-32): generate, solve, report."""
    pm, pw = generate_workload(workload_type, workload_size, group_size, seed)
    smp = SmpObj.create_from_flat_lists(pm, pw, workload_size)
    with SmpEngineBamboo(smp, workload_size, **opts) as eng:
        extra = f", group_size: {group_size}" if workload_type == WorkloadType.RANDOM else ""
        print(f"Executing {eng.name} with workload: {WorkloadType(workload_type).name.title()}, "
              f"size: {workload_size}{extra}")
        match = eng.find_stable_matching()
        eng.print_profiling_info()
        return match, eng.profile()


class SmpBatch:
    """Many independent instances solved in one persistent launch (smp_batch_*)."""

    AUTO = _abi.SMP_BATCH_AUTO
    FUSED = _abi.SMP_BATCH_FUSED
    POOL = _abi.SMP_BATCH_POOL
    WOMEN_ROWS = _abi.SMP_RANK_WOMEN_ROWS
    MEN_ROWS = _abi.SMP_RANK_MEN_ROWS

    def __init__(self, sizes, pref_m_list, pref_w_list, **opts):
        sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        self.count = int(sizes.size)
        self.sizes = sizes
        sq = sizes.astype(np.int64) ** 2
        self.offsets = np.zeros(self.count, dtype=np.int64)
        if self.count > 1:
            self.offsets[1:] = np.cumsum(sq)[:-1]
        self.moffsets = np.zeros(self.count, dtype=np.int64)
        if self.count > 1:
            self.moffsets[1:] = np.cumsum(sizes.astype(np.int64))[:-1]
        if isinstance(pref_m_list, np.ndarray) and pref_m_list.ndim == 1:
            pm_all = np.ascontiguousarray(pref_m_list, dtype=np.int32)
            pw_all = np.ascontiguousarray(pref_w_list, dtype=np.int32)
        else:
            pm_all = np.concatenate([np.asarray(a, dtype=np.int32).reshape(-1) for a in pref_m_list])
            pw_all = np.concatenate([np.asarray(a, dtype=np.int32).reshape(-1) for a in pref_w_list])
        if pm_all.size != int(sq.sum()) or pw_all.size != int(sq.sum()):
            raise ValueError("concatenated lists do not match sizes")
        o = _abi.SmpBatchOpts()
        _lib.smp_batch_default_opts(ctypes.byref(o))
        for k, v in opts.items():
            setattr(o, k, int(v))
        self._h = ctypes.c_void_p()
        _check(_lib.smp_batch_create(self.count, _i32p(sizes),
                                     self.offsets.ctypes.data_as(_abi.i64p), _i32p(pm_all),
                                     _i32p(pw_all), ctypes.byref(o), ctypes.byref(self._h)),
               "smp_batch_create")

    def exec(self) -> None:
        _check(_lib.smp_batch_exec(self._h), "smp_batch_exec")

    def exec_looped(self) -> None:
        _check(_lib.smp_batch_exec_looped(self._h), "smp_batch_exec_looped")

    def result(self) -> list[np.ndarray]:
        out = np.empty(int(self.sizes.sum()), dtype=np.int32)
        _check(_lib.smp_batch_result(self._h, _i32p(out)), "smp_batch_result")
        return [out[o:o + n] for o, n in zip(self.moffsets, self.sizes)]

    def profile(self) -> dict:
        t = _abi.SmpBatchTimes()
        _check(_lib.smp_batch_profile(self._h, ctypes.byref(t)), "smp_batch_profile")
        return t.as_dict()

    def close(self) -> None:
        if self._h:
            _lib.smp_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SmpServer:
    """Request-serving loop (smp_server_*): continuous batching or the plain baseline.

    submit() never blocks on the GPU; wait()/poll() return (match, info)."""

    BATCHED = _abi.SMP_SERVE_BATCHED
    PLAIN = _abi.SMP_SERVE_PLAIN

    def __init__(self, **opts):
        o = _abi.SmpServerOpts()
        _lib.smp_server_default_opts(ctypes.byref(o))
        for k, v in opts.items():
            setattr(o, k, int(v))
        self.max_n = o.max_n
        self._h = ctypes.c_void_p()
        _check(_lib.smp_server_create(ctypes.byref(o), ctypes.byref(self._h)), "smp_server_create")
        self._n = {}
        self._keep = {}

    def submit(self, pref_m, pref_w, n: int) -> int:
        """Host lists (numpy int32, copied) or DeviceArray pairs (borrowed)."""
        t = ctypes.c_uint64()
        if isinstance(pref_m, DeviceArray):
            st = _lib.smp_server_submit(self._h, pref_m.ptr, pref_w.ptr, n, 1, ctypes.byref(t))
            keep = (pref_m, pref_w)
        else:
            pm = np.ascontiguousarray(pref_m, dtype=np.int32).reshape(-1)
            pw = np.ascontiguousarray(pref_w, dtype=np.int32).reshape(-1)
            st = _lib.smp_server_submit(self._h, pm.ctypes.data_as(ctypes.c_void_p),
                                        pw.ctypes.data_as(ctypes.c_void_p), n, 0, ctypes.byref(t))
            keep = None
        _check(st, "smp_server_submit")
        self._n[t.value] = n
        if keep is not None:
            self._keep[t.value] = keep
        return t.value

    def _finish(self, ticket, st, out, info):
        self._n.pop(ticket, None)
        self._keep.pop(ticket, None)
        _check(st, "smp_server request")
        return out, info.as_dict()

    def poll(self, ticket: int):
        out = np.empty(self._n[ticket], dtype=np.int32)
        info = _abi.SmpRequestInfo()
        st = _lib.smp_server_poll(self._h, ticket, _i32p(out), ctypes.byref(info))
        if st == _abi.SMP_PENDING:
            return None
        return self._finish(ticket, st, out, info)

    def wait(self, ticket: int, timeout_ms: int = -1):
        out = np.empty(self._n[ticket], dtype=np.int32)
        info = _abi.SmpRequestInfo()
        st = _lib.smp_server_wait(self._h, ticket, _i32p(out), ctypes.byref(info), timeout_ms)
        return self._finish(ticket, st, out, info)

    def stats(self) -> dict:
        s = _abi.SmpServerStats()
        _check(_lib.smp_server_stats_get(self._h, ctypes.byref(s)), "smp_server_stats_get")
        return s.as_dict()

    def close(self) -> None:
        if self._h:
            _lib.smp_server_destroy(self._h)
            self._h = ctypes.c_void_p()
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_copy(dst: DeviceArray, src: DeviceArray) -> None:
    """Streaming device-to-device copy by the library's own kernel."""
    _check(_lib.smp_device_copy(dst.ptr, src.ptr, min(dst.nbytes, src.nbytes) & ~15), "smp_device_copy")


def device_count() -> int:
    c = ctypes.c_int32(0)
    st = _lib.smp_device_count(ctypes.byref(c))
    return int(c.value) if st == _abi.SMP_OK else 0


def version() -> str:
    return _lib.smp_version().decode()
