"""The C-ABI library loads, exports every entry point include/smp_hip.h declares,
and its struct layouts agree with the Python declarations (no GPU calls)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "smp_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s+(smp_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(smp):
    from pactae_amd import _abi
    lib = ctypes.CDLL(_abi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {s[0] for s in _abi.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_package_exposes_its_modules(smp):
    """Both import forms reach the same module (the package imports them lazily)."""
    import pactae_amd
    from pactae_amd import smp as a
    assert pactae_amd.smp is a
    from pactae_amd import build as b
    assert callable(b.build)


def test_struct_layouts_match_header(tmp_path):
    from pactae_amd import _abi
    src = tmp_path / "sz.c"
    src.write_text('#include "smp_hip.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(smp_opts), sizeof(smp_times),'
                   ' sizeof(smp_batch_opts), sizeof(smp_batch_times), offsetof(smp_times, proposals),'
                   ' offsetof(smp_times, is_perfect), offsetof(smp_opts, init_mode),'
                   ' offsetof(smp_times, bucket_kernel_ms), offsetof(smp_times, sparse_init),'
                   ' offsetof(smp_times, rank_matrix)); return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    want = [ctypes.sizeof(_abi.SmpOpts), ctypes.sizeof(_abi.SmpTimes), ctypes.sizeof(_abi.SmpBatchOpts),
            ctypes.sizeof(_abi.SmpBatchTimes), _abi.SmpTimes.proposals.offset,
            _abi.SmpTimes.is_perfect.offset, _abi.SmpOpts.init_mode.offset,
            _abi.SmpTimes.bucket_kernel_ms.offset, _abi.SmpTimes.sparse_init.offset,
            _abi.SmpTimes.rank_matrix.offset]
    assert [int(x) for x in out] == want


def test_errors_are_reported_not_thrown(smp):
    # Host-side argument checks return a status and a message; no device needed.
    import numpy as np
    with pytest.raises(smp.SmpError) as ei:
        smp.SmpEngineBamboo(smp.SmpObj.create_from_flat_lists(np.zeros(0, np.int32),
                                                              np.zeros(0, np.int32), 0), 0)
    assert ei.value.status == 1
    assert "n must be" in str(ei.value)


def test_unknown_options_are_rejected(smp):
    # Option values are checked before any device is touched.
    import numpy as np
    n = 4
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=1)
    obj = smp.SmpObj.create_from_flat_lists(pm, pw, n)
    for bad in (dict(init_mode=3), dict(init_mode=-1), dict(schedule=6), dict(schedule=-1),
                dict(tail_window_entries=-1), dict(start_order=2), dict(start_order=-1)):
        with pytest.raises(smp.SmpError) as ei:
            smp.SmpEngineBamboo(obj, n, **bad)
        assert ei.value.status == 1, bad
        assert "init_mode, schedule or start_order" in str(ei.value)
    assert [m.value for m in smp.InitMode] == [0, 1, 2]


def test_version(smp):
    assert "gfx950" in smp.version()


def test_server_struct_layouts(tmp_path):
    from pactae_amd import _abi
    src = tmp_path / "sz2.c"
    src.write_text('#include "smp_hip.h"\n#include <stdio.h>\n'
                   'int main(void){printf("%zu %zu %zu\\n", sizeof(smp_server_opts), sizeof(smp_request_info),'
                   ' sizeof(smp_server_stats)); return 0;}\n')
    exe = tmp_path / "sz2"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert [int(x) for x in out] == [ctypes.sizeof(_abi.SmpServerOpts), ctypes.sizeof(_abi.SmpRequestInfo),
                                     ctypes.sizeof(_abi.SmpServerStats)]
