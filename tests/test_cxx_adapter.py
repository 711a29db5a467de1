"""The C++ adapter a maintainer would add to the reference (tests/cxx/
smp_engine_bamboo_mi355x.h, the code INTEGRATION.md shows), compiled with g++
against include/smp_hip.h and linked to libsmp_hip.so, driven the way the
reference's executor drives an engine (tests/cxx/adapter_driver.cpp).

CPU: the adapter and driver compile and link.  GPU: the driver's matchings equal
the committed golden fixtures and the oracle's, and the framework's "perfect
matching" line appears exactly for the instance whose first choices differ."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))


def build_driver(tmp_path):
    from pactae_amd import _abi
    exe = tmp_path / "adapter_driver"
    lib_dir = os.path.dirname(_abi.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(HERE, "cxx"), os.path.join(HERE, "cxx", "adapter_driver.cpp"),
                    "-o", str(exe), "-L", lib_dir, "-lsmp_hip", f"-Wl,-rpath,{lib_dir}"], check=True)
    return exe


def test_adapter_compiles_and_links(tmp_path, smp):
    exe = build_driver(tmp_path)
    assert exe.exists()


def run_driver(exe, tmp_path, n, pm, pw, tag):
    lists = tmp_path / f"{tag}.lists"
    out = tmp_path / f"{tag}.match"
    np.concatenate([np.asarray(pm, np.int32).ravel(), np.asarray(pw, np.int32).ravel()]).tofile(lists)
    r = subprocess.run([str(exe), str(n), str(lists), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.fromfile(out, dtype=np.int32), r.stdout


@pytest.mark.gpu
def test_adapter_matches_oracle_and_fixtures(gpu, oracle, tmp_path):
    smp = gpu
    exe = build_driver(tmp_path)
    perfect_line = "Perfect matching identified; skipping execution steps."
    n = 2000
    pm, pw = smp.generate(smp.Workload.UNIFORM, n, seed=0x5EED0001)
    got, out = run_driver(exe, tmp_path, n, pm, pw, "uniform")
    assert np.array_equal(got, oracle.gs_seq(n, pm, pw)[0])
    assert perfect_line not in out and "library: is_perfect 0" in out
    assert "BambooSMP-MI355X Total Time:" in out
    pm, pw = smp.generate(smp.Workload.SHIFTED, n, seed=0x5EED0003)
    got, out = run_driver(exe, tmp_path, n, pm, pw, "shifted")
    assert np.array_equal(got, pm.reshape(n, n)[:, 0])
    assert perfect_line in out and "library: is_perfect 1" in out
    with open(os.path.join(HERE, "golden", "matchings.json")) as f:
        rows = json.load(f)
    for i, r in enumerate(rows[::12]):
        if r["pref_m"] is not None:  # small fixtures hold their lists
            pm, pw = np.asarray(r["pref_m"], np.int32), np.asarray(r["pref_w"], np.int32)
        else:  # the others their generator's arguments (digests pinned in test_workloads.py)
            pm, pw = smp.generate(r["kind"], r["n"], seed=r["seed"])
        got, _ = run_driver(exe, tmp_path, r["n"], pm, pw, f"golden{i}")
        assert got.tolist() == r["match"], (r["kind"], r["n"], r["seed"])
