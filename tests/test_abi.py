"""The C-ABI library loads and exports every function include/*.h declares; calls
that need no GPU behave (argument checks, no-device status, the host-only reader
and RNG jump)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from ffea_ci_demo_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z_0-9 ]*?\**\s*\b(ffea_[a-z_0-9]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 25
    L = _native.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_invalid_arguments_rejected_without_device():
    L = _native.lib()
    assert L.ffea_gpu_blob_step(None, 1) == _native.ERR_INVALID
    assert L.ffea_gpu_blob_update_positions(None) == _native.ERR_INVALID
    L.ffea_gpu_blob_destroy(None)


def test_rng_advance_matches_oracle(oracle_mod):
    L = _native.lib()
    L.ffea_gpu_advance_rng_state.argtypes = [np.ctypeslib.ndpointer(np.uint32), C.c_longlong,
                                             np.ctypeslib.ndpointer(np.uint32)]
    s0 = np.array([11, 22, 33, 44, 55, 66], np.uint32)
    out = np.zeros(6, np.uint32)
    for n in (0, 1, 7, 7 * 1296, 123456789):
        assert L.ffea_gpu_advance_rng_state(s0, n, out) == 0
        g = oracle_mod.Rng(state=s0)
        if n:
            g.advance(0, n) if n < 2 ** 31 else None
        assert np.array_equal(out, g.state())


def test_cpp_wrapper_and_example_build(tmp_path):
    """include/ffea_gpu.hpp and the script-runner example compile and link against the
    library (run on the GPU in tests/test_gpu_cpp.py)"""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    for name in ("run_script", "solver_drop_in"):
        out = str(tmp_path / name)
        cmd = [gxx, "-std=c++17", "-I", os.path.join(ROOT, "include"),
               os.path.join(ROOT, "ffea_ci_demo_amd", "examples", name + ".cpp"),
               "-L", os.path.join(ROOT, "ffea_ci_demo_amd"), "-lffea_gpu",
               "-Wl,-rpath," + os.path.join(ROOT, "ffea_ci_demo_amd"), "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_rng_stream_state_matches_oracle_streams(oracle_mod):
    """ffea_gpu_rng_stream_state(seed, k) is the k-th RngStream after SetPackageSeed(seed)
    (streams 2^127 draws apart); the oracle's package streams are pinned to RngStream"""
    for seed in ([12345] * 6, [1, 2, 3, 4, 5, 6], [4294967086, 7, 9, 4294944442, 11, 1]):
        pk = oracle_mod.Package()
        assert pk.set_seed(seed)
        streams = [oracle_mod.Rng(package=pk) for _ in range(6)]
        for k, g in enumerate(streams):
            assert np.array_equal(_native.rng_stream_state(seed, k), g.state()), (seed, k)
    # L'Ecuyer's published second stream of the default package seed 12345 x 6
    assert list(_native.rng_stream_state([12345] * 6, 1)) == [3692455944, 1366884236, 2968912127,
                                                              335948734, 4161675175, 475798818]


def test_ensemble_workload_streams_are_distinct():
    from ffea_ci_demo_amd import workloads
    specs = workloads.ensemble(6, nmin=2, nmax=4, seed=1)
    states = {tuple(s.rng_state) for s in specs}
    assert len(states) == 6
    ns = [round(s.num_elements / 6 ** 0) for s in specs]
    assert all(n in {6 * k ** 3 for k in (2, 3, 4)} for n in ns)
