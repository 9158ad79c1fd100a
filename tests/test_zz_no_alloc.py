"""The allocation-free strain-energy entry point (ffea_gpu_batch_strain_energy_noalloc)
against the allocating one (ffea_gpu_batch_strain_energy), through the C ABI: equal bit for
bit on two calls in a row, the second on a smaller handle, and a handle too large for its
scratch refused."""
import ctypes as C

import numpy as np
import pytest

from ffea_ci_demo_amd import _native, workloads
from ffea_ci_demo_amd.blob import GpuBatch, GpuBlob

pytestmark = pytest.mark.gpu


def _energies(handle, nblob):
    L = _native.lib()
    a = np.zeros(nblob)
    b = np.full(nblob, np.nan)
    assert L.ffea_gpu_batch_strain_energy(handle.h, a) == _native.OK
    rc = L.ffea_gpu_batch_strain_energy_noalloc(handle.h, b)
    assert rc == _native.OK, L.ffea_gpu_blob_last_error(handle.h)
    return a, b


def test_noalloc_strain_energy_equals_allocating_one():
    specs = workloads.ensemble(3, nmin=6, nmax=10, seed=3)
    print(f"[case] noalloc strain energy: 3-blob ensemble, {sum(s.num_elements for s in specs)} tets", flush=True)
    with GpuBatch(specs, precision="fp64") as g:
        g.step(3)
        a, b = _energies(g, 3)
        assert np.all(a > 0) and np.array_equal(a, b), (a, b)
    spec = workloads.cube_blob(4, rng_seed=3)
    print(f"[case] noalloc strain energy: then one {spec.num_elements}-tet blob, mixed", flush=True)
    with GpuBlob.from_spec(spec, precision="mixed") as g:
        g.step(3)
        a, b = _energies(g, 1)
        assert a[0] > 0 and np.array_equal(a, b), (a, b)
        g.step(2)
        a, b = _energies(g, 1)
        assert np.array_equal(a, b)


def test_noalloc_refuses_beyond_its_scratch():
    spec = workloads.cube_blob(112, calc_noise=0)  # 8,429,568 tets: more than 8,388,608 doubles
    print(f"[case] noalloc refusal: {spec.num_elements}-tet blob", flush=True)
    with GpuBlob.from_spec(spec, precision="mixed") as g:
        out = np.zeros(1)
        rc = _native.lib().ffea_gpu_batch_strain_energy_noalloc(g.h, out)
        assert rc == _native.ERR_INVALID
        assert b"more than its" in _native.lib().ffea_gpu_blob_last_error(g.h)
