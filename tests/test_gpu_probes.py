"""The achievable-bandwidth probes (bench.py reports them beside the roofline): each
returns a positive mean launch time whose rate is below the HBM peak of 8 TB/s."""
import pytest

from ffea_ci_demo_amd import _native

pytestmark = pytest.mark.gpu

PEAK_GBS = 8000.0


def test_copy_and_read_probes_report_physical_rates():
    nbytes = 1 << 28
    copy_ms = _native.copy_probe(nbytes, iters=5)
    read_ms = _native.read_probe(nbytes, iters=5)
    copy_gbs = 2 * nbytes / (copy_ms / 1e3) / 1e9
    read_gbs = nbytes / (read_ms / 1e3) / 1e9
    print(f"copy {copy_gbs:.0f} GB/s (in + out), read {read_gbs:.0f} GB/s", flush=True)
    assert 0.0 < copy_gbs < PEAK_GBS
    assert 0.0 < read_gbs < PEAK_GBS


def test_probes_refuse_empty_sizes():
    with pytest.raises(RuntimeError):
        _native.copy_probe(8)
    with pytest.raises(RuntimeError):
        _native.read_probe(0)
