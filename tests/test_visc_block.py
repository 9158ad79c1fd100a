"""The assembly forms each 3x3 block with visc_block, which picks the two corners' shape
gradients once; the element pass and the reference's operation order are visc_entry, one
entry at a time in create_viscosity_matrix's order (tetra_element_linear.cpp:298-327).
Every block entry must be equal bit for bit, for all 16 corner pairs, in FP32 and FP64:
built on the host from the same header with hipcc and run here (no GPU needed)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_visc_block_equals_visc_entry(tmp_path):
    exe = str(tmp_path / "visc_block_check")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-I",
                    os.path.join(ROOT, "ffea_ci_demo_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "visc_block_check.cpp"), "-o", exe],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "float 0 double 0" in r.stdout
