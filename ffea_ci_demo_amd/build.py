"""Builds the in-tree native library libffea_gpu.so (HIP for gfx950, hipcc)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libffea_gpu.so")
SOURCES = [os.path.join(HERE, "csrc", "ffea_blob.hip"), os.path.join(HERE, "csrc", "ffea_script.cpp"),
           os.path.join(HERE, "csrc", "ffea_world.cpp"), os.path.join(HERE, "csrc", "amg_setup.cpp")]
HEADERS = [os.path.join(HERE, "csrc", "mrg32k3a.h"), os.path.join(HERE, "csrc", "kernels.h"),
           os.path.join(HERE, "csrc", "amg_kernels.h"), os.path.join(HERE, "csrc", "amg_setup.h"),
           os.path.join(HERE, "csrc", "script_internal.h"), os.path.join(ROOT, "include", "ffea_gpu.h"),
           os.path.join(ROOT, "include", "ffea_script.h"), os.path.join(ROOT, "include", "ffea_world.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc, *HIPCC_FLAGS, "-o", LIB + ".tmp", *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
