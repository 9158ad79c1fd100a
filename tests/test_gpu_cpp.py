"""The C++ drop-in (include/ffea_gpu.hpp) driving a script on the GPU: the example
program runs the reference's World loop for a script, writes its output files and
reports the same positions as the Python binding."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from ffea_ci_demo_amd import meshgen
from ffea_ci_demo_amd.script import Script

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu


def test_cpp_script_runner(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "run_script")
    subprocess.run([gxx, "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "ffea_ci_demo_amd", "examples", "run_script.cpp"),
                    "-L", os.path.join(ROOT, "ffea_ci_demo_amd"), "-lffea_gpu",
                    "-Wl,-rpath," + os.path.join(ROOT, "ffea_ci_demo_amd"), "-o", exe], check=True)
    mesh = meshgen.kuhn_mesh(3, seed=2)
    prefix = str(tmp_path / "b")
    meshgen.write_blob_files(mesh, prefix)
    path = str(tmp_path / "b.ffea")
    meshgen.write_ffea_script(path, prefix, num_steps=9, check=5, rng_seed=5, epsilon=1e-10, max_iterations_cg=20000)
    print("[case] C++ runner on a 3^3 cube script, 10 updates", flush=True)
    r = subprocess.run([exe, path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    last = r.stdout.strip().splitlines()[-1]
    assert last.startswith("done: 10 updates")
    for name in ("b_trajectory.out", "b_measurement.out", "b_measurement.fdm", "b.fcp"):
        assert os.path.getsize(tmp_path / name) > 0, name
    # -d: no detailed measurement file, as ffea -d
    meshgen.write_blob_files(mesh, str(tmp_path / "c"))
    path_c = str(tmp_path / "c.ffea")
    meshgen.write_ffea_script(path_c, str(tmp_path / "c"), num_steps=1, check=1, rng_seed=5)
    rc = subprocess.run([exe, "-d", path_c], capture_output=True, text=True, timeout=120)
    assert rc.returncode == 0, rc.stderr
    assert os.path.exists(tmp_path / "c_measurement.out") and not os.path.exists(tmp_path / "c_measurement.fdm")
    g = Script(path).gpu_blob(0)
    g.step(10)
    x0 = g.positions()[0]
    got = [float(v) for v in last.rsplit("(", 1)[1].rstrip(")").split(",")]
    assert np.allclose(got, x0, rtol=1e-9, atol=0)


@pytest.mark.parametrize("calc_stokes", [0, 1])
def test_cpp_solver_drop_in(tmp_path, calc_stokes):
    """ffea_gpu::NoMassSolver (the Solver-level drop-in) solves a GPU blob's right-hand side
    at the positions of its element pass to the velocities the blob's own step found"""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "solver_drop_in")
    subprocess.run([gxx, "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "ffea_ci_demo_amd", "examples", "solver_drop_in.cpp"),
                    "-L", os.path.join(ROOT, "ffea_ci_demo_amd"), "-lffea_gpu",
                    "-Wl,-rpath," + os.path.join(ROOT, "ffea_ci_demo_amd"), "-o", exe], check=True)
    mesh = meshgen.kuhn_mesh(4, seed=3)
    prefix = str(tmp_path / "b")
    meshgen.write_blob_files(mesh, prefix, stokes_radius=0.4 if calc_stokes else 0.0)
    path = str(tmp_path / "b.ffea")
    meshgen.write_ffea_script(path, prefix, num_steps=5, check=5, rng_seed=9, epsilon=1e-10,
                              max_iterations_cg=20000, calc_stokes=calc_stokes)
    print(f"[case] C++ solver drop-in, 4^3 cube, calc_stokes={calc_stokes}", flush=True)
    r = subprocess.run([exe, path, "4"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    last = r.stdout.strip().splitlines()[-1]
    assert last.startswith("done: 4 steps")
    assert float(last.rsplit(" ", 1)[1]) <= 1e-13, r.stdout
