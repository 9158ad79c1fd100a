"""ORACLE — test infrastructure only: the CPU baseline that bench.py reports.

Times whole steps of the oracle's FP64 restatement of the reference's blob step on this
host's cores: the OpenMP build (libffea_oracle_omp.so, compiled here with the reference's
Release flags -O3 -ffast-math -march=native plus -fopenmp; its element loop, matrix build,
SpMV, dot products and vector updates run in parallel as the reference's OpenMP loops do,
Blob.cpp:419-481 and NoMassCGSolver.cpp:199-281). Prints one JSON object.

  python oracle/cpu_baseline.py --n 100 --steps 1 [--lib PATH]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def build_omp(out):
    subprocess.run(["make", "-s", "-C", HERE, "omp", f"OMP_OUT={out}"], check=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=100)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--lib", default=None, help="an OpenMP build to load (default: build one here)")
    args = ap.parse_args()
    lib = args.lib
    tmp = None
    if lib is None:
        tmp = tempfile.mkdtemp(prefix="ffea_oracle_omp_")
        lib = os.path.join(tmp, "libffea_oracle_omp.so")
        build_omp(lib)
    os.environ["FFEA_ORACLE_LIB"] = lib
    sys.path.insert(0, HERE)
    sys.path.insert(0, os.path.dirname(HERE))
    import oracle
    from ffea_ci_demo_amd import workloads
    spec = workloads.cube_blob(args.n, name=f"cube{args.n}")
    t0 = time.perf_counter()
    o = oracle.Blob(spec.pos, spec.top, spec.A, spec.B, spec.G, spec.E, **spec.oracle_kwargs())
    t_create = time.perf_counter() - t0
    iters = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rc = o.step()
        if rc != 0:
            raise SystemExit(f"oracle step failed ({rc})")
        iters.append(o.cg_iterations())
    wall = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    print(json.dumps({"element_steps_per_s": spec.num_elements * args.steps / wall, "s_per_step": wall / args.steps,
                      "cg_iterations": iters, "threads": threads, "host_cpus": os.cpu_count(),
                      "cpu_model": cpu_model(), "create_s": t_create, "num_elements": spec.num_elements}))


if __name__ == "__main__":
    main()
