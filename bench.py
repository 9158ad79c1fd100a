"""Blob time-step benchmark: element-steps/s on one MI355X.

Workload (default): the 100x100x100 jittered-cube blob — 6.0M quadratic tets,
1.03M corner nodes — with the reference material (shear/bulk modulus 3.6e8 /
2.2e9 Pa, viscosities 1e-3 Pa s), thermal noise on at 300 K, dt = 1e-14 s, CG
tolerance 1e-10, steric/electrostatics off. One step = one full blob update:
element forces, gather, viscosity-matrix assembly, the CG to tolerance, Euler
update. Inputs are resident in HBM before the timed region, which is bracketed by a
synchronisation of the library's HIP stream on both sides (all work runs on it; one
GPU, so there is no barrier across ranks). torch is not used: it bundles a HIP runtime
of its own that must not share the process with the library's.

Prints one JSON line (rank 0). `value` is the headline form's element-steps/s (default:
FP32-mixed with the multigrid preconditioner); `forms` holds every form measured (FP64
and mixed with the reference's Jacobi-PCG and with multigrid, each autotuned before its
timed steps, with its position error against FP64). `roofline` is for the dominant
kernel (the CG SpMV step), from HIP events recorded on the library's stream over the
timed region. `cpu_baseline` times whole steps of the oracle (the FP64 C restatement of
the reference algorithm) on the host's cores, OpenMP build, beside a single-thread
figure.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 100] [--forms fp64,mixed]
  python bench.py --workload ensemble [--blobs 512] [--forms mixed+mg] [--no-graph]
                                              # many blobs in one handle
  python bench.py --workload adjoint          # gradients (config-3 blob), +ref forms
  python bench.py --workload illcond          # the ill-conditioned blob, 2000 steps
  python bench.py --workload masslumped       # the lumped-mass (inertial) solver
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "element-steps/s on 1 MI355X (FP32-mixed vs FP64) + HBM roofline fraction"
PEAK_HBM_GBS = 8000.0
DEFAULT_FORMS = "fp64,mixed,fp64+mg,mixed+mg"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "cg_spmv_cube100_traffic.json")
ENSEMBLE_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "ensemble512_traffic.json")


def profiled_traffic(path, workload, form):
    """HBM bytes per SpMV launch that rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this
    workload and form recorded (scaled by a 1 GiB copy), or None"""
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    if t.get("workload") != workload or form not in t.get("forms", {}):
        return None
    return t["forms"][form]["hbm_bytes_per_launch"]


def spmv_alg_bytes(info, w, projected=False):
    """algorithmic bytes of one CG SpMV-step launch: the viscosity matrix (9 values of
    w bytes + a 4-byte block column per 3x3 block) and five FP64 node 3-vectors (z read
    once, p and q read and written); the mixed form (w = 4) gathers z's packed FP32 copy
    (12 B per row, read once), and its projected form of a free blob also reads the FP64
    positions for its rigid-mode sums"""
    nl = info["num_linear_nodes"]
    return info["num_blocks"] * (9 * w + 4) + (6 if projected else 5) * 24 * nl + (12 * nl if w == 4 else 0)


def projects(specs, prec):
    """the mixed form projects rigid modes out of every SpMV for free blobs (no pins, no
    Stokes drag)"""
    return prec == "mixed" and all(len(s.pinned) == 0 and not s.calc_stokes for s in specs)


def iter_alg_bytes(info, w):
    """the per-CG-iteration model: Nb (9w + 4) + 216 Nl (matrix + 9 FP64 vector passes);
    the mixed form adds z's FP32 copy, written by the vector step and read by the SpMV
    (24 Nl)"""
    nl = info["num_linear_nodes"]
    return info["num_blocks"] * (9 * w + 4) + 216 * nl + (24 * nl if w == 4 else 0)


def step_alg_bytes(info, w, k):
    """the whole-step model B_const + k B_iter: per element its record (32 + 58 w: corner
    indices, rest state and material read, forces and shape data written and read back),
    per corner node 120 B (positions, right-hand side, velocity), the block values written
    (9 w per block), and k CG iterations; a multigrid V-cycle's own bytes are not in it"""
    ne, nl, nb = info["num_elements"], info["num_linear_nodes"], info["num_blocks"]
    return ne * (32 + 58 * w) + 120 * nl + 9 * w * nb + k * iter_alg_bytes(info, w)


def split_form(form):
    """"fp64" / "mixed" (the reference's Jacobi-PCG), "fp64+mg" / "mixed+mg" (multigrid-
    preconditioned CG); a "+ref" suffix adds FP64 refinement of the mixed form's solves:
    (precision, preconditioner)"""
    parts = form.split("+")
    return parts[0], ("multigrid" if "mg" in parts[1:] else "jacobi")


def refined(form):
    return "ref" in form.split("+")[1:]


def run_form(spec, form, steps, warmup, tune=True):
    from ffea_ci_demo_amd.blob import GpuBlob
    prec, pc = split_form(form)
    t0 = time.perf_counter()
    g = GpuBlob.from_spec(spec, precision=prec)
    t_create = time.perf_counter() - t0
    mg = None
    t_mg = 0.0
    if pc == "multigrid":
        t0 = time.perf_counter()
        mg = g.set_preconditioner("multigrid")
        t_mg = time.perf_counter() - t0
        print(f"[bench] multigrid: levels {mg['rows']} (blocks {mg['block']}), coarsest {mg['coarsest_unknowns']} "
              f"unknowns, set-up {mg['setup_seconds']:.1f}s", file=sys.stderr, flush=True)
    info = g.info()
    print(f"[bench] {spec.name} form={form} created in {t_create:.1f}s: Ne={info['num_elements']} "
          f"Nl={info['num_linear_nodes']} blocks={info['num_blocks']} device_bytes={info['device_bytes']/1e9:.2f}GB",
          file=sys.stderr, flush=True)
    for _ in range(warmup):
        g.step(1)
    if tune:
        g.autotune(10)  # row grid, SpMV row unroll, assembly lanes: result-neutral
        g.step(1)
    shape = g.launch_shape()
    _, tot0 = g.cg_iterations()
    g.set_timing(("element", "assemble", "gather_cg_start", "cg_spmv", "cg_vector"))
    g.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        g.step(1)
    g.synchronize()
    wall = time.perf_counter() - t0
    tim = g.timing()
    _, tot1 = g.cg_iterations()
    iters = tot1 - tot0
    w = 8 if prec == "fp64" else 4
    proj = projects([spec], prec)
    spmv_ms, spmv_n = tim["cg_spmv"]
    vec_ms, vec_n = tim["cg_vector"]
    avg_spmv_s = spmv_ms / max(spmv_n, 1) / 1e3
    ne = info["num_elements"]
    per_step = {k: round(tim[k][0] / max(tim[k][1], 1), 4) for k in ("element", "assemble", "gather_cg_start")}
    res = {
        "launch_shape": shape,
        "phase_ms_per_launch": per_step,
        "element_steps_per_s": ne * steps / wall,
        "ms_per_step": wall / steps * 1e3,
        "cg_iterations_per_step": iters / steps,
        "spmv_avg_us": avg_spmv_s * 1e6,
        "vector_avg_us": vec_ms / max(vec_n, 1) * 1e3,
        "spmv_alg_bytes": spmv_alg_bytes(info, w, proj),
        "spmv_achieved_GBs": spmv_alg_bytes(info, w, proj) / avg_spmv_s / 1e9,
        "iteration_model_GBs": iter_alg_bytes(info, w) * iters / wall / 1e9,
        "step_model_bytes": step_alg_bytes(info, w, iters / steps),
        "step_model_GBs": step_alg_bytes(info, w, iters / steps) * steps / wall / 1e9,
        "step_model_frac": step_alg_bytes(info, w, iters / steps) * steps / wall / 1e9 / PEAK_HBM_GBS,
        "setup_s": t_create + t_mg,
        "element_steps_per_s_1000_steps_with_setup": ne * 1000 / (1000 * wall / steps + t_create + t_mg),
        "info": info,
    }
    if mg is not None:
        res["vector_avg_us"] = None
        res["vector_and_vcycle_avg_us"] = vec_ms / max(vec_n, 1) * 1e3
        res["multigrid"] = {k: mg[k] for k in ("rows", "block", "blocks", "transfer_blocks", "coarsest_unknowns",
                                                "setup_seconds")}
    print(f"[bench] form={form} {res['ms_per_step']:.1f} ms/step, k={res['cg_iterations_per_step']:.1f}, "
          f"spmv {res['spmv_avg_us']:.1f} us = {res['spmv_achieved_GBs']:.0f} GB/s", file=sys.stderr, flush=True)
    positions = g.positions()
    g.close()
    return res, positions


def run_ensemble(specs, form, steps, warmup, tune_iters=20, graph=True):
    """all blobs in one handle (+mg: one multigrid hierarchy per blob): autotuned row grid,
    one eager timed step for the SpMV average, then graph-captured (or, graph=False, eager)
    steps timed end to end"""
    from ffea_ci_demo_amd.blob import GpuBatch
    prec, pc = split_form(form)
    t0 = time.perf_counter()
    g = GpuBatch(specs, precision=prec)
    t_create = time.perf_counter() - t0
    info = g.info()
    print(f"[bench] ensemble of {len(specs)} blobs form={form} created in {t_create:.1f}s: Ne={info['num_elements']} "
          f"Nl={info['num_linear_nodes']} chunks={info['num_chunks']} device_bytes={info['device_bytes']/1e9:.2f}GB",
          file=sys.stderr, flush=True)
    mg = None
    if pc == "multigrid":
        t0 = time.perf_counter()
        mg = g.set_preconditioner("multigrid")
        print(f"[bench] ensemble multigrid: {mg['num_levels']} levels, rows {mg['rows']}, coarsest "
              f"{mg['coarsest_unknowns']} unknowns, set-up {time.perf_counter() - t0:.1f}s", file=sys.stderr,
              flush=True)
    g.step(1)
    chosen = g.autotune(tune_iters)
    shape = g.launch_shape()
    print(f"[bench] autotuned launch shape: {shape}", file=sys.stderr, flush=True)
    g.set_timing(("cg_spmv", "cg_vector"))
    g.step(1)
    tim = g.timing()
    g.set_timing(())
    g.use_graph(graph)
    for _ in range(warmup):
        g.step(1)
    last0, tot0 = g.cg_iterations()
    g.synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    g.synchronize()
    wall = time.perf_counter() - t0
    last1, tot1 = g.cg_iterations()
    k = (tot1 - tot0) / steps
    info = g.info()
    w = 8 if prec == "fp64" else 4
    proj = projects(specs, prec)
    spmv_ms, spmv_n = tim["cg_spmv"]
    avg_spmv_s = spmv_ms / max(spmv_n, 1) / 1e3
    ne = info["num_elements"]
    res = {
        "element_steps_per_s": ne * steps / wall,
        "ms_per_step": wall / steps * 1e3,
        "cg_iterations_per_step_max": float(k.max()),
        "cg_iterations_per_step_median": float(np.median(k)),
        "cg_iterations_per_step_min": float(k.min()),
        "spmv_avg_us_all_blobs_running": avg_spmv_s * 1e6,
        "spmv_alg_bytes": spmv_alg_bytes(info, w, proj),
        "spmv_achieved_GBs": spmv_alg_bytes(info, w, proj) / avg_spmv_s / 1e9,
        "row_workgroups": chosen,
        "launch_shape": shape,
        "graph_cg_iterations": info["graph_iterations"],
        "graph": graph,
        "info": info,
    }
    if mg is not None:
        res["multigrid"] = {k: mg[k] for k in ("num_levels", "rows", "coarsest_unknowns", "setup_seconds")}
    print(f"[bench] ensemble form={form} {res['ms_per_step']:.1f} ms/step, k max/median "
          f"{res['cg_iterations_per_step_max']:.0f}/{res['cg_iterations_per_step_median']:.0f}, "
          f"spmv {res['spmv_avg_us_all_blobs_running']:.1f} us = {res['spmv_achieved_GBs']:.0f} GB/s",
          file=sys.stderr, flush=True)
    g.close()
    return res


def cpu_baseline(spec, k_per_step, cg_iters, n=None):
    """The oracle (the FP64 C restatement of the reference's algorithm) on this host's CPU:
    whole steps of the same blob on every core the process may use (OpenMP build with the
    reference's Release flags, oracle/cpu_baseline.py, in a child process), and the
    single-thread figure of one step with the CG capped at cg_iters iterations (a full
    step is element pass + gather/build + k iterations)"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    t0 = time.perf_counter()
    o = oracle.Blob(spec.pos, spec.top, spec.A, spec.B, spec.G, spec.E, **spec.oracle_kwargs())
    t_create = time.perf_counter() - t0
    te, tb, ti = o.time_step(cg_iters)
    del o
    t_step = te + tb + k_per_step * ti
    print(f"[bench] cpu oracle: create {t_create:.1f}s element {te:.2f}s build {tb:.2f}s cg iter {ti*1e3:.1f}ms",
          file=sys.stderr, flush=True)
    single = {"value": spec.num_elements / t_step, "unit": "element-steps/s", "cores": 1,
              "sample": (f"one step on 1 thread (-O2, strict IEEE): element pass {te:.2f}s + gather/matrix build "
                         f"{tb:.2f}s + {cg_iters} CG iterations at {ti*1e3:.1f} ms each, scaled to the "
                         f"{k_per_step:.0f} iterations/step the GPU FP64 run needed")}
    omp = None
    if n is not None:
        try:
            r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"), "--n", str(n),
                                "--steps", "1"], capture_output=True, text=True, timeout=600)
            if r.returncode == 0:
                omp = json.loads(r.stdout.strip().splitlines()[-1])
            else:
                print(f"[bench] OpenMP CPU baseline failed: {r.stderr[-500:]}", file=sys.stderr, flush=True)
        except (subprocess.TimeoutExpired, ValueError) as ex:
            print(f"[bench] OpenMP CPU baseline failed: {ex}", file=sys.stderr, flush=True)
    if omp is None:
        return dict(single, kind="port")
    print(f"[bench] cpu oracle OpenMP: {omp['threads']} threads, {omp['s_per_step']:.2f} s/step "
          f"({omp['cg_iterations']} CG iterations)", file=sys.stderr, flush=True)
    return {
        "value": omp["element_steps_per_s"],
        "unit": "element-steps/s",
        "cores": omp["threads"],
        "kind": "port",
        "sample": (f"one whole step of the same blob ({omp['cg_iterations'][0]} Jacobi-PCG iterations, measured, "
                   f"{omp['s_per_step']:.1f} s) on the oracle's OpenMP build (-O3 -ffast-math -march=native "
                   f"-fopenmp, the reference's Release flags), {omp['threads']} threads of a "
                   f"{omp['host_cpus']}-CPU host"),
        "cpu_model": omp["cpu_model"],
        "single_thread": single,
    }


def cpu_baseline_ensemble(specs, stride=32):
    """The oracle on one host core over a bounded sample of the ensemble: every stride-th
    blob by size (so the sample spans the size range), one whole step each, the
    reference's Jacobi-PCG run to convergence; element-steps/s = sampled tets / time"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    order = sorted(specs, key=lambda s: s.num_elements)
    sample = order[stride // 2::stride]
    ne, secs, iters = 0, 0.0, []
    for spec in sample:
        o = oracle.Blob(spec.pos, spec.top, spec.A, spec.B, spec.G, spec.E, **spec.oracle_kwargs())
        t0 = time.perf_counter()
        if o.step() != 0:
            raise RuntimeError(f"oracle step failed on {spec.name}")
        secs += time.perf_counter() - t0
        ne += spec.num_elements
        iters.append(o.cg_iterations())
        del o
    print(f"[bench] cpu oracle: {len(sample)} blobs, {ne} tets, {secs:.1f} s, Jacobi-PCG iterations "
          f"{min(iters)}..{max(iters)}", file=sys.stderr, flush=True)
    return {"value": ne / secs, "unit": "element-steps/s", "cores": 1, "kind": "port",
            "sample": (f"one whole step (element pass, matrix build, Jacobi-PCG to convergence: "
                       f"{min(iters)}..{max(iters)} iterations, median {int(np.median(iters))}) of every "
                       f"{stride}th blob by size, {len(sample)} blobs, {ne} tets, {secs:.1f} s on 1 thread "
                       f"(-O2, strict IEEE)")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=100, help="cubes per side (100: the 6.0M-tet blob)")
    ap.add_argument("--forms", default=DEFAULT_FORMS,
                    help="fp64 / mixed: the reference's Jacobi-PCG; +mg: multigrid-preconditioned CG")
    ap.add_argument("--headline", default=None, help="form reported as value (default: last of --forms)")
    ap.add_argument("--workload", default="cube",
                    choices=("cube", "ensemble", "adjoint", "illcond", "drift", "masslumped"))
    ap.add_argument("--run-steps", type=int, default=None,
                    help="adjoint: trajectory length K (200); illcond/drift: steps compared (2000/1000)")
    ap.add_argument("--blobs", type=int, default=512, help="ensemble: number of blobs")
    ap.add_argument("--fd-forms", default="fp64",
                    help="adjoint: forms whose gradient is checked by central differences (CG epsilon 1e-13)")
    ap.add_argument("--cpu-cg-iters", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="ensemble: eager steps instead of graph-captured ones")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    from ffea_ci_demo_amd import _native, workloads

    if args.workload == "ensemble":
        return main_ensemble(args)
    if args.workload == "adjoint":
        return main_adjoint(args)
    if args.workload in ("illcond", "drift"):
        return main_drift(args)
    if args.workload == "masslumped":
        return main_masslumped(args)
    t0 = time.perf_counter()
    spec = workloads.cube_blob(args.n, name=f"cube{args.n}_{6 * args.n ** 3}tets")
    print(f"[bench] mesh n={args.n} built in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    forms = [f for f in args.forms.split(",") if f]
    results, pos = {}, {}
    for f in forms:
        results[f], pos[f] = run_form(spec, f, args.steps, args.warmup)
    if "fp64" in pos:
        for f in forms:
            if f != "fp64":
                d = np.abs(pos[f] - pos["fp64"]).max() / np.abs(pos["fp64"]).max()
                results[f]["position_rel_err_vs_fp64"] = float(d)
    head = args.headline or forms[-1]
    h = results[head]
    w = 8 if split_form(head)[0] == "fp64" else 4
    copy_bytes = 1 << 30
    copy_ms = _native.copy_probe(copy_bytes, iters=20)
    read_ms = _native.read_probe(copy_bytes, iters=20)
    traffic = profiled_traffic(TRAFFIC_FILE, spec.name, head)
    roof = {
        "bound": "hbm",
        "kernel": "k_cg_spmv",
        "achieved": round(h["spmv_achieved_GBs"], 1),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(h["spmv_achieved_GBs"] / PEAK_HBM_GBS, 4),
        "traffic": traffic,
        "alg_bytes_per_launch": h["spmv_alg_bytes"],
        "avg_launch_us": round(h["spmv_avg_us"], 2),
        "copy_probe_GBs": round(2 * copy_bytes / (copy_ms / 1e3) / 1e9, 1),
        "read_probe_GBs": round(copy_bytes / (read_ms / 1e3) / 1e9, 1),
    }
    cpu = None
    if not args.no_cpu_baseline:
        k_ref = results["fp64"]["cg_iterations_per_step"] if "fp64" in results else h["cg_iterations_per_step"]
        cpu = cpu_baseline(spec, k_ref, args.cpu_cg_iters, n=args.n)
    ratios = {}
    for a, b in (("fp64", "mixed"), ("fp64+mg", "mixed+mg"), ("fp64", "mixed+mg"), ("fp64", "fp64+mg")):
        if a in results and b in results:
            ratios[f"{b}_vs_{a}"] = round(results[a]["ms_per_step"] / results[b]["ms_per_step"], 3)
    out = {
        "metric": METRIC,
        "value": round(h["element_steps_per_s"], 1),
        "unit": "element-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if split_form(head)[0] == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"{spec.name}: {args.n}^3 jittered unit cubes (edge 1 nm) x 6 quadratic tets, "
                        f"{spec.num_elements} tets, {results[head]['info']['num_linear_nodes']} corner nodes; "
                        "thermal noise at 300 K, dt 1e-14 s, CG epsilon 1e-10, no pins",
            "form": head,
            "cg_iterations_per_step": round(h["cg_iterations_per_step"], 1),
            "precision": ("element algebra, shape data and viscosity matrix FP32; the matrix products gather "
                          "an FP32 copy of their input vector and accumulate in FP64; CG vectors, inner products, "
                          "residual norms and positions FP64" if split_form(head)[0] == "mixed" else "FP64") +
                         ("; multigrid preconditioner matrices FP16 with a power-of-two scale per matrix, FP64 "
                          "accumulation" if split_form(head)[1] == "multigrid" else ""),
            "parallelism": "single GPU",
        },
        "speedup_ratios": ratios,
        "step_model": {"bytes_per_step": h["step_model_bytes"], "achieved_GBs": round(h["step_model_GBs"], 1),
                       "frac": round(h["step_model_frac"], 4),
                       "note": "B_const + k B_iter per step over the step time; a V-cycle's own bytes are not "
                               "counted"},
        "setup": {"setup_s": round(h["setup_s"], 2),
                  "element_steps_per_s_1000_steps_with_setup": round(h["element_steps_per_s_1000_steps_with_setup"],
                                                                     1)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "forms": {f: {k: (float(f"{v:.5g}") if isinstance(v, float) else v) for k, v in r.items() if k != "info"}
                  for f, r in results.items()},
    }
    print(json.dumps(out), flush=True)


def element_pass_bytes(info, w, inertial):
    """algorithmic bytes of one element-pass launch: per element the corner indices (16 B),
    J^-1_0, vol_0, G, E, A, B (14 w), last_det read and written (16 B), forces (12 w) and the
    shape record (16 w) written; per corner node its position (24 B) and, for the inertial
    solver, its velocity (24 B), each read once"""
    ne, nl = info["num_elements"], info["num_linear_nodes"]
    return ne * (32 + 14 * w + 12 * w + 16 * w) + nl * (48 if inertial else 24)


def main_masslumped(args):
    """the inertial solvers on the 6.0M-tet blob: forms fp64 / mixed are MassLumpedSolver's
    explicit steps (element pass with the viscous term V v, gather, a = f/m, v += a dt,
    x += v dt); fp64+cg / mixed+cg are ConjugateGradientSolver's (the consistent mass matrix
    solved by Jacobi-PCG each step, then the same update)"""
    from ffea_ci_demo_amd import _native, units, workloads
    from ffea_ci_demo_amd.blob import GpuBlob
    spec = workloads.cube_blob(args.n, name=f"cube{args.n}_{6 * args.n ** 3}tets")
    rho = 1500.0 * units.VOLUME / units.MASS
    forms = [f for f in args.forms.split(",") if f] if args.forms != DEFAULT_FORMS else ["fp64", "mixed"]
    results, pos = {}, {}
    for form in forms:
        prec, _, kind = form.partition("+")
        g = GpuBlob.from_spec(spec, precision=prec)
        g.set_solver("cg" if kind == "cg" else "masslumped", rho)
        info = g.info()
        for _ in range(args.warmup):
            g.step(1)
        g.set_timing(("element",))
        g.synchronize()
        t0 = time.perf_counter()
        g.step(args.steps)
        g.synchronize()
        wall = time.perf_counter() - t0
        el_ms, el_n = g.timing()["element"]
        w = 8 if prec == "fp64" else 4
        el_s = el_ms / max(el_n, 1) / 1e3
        results[form] = {"element_steps_per_s": spec.num_elements * args.steps / wall,
                         "ms_per_step": wall / args.steps * 1e3, "element_pass_avg_us": el_s * 1e6,
                         "element_pass_alg_bytes": element_pass_bytes(info, w, True),
                         "element_pass_achieved_GBs": element_pass_bytes(info, w, True) / el_s / 1e9}
        if kind == "cg":
            results[form]["cg_iterations_last_step"] = g.cg_iterations()[0]
        pos[form] = g.positions()
        print(f"[bench] masslumped form={form} {results[form]['ms_per_step']:.3f} ms/step, element pass "
              f"{el_s * 1e6:.0f} us = {results[form]['element_pass_achieved_GBs']:.0f} GB/s", file=sys.stderr,
              flush=True)
        g.close()
    if "fp64" in pos:
        for f in forms:
            if f != "fp64":
                results[f]["position_rel_err_vs_fp64"] = float(np.abs(pos[f] - pos["fp64"]).max() /
                                                               np.abs(pos["fp64"]).max())
    head = args.headline or forms[-1]
    h = results[head]
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        o = oracle.Blob(spec.pos, spec.top, spec.A, spec.B, spec.G, spec.E, **spec.oracle_kwargs())
        o.set_mass_lumped(np.full(spec.num_elements, rho))
        t0 = time.perf_counter()
        assert o.step() == 0
        t_step = time.perf_counter() - t0
        del o
        cpu = {"value": spec.num_elements / t_step, "unit": "element-steps/s", "cores": 1, "kind": "port",
               "sample": f"one lumped-mass step of the same blob on the oracle (FP64 C restatement, 1 thread): "
                         f"{t_step:.2f}s"}
    out = _line(round(h["element_steps_per_s"], 1), "element-steps/s", round(h["ms_per_step"], 4), args.steps,
                args.warmup, "f64" if head.startswith("fp64") else "f32+f64",
                {"workload": f"{spec.name}: {args.n}^3 jittered unit cubes x 6 quadratic tets, {spec.num_elements} "
                             "tets, density 1500 kg/m^3, thermal noise at 300 K, dt 1e-14 s; lumped-mass solver "
                             "(explicit inertial step, no linear solve)", "form": head,
                 "parallelism": "single GPU"},
                {f: {k: (float(f"{v:.5g}") if isinstance(v, float) else v) for k, v in r.items()}
                 for f, r in results.items()})
    out["roofline"] = {"bound": "hbm", "kernel": "k_element", "achieved": round(h["element_pass_achieved_GBs"], 1),
                       "peak": PEAK_HBM_GBS, "unit": "GB/s",
                       "frac": round(h["element_pass_achieved_GBs"] / PEAK_HBM_GBS, 4), "traffic": None,
                       "alg_bytes_per_launch": h["element_pass_alg_bytes"],
                       "avg_launch_us": round(h["element_pass_avg_us"], 2)}
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


def main_ensemble(args):
    from ffea_ci_demo_amd import _native, workloads
    t0 = time.perf_counter()
    specs = workloads.config2(args.blobs)
    ne_tot = sum(s.num_elements for s in specs)
    ns = [round((s.num_elements / 6) ** (1 / 3)) for s in specs]
    print(f"[bench] {len(specs)} meshes ({ne_tot} tets, n {min(ns)}..{max(ns)}) built in "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    form = (args.headline or "mixed+mg") if args.forms == DEFAULT_FORMS else args.forms.split(",")[-1]
    h = run_ensemble(specs, form, args.steps, args.warmup, graph=not args.no_graph)
    copy_bytes = 1 << 30
    copy_ms = _native.copy_probe(copy_bytes, iters=20)
    read_ms = _native.read_probe(copy_bytes, iters=20)
    roof = {
        "bound": "hbm",
        "kernel": "k_cg_spmv",
        "achieved": round(h["spmv_achieved_GBs"], 1),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(h["spmv_achieved_GBs"] / PEAK_HBM_GBS, 4),
        "traffic": profiled_traffic(ENSEMBLE_TRAFFIC_FILE, f"ensemble{len(specs)}", form),
        "alg_bytes_per_launch": h["spmv_alg_bytes"],
        "avg_launch_us": round(h["spmv_avg_us_all_blobs_running"], 2),
        "copy_probe_GBs": round(2 * copy_bytes / (copy_ms / 1e3) / 1e9, 1),
        "read_probe_GBs": round(copy_bytes / (read_ms / 1e3) / 1e9, 1),
    }
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_ensemble(specs)
    out = {
        "metric": METRIC,
        "value": round(h["element_steps_per_s"], 1),
        "unit": "element-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if split_form(form)[0] == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"ensemble of {len(specs)} independent jittered-cube blobs, n^3 cubes with n = "
                        f"round(exp(U(ln 8, ln 40))) ({min(ns)}..{max(ns)}), {ne_tot} tets in all; shear and bulk "
                        "moduli each log-uniform in [1e7, 1e9] Pa; one MRG32k3a stream per blob; noise at 300 K, "
                        "dt 1e-14 s, CG epsilon 1e-10; one handle, " +
                        ("graph-captured steps" if not args.no_graph else "eager steps"),
            "form": form,
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "forms": {form: {k: (float(f"{v:.5g}") if isinstance(v, float) else v) for k, v in h.items() if k != "info"}},
    }
    print(json.dumps(out), flush=True)


def _line(value, unit, ms, steps, warmup, dtype, config, forms, higher=True, roof=None, cpu=None):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic", "config": config, "roofline": roof, "cpu_baseline": cpu, "forms": forms}


def main_adjoint(args):
    """Differentiation blob: 50^3 jittered cubes (750k tets), per-element log-normal moduli,
    bottom face pinned, +1 pN z on the top-face corner nodes, noise off. Target field: the
    positions after K steps with a second moduli seed. Loss: sum over corner nodes of
    |pos_K - target|^2. Gradient in G, E, A, B per element by the adjoint with checkpoints
    every 20 steps in every form of --forms (+ref: FP64 refinement of the mixed form's
    solves to 1e-10), each against the FP64 form's. Directional central differences of the
    loss check the FP64 gradient, one random direction in each of G, E, A and B, with the
    CG tolerance at 1e-13 for both the differences and the adjoint they check."""
    from ffea_ci_demo_amd import workloads
    from ffea_ci_demo_amd.blob import GpuBlob
    n = args.n if args.n != 100 else 50
    K = args.run_steps or 200
    interval = 20
    spec = workloads.config3(n=n)
    tgt_spec = workloads.config3(n=n, moduli_seed=12)
    corner = spec.mesh.is_linear
    with GpuBlob.from_spec(tgt_spec) as g:
        g.set_preconditioner("multigrid")
        g.step(K)
        target = g.positions()

    def loss(pos):
        d = (pos - target) * corner[:, None]
        return float((d ** 2).sum()), 2.0 * d

    def make(form, s=spec):
        prec, pc = split_form(form)
        g = GpuBlob.from_spec(s, precision=prec)
        if pc == "multigrid":
            g.set_preconditioner("multigrid")
        if refined(form):
            g.set_refinement(1e-10)
        return g

    forms, grads = {}, {}
    form_list = [f for f in args.forms.split(",") if f]
    if "fp64" not in form_list and "fp64+mg" not in form_list:
        form_list.insert(0, "fp64+mg")
    for form in form_list:
        with make(form) as g:
            g.step(2)  # warm-up (kernels loaded, iteration count learned)
            g.set_positions(spec.pos)
            g.synchronize()
            t0 = time.perf_counter()
            g.run_checkpointed(K, interval)
            t_fwd = time.perf_counter() - t0
            it_last = g.cg_iterations()[0]
            L, seed = loss(g.positions())
            t0 = time.perf_counter()
            dA, dB, dG, dE, _ = g.adjoint(seed)
            t_bwd = time.perf_counter() - t0
            grads[form] = {"A": dA, "B": dB, "G": dG, "E": dE}
            forms[form] = {"loss": L, "forward_ms_per_step": t_fwd / K * 1e3, "reverse_ms_per_step": t_bwd / K * 1e3,
                           "gradient_element_steps_per_s": spec.num_elements * K / (t_fwd + t_bwd),
                           "cg_iterations_last_forward_step": it_last}
            if refined(form):
                forms[form]["refinement_correction_solves"] = g.refinement_sweeps()[1]
            print(f"[bench] adjoint {form}: forward {t_fwd:.2f}s reverse {t_bwd:.2f}s loss {L:.6e}",
                  file=sys.stderr, flush=True)
    ref_form = "fp64" if "fp64" in grads else "fp64+mg"
    for f in form_list:
        if f == ref_form:
            continue
        forms[f]["loss_rel_err_vs_fp64"] = abs(forms[f]["loss"] - forms[ref_form]["loss"]) / abs(forms[ref_form]["loss"])
        for name in "ABGE":
            a, b = grads[f][name], grads[ref_form][name]
            forms[f][f"grad_{name}_rel_err_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    # finite-difference checks of the FP64 forms' adjoints at CG tolerance 1e-13
    fd = {}
    for form in [f for f in args.fd_forms.split(",") if f]:
        s13 = workloads.config3(n=n, epsilon=1e-13)
        with make(form, s13) as g:
            g.run_checkpointed(K, interval)
            _, seed = loss(g.positions())
            ga = dict(zip("ABGE", g.adjoint(seed)[:4]))
            r = np.random.default_rng(7)
            h = 1e-4
            errs = {}
            for name in "GEAB":
                d = r.uniform(-1.0, 1.0, spec.num_elements)
                base = {"A": s13.A, "B": s13.B, "G": s13.G, "E": s13.E}
                vals = []
                for sgn in (1.0, -1.0):
                    m = dict(base)
                    m[name] = base[name] * (1 + sgn * h * d)
                    g.set_material(m["A"], m["B"], m["G"], m["E"])
                    g.set_positions(s13.pos)
                    g.step(K)
                    vals.append(loss(g.positions())[0])
                fdv = (vals[0] - vals[1]) / (2 * h)
                ad = float((ga[name] * base[name] * d).sum())
                errs[name] = abs(fdv - ad) / abs(fdv)
                print(f"[bench] fd check {form} d{name}: fd {fdv:.9e} adjoint {ad:.9e} rel err {errs[name]:.2e}",
                      file=sys.stderr, flush=True)
            fd[form] = {"epsilon": 1e-13, "h": h, "rel_err": errs, "max_rel_err": max(errs.values())}
    head = args.headline or ("fp64+mg" if "fp64+mg" in forms else form_list[0])
    h = forms[head]
    out = _line(round(h["gradient_element_steps_per_s"], 1), "element-steps/s", round(h["forward_ms_per_step"] +
                h["reverse_ms_per_step"], 3), K, 2, "f64" if split_form(head)[0] == "fp64" else "f32+f64",
                {"workload": f"differentiation blob: {n}^3 jittered cubes, {spec.num_elements} tets, log-normal G/E "
                             "(median 3.6e8/2.2e9 Pa, sigma 0.5), bottom face pinned, +1 pN z on top-face corner "
                             f"nodes, noise off; loss vs the positions of a second moduli seed after K={K} steps; "
                             f"adjoint with checkpoints every {interval} steps; value = (forward + reverse) "
                             f"element-steps/s of the {head} gradient", "form": head, "parallelism": "single GPU"},
                {f: {k: (float(f"{v:.5g}") if isinstance(v, float) else v) for k, v in r.items()}
                 for f, r in forms.items()})
    out["fd_check"] = fd
    print(json.dumps(out), flush=True)


def main_drift(args):
    """the FP64 form (the reference's algorithm) and the other forms of --forms stepped
    side by side on the same blob; each form's position error against FP64 and the CG
    iteration counts along the run"""
    from ffea_ci_demo_amd import workloads
    from ffea_ci_demo_amd.blob import FFEAException, GpuBlob
    if args.workload == "illcond":
        spec = workloads.config4(n=args.n if args.n != 100 else 40)
        K = args.run_steps or 2000
    else:
        spec = workloads.cube_blob(args.n, name=f"cube{args.n}_{6 * args.n ** 3}tets")
        K = args.run_steps or 1000
    others = [f for f in args.forms.split(",") if f and f != "fp64"] or ["mixed"]
    marks = sorted({m for m in [1, 10, K] + list(range(50, K + 1, 50)) if m <= K})
    handles = {"fp64": GpuBlob.from_spec(spec, precision="fp64")}
    for f in others:
        prec, pc = split_form(f)
        handles[f] = GpuBlob.from_spec(spec, precision=prec)
        if pc == "multigrid":
            handles[f].set_preconditioner("multigrid")
    for h in handles.values():
        h.use_graph(True)
    rows = {f: [] for f in others}
    wall = {f: 0.0 for f in handles}
    done, aborted = 0, None
    for mark in marks:
        n = mark - done
        try:
            for f, h in handles.items():
                t0 = time.perf_counter()
                h.step(n)
                h.synchronize()
                wall[f] += time.perf_counter() - t0
        except FFEAException as ex:  # the reference aborts the same way (Blob.cpp:484-490, NoMassCGSolver.cpp:187)
            aborted = {"after_step": done, "within_next": n, "form": f, "message": str(ex).strip()}
            print(f"[bench] {f} aborted after step {done} (+<= {n}): {str(ex).strip()}", file=sys.stderr, flush=True)
            break
        done = mark
        pd = handles["fp64"].positions()
        kd = handles["fp64"].cg_iterations()[0]
        msg = []
        for f in others:
            err = float(np.abs(handles[f].positions() - pd).max() / np.abs(pd).max())
            k = handles[f].cg_iterations()[0]
            rows[f].append({"step": mark, "pos_rel_err_vs_fp64": err, "k_fp64": kd, "k": k})
            msg.append(f"{f} rel err {err:.3e} k {k}")
        print(f"[bench] {spec.name} step {mark}: k_fp64 {kd}; " + "; ".join(msg), file=sys.stderr, flush=True)
    for h in handles.values():
        h.close()
    K = done
    forms = {f: {"element_steps_per_s": spec.num_elements * K / wall[f], "ms_per_step": wall[f] / K * 1e3}
             for f in handles}
    for f in others:
        forms[f]["drift"] = rows[f]
    if aborted:
        forms["aborted"] = aborted
    head = args.headline or others[-1]
    out = _line(round(forms[head]["element_steps_per_s"], 1), "element-steps/s",
                round(forms[head]["ms_per_step"], 3), K, 0, "f32+f64" if split_form(head)[0] == "mixed" else "f64",
                {"workload": f"{spec.name}: {spec.num_elements} tets, {K} steps of FP64 and {', '.join(others)} side by "
                             "side (graph-captured steps); positions relative error max|form-fp64|/max|fp64|",
                 "form": head, "parallelism": "single GPU"},
                {f: {k: (float(f"{v:.5g}") if isinstance(v, float) else v) for k, v in r.items()}
                 for f, r in forms.items()})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
