"""The ill-conditioned blob (shear 1e7 Pa, bulk 1e10 Pa, 2% of the cubes squashed to
height 1/20, noise on) on the GPU against the oracle.

Tolerances: right-hand side in lockstep within 1e-12 (given identical positions and
draws), FP64 positions within 1e-10 after 60 steps, the FP32-mixed form's positions
within 1e-5 of the oracle's; relative = max|a-b| / max|b|.
"""
import numpy as np
import pytest

from conftest import rel_err
from ffea_ci_demo_amd import workloads
from ffea_ci_demo_amd.blob import GpuBlob

pytestmark = pytest.mark.gpu


def _pair(oracle_mod, spec, precision):
    print(f"[case] {spec.name} ne={spec.num_elements} precision={precision}", flush=True)
    o = oracle_mod.Blob(spec.pos, spec.top, spec.A, spec.B, spec.G, spec.E, **spec.oracle_kwargs())
    return o, GpuBlob.from_spec(spec, precision=precision)


def test_illcond_rhs_lockstep(oracle_mod):
    spec = workloads.config4(n=8, rng_seed=21)
    o, g = _pair(oracle_mod, spec, "fp64")
    for k in range(12):
        g.set_positions(o.positions())
        g.set_rng_state(o.rng_state())
        assert o.step() == 0
        g.step(1)
        e = rel_err(g.rhs(), o.rhs())
        assert e <= 1e-12, (k, e)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-5)])
def test_illcond_positions(oracle_mod, precision, tol):
    spec = workloads.config4(n=8, rng_seed=22)
    o, g = _pair(oracle_mod, spec, precision)
    nsteps = 60
    for _ in range(nsteps):
        assert o.step() == 0
    g.step(nsteps)
    e = rel_err(g.positions(), o.positions())
    print(f"[illcond] {precision}: positions rel err {e:.3e} after {nsteps} steps", flush=True)
    assert e <= tol
    assert abs(g.cg_iterations()[0] - o.cg_iterations()) <= (1 if precision == "fp64" else 3)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-5)])
def test_illcond_multigrid(oracle_mod, precision, tol):
    """a blob large enough for the multigrid preconditioner (1331 corner nodes)"""
    spec = workloads.config4(n=10, rng_seed=23)
    o, g = _pair(oracle_mod, spec, precision)
    g.set_preconditioner("multigrid")
    g.use_graph(True)
    nsteps = 50
    for _ in range(nsteps):
        assert o.step() == 0
    g.step(nsteps)
    e = rel_err(g.positions(), o.positions())
    print(f"[illcond] {precision}+multigrid: positions rel err {e:.3e}, iterations {g.cg_iterations()[0]} "
          f"(oracle {o.cg_iterations()})", flush=True)
    assert e <= tol
    assert g.cg_iterations()[0] * 3 <= o.cg_iterations()
