"""Rotational constant forces (Blob::apply_ctforces CTF 2, Blob.cpp:1684-1719) on the HIP
path against the oracle: per-step right-hand side within 1e-12 relative given identical
positions and draws, FP64 positions within 1e-10 relative after N steps. The forces act on
every node (corner and midside) about one axis, plus repeated and 't' entries on single
nodes, on top of linear node forces; midside shares go through the per-step
linearise_force records. A blob loaded from a .ffea script with a rotational .ctforces
section matches the oracle built from the same arrays."""
import numpy as np
import pytest

from conftest import rel_err
from ffea_ci_demo_amd import meshgen, workloads
from ffea_ci_demo_amd.blob import FFEAException, GpuBlob
from ffea_ci_demo_amd.script import Script
from test_layout import stokes_spec

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12
POS_TOL = 1e-10


def loaded(spec, F=3e-3, torque=False):
    """rotational forces on every node about an axis through a point off the mesh nodes,
    extra entries on nodes 5 (twice) and the last node about a second axis, and linear
    forces on nodes 5 and 9"""
    nt = spec.pos.shape[0]
    c = spec.pos.mean(axis=0) + np.array([0.0123, -0.0311, 0.0])
    d1 = np.array([0.2, -0.1, 1.0]) / np.linalg.norm([0.2, -0.1, 1.0])
    d2 = np.array([1.0, 0.0, 0.0])
    nodes = np.concatenate([np.arange(nt), [5, 5, nt - 1]]).astype(np.int32)
    axis = np.tile(np.concatenate([c, d1]), (len(nodes), 1))
    axis[nt + 1:] = np.concatenate([c + [0.0, 0.37, 0.11], d2])
    Fs = np.full(len(nodes), F)
    Fs[nt:] = [-0.5 * F, 2.0 * F, 1.5 * F]
    types = ["pt" if torque else "pf"] * nt + ["pt", "pf", "pt"]
    spec.rotational_forces = (nodes, Fs, axis, types)
    spec.ctf_node = np.array([5, 9], np.int32)
    spec.ctf_force = np.array([[1e-3, -2e-3, 0.5e-3], [0.0, 0.0, 1e-3]])
    return spec


def pair(oracle_mod, spec, precision="fp64"):
    print(f"[case] rotational forces {spec.name} ne={spec.num_elements} entries={len(spec.rotational_forces[0])} "
          f"{precision}", flush=True)
    o = oracle_mod.Blob(spec.pos, spec.top, spec.A, spec.B, spec.G, spec.E, **spec.oracle_kwargs())
    return o, GpuBlob.from_spec(spec, precision=precision)


def lockstep(o, g, nsteps):
    for k in range(nsteps):
        g.set_positions(o.positions())
        g.set_rng_state(o.rng_state())
        assert o.step() == 0
        g.step(1)
        e = rel_err(g.rhs(), o.rhs())
        assert e <= FORCE_TOL, (k, e)


def pinned_spec(n=5, noise=1):
    spec = workloads.config3(n=n, epsilon=1e-12)
    spec.calc_noise = noise
    return spec


@pytest.mark.parametrize("torque", [False, True])
def test_pinned_blob_lockstep_and_positions(oracle_mod, torque):
    spec = loaded(pinned_spec(), torque=torque)
    o, g = pair(oracle_mod, spec)
    lockstep(o, g, 5)
    for _ in range(20):
        assert o.step() == 0
    g.step(20)
    assert rel_err(g.positions(), o.positions()) <= POS_TOL
    g.close()


def test_free_blob_turns(oracle_mod):
    """no pins, Stokes drag, noise off: the blob turns about the axis the way the forces push"""
    spec = loaded(stokes_spec(5, seed=3, noise=0), F=1e-2)
    spec.pinned = np.zeros(0, np.int32)
    o, g = pair(oracle_mod, spec)
    x0 = g.positions()
    for _ in range(10):
        assert o.step() == 0
    g.step(10)
    x1 = g.positions()
    assert rel_err(x1, o.positions()) <= POS_TOL
    c, d = spec.rotational_forces[2][0, :3], spec.rotational_forces[2][0, 3:]
    turn = np.cross(x0 - c, x1 - x0) @ d  # per node: positive when it moves the 'f' way round d
    assert np.median(turn) > 0
    g.close()


def test_stokes_noise_surface_and_rotational_forces(oracle_mod):
    spec = stokes_spec(5, seed=6)
    z = spec.pos[:, 2]
    spec.pinned = np.nonzero(z <= z.min() + 1e-9)[0].astype(np.int32)
    spec = loaded(spec)
    zmax = z.max()
    spec.surface_forces = workloads.surface_force_sets(spec, [(np.array([3e-12, -1e-12, 2e-12]),
                                                               lambda cc: cc[:, 2] > zmax - 1e-6)])
    o, g = pair(oracle_mod, spec)
    lockstep(o, g, 5)
    for _ in range(10):
        assert o.step() == 0
    g.step(10)
    assert rel_err(g.positions(), o.positions()) <= POS_TOL
    g.close()


def test_mixed_form_close(oracle_mod):
    spec = loaded(pinned_spec(noise=0))
    o, g = pair(oracle_mod, spec, "mixed")
    for _ in range(10):
        assert o.step() == 0
    g.step(10)
    assert rel_err(g.positions(), o.positions()) <= 1e-8
    g.close()


def test_with_multigrid_preconditioner_and_graph(oracle_mod):
    spec = loaded(pinned_spec(n=11))
    o, g = pair(oracle_mod, spec)
    g.set_preconditioner("multigrid")
    g.use_graph(True)
    for _ in range(6):
        assert o.step() == 0
    g.step(6)
    assert rel_err(g.rhs(), o.rhs()) <= 1e-10
    assert rel_err(g.positions(), o.positions()) <= POS_TOL
    g.close()


def test_script_blob_matches_oracle(oracle_mod, tmp_path):
    mesh = meshgen.kuhn_mesh(3, seed=2, jitter=0.1)
    prefix = str(tmp_path / "rb")
    z = mesh.pos[:, 2]
    meshgen.write_blob_files(mesh, prefix, pinned=np.nonzero(z <= z.min() + 1e-9)[0])
    meshgen.write_full_ctforces(str(tmp_path / "rb.ctforces"),
                                node_forces=[(2e-12, (0, 0, 1.0), 7)],
                                rot_forces=[(5e-12, "pf", (1.3e-9, 1.1e-9, 0.0), (0, 0.3, 1), "all"),
                                            (1e-12, "pt", (0.0, 0.0, 0.0), (1, 0, 0), 11)])
    script = str(tmp_path / "rb.ffea")
    meshgen.write_ffea_script(script, prefix, calc_ctforces=1, ctforces_fname="rb.ctforces")
    S = Script(script)
    spec = S.blob_spec(0)
    print(f"[case] rotational forces from script ne={spec.num_elements}", flush=True)
    o = oracle_mod.Blob(spec.pos, spec.top, spec.A, spec.B, spec.G, spec.E, **spec.oracle_kwargs())
    g = S.gpu_blob(0)
    for _ in range(8):
        assert o.step() == 0
    g.step(8)
    assert rel_err(g.rhs(), o.rhs()) <= 1e-10
    assert rel_err(g.positions(), o.positions()) <= POS_TOL
    g.close()


def test_refusals():
    spec = loaded(pinned_spec(n=3, noise=0))
    nodes, F, axis, types = spec.rotational_forces
    for bad in [(nodes, F, axis, ["nf"] + types[1:]), (nodes, F, axis, ["pq"] + types[1:]),
                (np.where(nodes == 0, spec.pos.shape[0], nodes), F, axis, types)]:
        s2 = loaded(pinned_spec(n=3, noise=0))
        s2.rotational_forces = bad
        with pytest.raises(FFEAException):
            GpuBlob.from_spec(s2)
    with GpuBlob.from_spec(spec) as g:  # position-dependent forces: no adjoint
        with pytest.raises(FFEAException, match="do not depend on positions"):
            g.run_checkpointed(4, 2)
