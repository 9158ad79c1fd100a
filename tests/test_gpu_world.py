"""The World loop (include/ffea_world.h): trajectory, measurement and checkpoint files
against the oracle stepped alongside, and the reference's own restart test
(tests/physics/sphere_63_120_nomass_restart: a 10-step run against a 5-step run restarted
to 10; checkpoints identical, measurements within 5e-4 relative, checkRuns.py:36-80) on
the reference's sphere_63_120 mesh.

Output tolerances follow the files' own precision: trajectory positions are printed with
%e (6 significant digits), measurements with %-14.6e.
"""
import os
import shutil

import numpy as np
import pytest

from conftest import rel_err
from ffea_ci_demo_amd import _native, units
from ffea_ci_demo_amd.blob import FFEAException, GpuBlob
from ffea_ci_demo_amd.script import Script
from ffea_ci_demo_amd.world import World

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "sphere_63_120")
PRINT_RTOL = 1e-6  # %e / %.6e keep 7 significant digits: half an ulp of the 7th is 5e-7 relative

# the <param> block of sphere_63_120_nomass_*.ffea (reference tests/physics/sphere_63_120_nomass_restart),
# plus calc_steric = 0: those scripts leave it at its default 1, the steric surface solver,
# which is not part of this build
BASE = [("calc_steric", "0"), ("restart", "0"), ("dt", "1e-14"), ("kT", "4.11e-21"), ("check", "1"), ("num_steps", "5"),
        ("rng_seed", "44"), ("trajectory_out_fname", None), ("measurement_out_fname", None),
        ("epsilon", "0.01"), ("max_iterations_cg", "1000"), ("kappa", "1e9"), ("epsilon_0", "1"),
        ("dielec_ext", "1"), ("do_stokes", "1"), ("stokes_visc", "1e-03"), ("calc_vdw", "0"),
        ("calc_noise", "1"), ("calc_kinetics", "0"), ("calc_es", "0"), ("es_update", "1"), ("es_N_x", "0"),
        ("es_N_y", "0"), ("es_N_z", "0"), ("sticky_wall_xz", "0"), ("es_h", "3")]


def sphere_script(d, name, *, centroids=((0.0, 0.0, 0.0),), traj=None, meas=None, extra=(), **over):
    """writes <d>/<name>.ffea over the sphere_63_120 files in <d>/structure"""
    s = d / "structure"
    if not s.exists():
        shutil.copytree(GOLD, s)
    vals = dict(BASE)
    vals["trajectory_out_fname"] = traj or f"{name}_trajectory.out"
    vals["measurement_out_fname"] = meas or f"{name}_measurement.out"
    vals.update({k: str(v) for k, v in over.items()})
    nb = len(centroids)
    lines = ["<param>"] + [f"\t<{k} = {vals[k]}>" for k, _ in BASE] + [f"\t<{k} = {v}>" for k, v in extra]
    lines += [f"\t<num_blobs = {nb}>", "\t<num_conformations = (%s)>" % ",".join(["1"] * nb),
              "\t<num_states = (%s)>" % ",".join(["1"] * nb), "</param>", "", "<system>"]
    for c in centroids:
        lines += ["\t<blob>", "\t\t<conformation>", "\t\t\t<motion_state = DYNAMIC>"]
        for key, ext in (("nodes", "node"), ("topology", "top"), ("surface", "surf"), ("material", "mat"),
                         ("stokes", "stokes"), ("vdw", "vdw"), ("pin", "pin")):
            lines.append(f"\t\t\t<{key} = structure/sphere_63_120.{ext}>")
        lines += ["\t\t</conformation>", "\t\t<solver = CG_nomass>", "\t\t<scale = 5e-9>",
                  "\t\t<centroid = (%r,%r,%r)>" % tuple(c), "\t</blob>"]
    lines.append("</system>")
    path = d / f"{name}.ffea"
    path.write_text("\n".join(lines) + "\n")
    return str(path)


def read_trajectory(path, num_blobs=1):
    """[(step, [node rows of each blob])] of a .ftj (World.cpp:527-550, 4461-4488)"""
    lines = open(path).read().split("\n")
    assert lines[0] == "FFEA_trajectory_file"
    i = lines.index("*") + 1
    frames = []
    while i < len(lines) and lines[i].startswith("Blob "):
        blobs, step = [], None
        for b in range(num_blobs):
            head = lines[i]
            assert head.startswith(f"Blob {b}, Conformation 0, step "), head
            step = int(head.rsplit(" ", 1)[1])
            assert lines[i + 1] == "DYNAMIC"
            j = i + 2
            rows = []
            while not (lines[j].startswith("Blob") or lines[j] == "*"):
                rows.append([float(x) for x in lines[j].split()])
                j += 1
            blobs.append(np.array(rows))
            i = j
        assert lines[i] == "*" and lines[i + 1] == "Conformation Changes:"
        for b in range(num_blobs):
            assert lines[i + 2 + b] == f"Blob {b}: Conformation 0 -> Conformation 0"
        assert lines[i + 2 + num_blobs] == "*"
        i += 3 + num_blobs
        frames.append((step, blobs))
    assert lines[i:] == [""], lines[i:i + 3]
    return frames


def read_measurements(path):
    txt = open(path).read().split("\n")
    k = txt.index("Measurements:")
    return txt[k + 1].split(), np.array([[float(x) for x in l.split()] for l in txt[k + 2:] if l.strip()])


def read_detailed(path, num_blobs, blob_cols=5, pair_cols=None):
    """header line and rows of a .fdm (World.cpp:606-646, 4504-4537, 4694-4721): per row the
    time, then each blob's fields (blob_cols of them), then each pair's (pair_cols[(i, j)] of
    them, default none); the text of every line is checked to end with the pair blocks'
    7-space leads"""
    txt = open(path).read().split("\n")
    assert txt[:3] == ["FFEA Detailed Measurement File", "", "Measurements:"], txt[:3]
    assert txt[-1] == ""
    pairs = [(i, j) for i in range(num_blobs) for j in range(i, num_blobs)]
    pc = pair_cols or {}
    rows = []
    for line in txt[4:-1]:
        vals = [float(x) for x in line.split()]
        assert len(vals) == 1 + blob_cols * num_blobs + sum(pc.get(q, 0) for q in pairs), line
        if not pc:
            assert line.endswith(" " * 7 * len(pairs)), repr(line[-30:])
        rows.append(vals)
    return txt[3], np.array(rows)


def close(a, b, rtol, atol):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return bool(np.all(np.abs(a - b) <= rtol * np.abs(b) + atol))


def oracle_for(oracle_mod, path, b=0):
    spec = Script(path).blob_spec(b)
    return spec, oracle_mod.Blob(spec.pos, spec.top, spec.A, spec.B, spec.G, spec.E, **spec.oracle_kwargs())


def test_strain_energy_matches_oracle(oracle_mod, tmp_path):
    path = sphere_script(tmp_path, "se")
    spec, o = oracle_for(oracle_mod, path)
    r = np.random.default_rng(4)
    p = spec.pos * np.array([1.02, 0.99, 1.01]) + r.normal(0.0, 0.01, spec.pos.shape)
    o.set_positions(p)
    for precision, tol in (("fp64", 1e-12), ("mixed", 1e-5)):
        print(f"[case] strain energy sphere_63_120 {precision}", flush=True)
        with GpuBlob.from_spec(spec, precision=precision) as g:
            assert g.strain_energy() == pytest.approx(0.0, abs=1e-12 * o.strain_energy())
            g.set_positions(p)
            e = g.strain_energy()
            assert abs(e - o.strain_energy()) <= tol * abs(o.strain_energy()), (precision, e, o.strain_energy())


def test_world_files_against_oracle(oracle_mod, tmp_path):
    path = sphere_script(tmp_path, "sphere", num_steps=6, check=2)
    spec, o = oracle_for(oracle_mod, path)
    print("[case] world sphere_63_120 6 steps check 2", flush=True)
    with World(path) as w:
        assert w.step == 0
        w.run()
        assert w.step == 7
        gp = w.blob(0).positions()
    frames = read_trajectory(str(tmp_path / "sphere_trajectory.out"))
    head, meas = read_measurements(str(tmp_path / "sphere_measurement.out"))
    assert head == ["Time", "StrainEnergy", "Centroid[0]", "Centroid[1]", "Centroid[2]", "RMSD"]
    assert [f[0] for f in frames] == [0, 2, 4, 6] and meas.shape == (4, 6)
    pos0 = spec.pos.copy()
    scale = np.abs(pos0).max() * units.LENGTH
    for n in range(7):
        if n % 2 == 0:
            state = o.rng_state()
            p = o.positions()
            f = frames[n // 2][1][0]
            assert f.shape == (spec.pos.shape[0], 10) and not f[:, 3:].any()
            assert close(f[:, :3], p * units.LENGTH, PRINT_RTOL, 1e-12 * scale), n
            strain = o.strain_energy() * units.ENERGY
            cog = p.mean(axis=0) * units.LENGTH
            rmsd = np.sqrt(((p - pos0) ** 2).sum(axis=1).mean()) * units.LENGTH
            row = meas[n // 2]
            assert close(row[0], n * spec.dt * units.TIME, PRINT_RTOL, 0.0)
            assert close(row[1], strain, 1e-5, 1e-9 * 4.11e-21), (n, row[1], strain)
            assert close(row[2:5], cog, PRINT_RTOL, 1e-12 * scale), (n, row[2:5], cog)
            assert close(row[5], rmsd, 1e-5, 1e-12 * scale), (n, row[5], rmsd)
        assert o.step() == 0
    assert rel_err(gp, o.positions()) <= 1e-10
    cp = open(tmp_path / "sphere.fcp").read().split("\n")
    assert cp[0] == "RNGStreams dedicated to the thermal stress: 1"
    assert [int(x) for x in cp[1].split()] == state.tolist()


def test_restart_matches_uninterrupted_run(tmp_path):
    """the reference's sphere_63_120_nomass_restart: runs I (0-10), II (0-5) and III
    (restart of II to 10) with one thermal stream"""
    full = sphere_script(tmp_path, "sphere_63_120_nomass_10steps", num_steps=10,
                         traj="sphere_63_120_nomass_I_trajectory.out", meas="sphere_63_120_nomass_I_measurement.out")
    first = sphere_script(tmp_path, "sphere_63_120_nomass_1-5steps", num_steps=5,
                          traj="sphere_63_120_nomass_II_trajectory.out",
                          meas="sphere_63_120_nomass_II_measurement.out")
    second = sphere_script(tmp_path, "sphere_63_120_nomass_6-10steps", num_steps=10, restart=1,
                           traj="sphere_63_120_nomass_II_trajectory.out",
                           meas="sphere_63_120_nomass_II_measurement.out",
                           extra=[("checkpoint_in", "sphere_63_120_nomass_1-5steps.fcp")])
    for path in (full, first, second):
        print(f"[case] world restart {os.path.basename(path)}", flush=True)
        with World(path) as w:
            if path == second:
                assert w.step == 5
            w.run()
            assert w.step == (6 if path == first else 11)
    a = open(tmp_path / "sphere_63_120_nomass_6-10steps.fcp").read().split("\n")
    b = open(tmp_path / "sphere_63_120_nomass_10steps.fcp").read().split("\n")
    assert a[:2] == b[:2]
    _, m1 = read_measurements(str(tmp_path / "sphere_63_120_nomass_I_measurement.out"))
    _, m2 = read_measurements(str(tmp_path / "sphere_63_120_nomass_II_measurement.out"))
    assert m1.shape == m2.shape == (11, 6)
    with np.errstate(divide="ignore", invalid="ignore"):
        d = np.where(m1 != 0, np.abs((m2 - m1) / m1), 0.0)
    print("restart DMax:", d.max())
    assert d.max() <= 5e-4
    # the detailed files, truncated and appended the same way (World.cpp:867-930)
    _, f1 = read_detailed(str(tmp_path / "sphere_63_120_nomass_I_measurement.fdm"), 1)
    _, f2 = read_detailed(str(tmp_path / "sphere_63_120_nomass_II_measurement.fdm"), 1)
    assert f1.shape == f2.shape == (11, 6)
    assert np.array_equal(f1[:, 0], m1[:, 0]) and np.array_equal(f2[:, 0], m2[:, 0])
    assert np.array_equal(f1[:5], f2[:5])  # before the restarted step: the same run
    with np.errstate(divide="ignore", invalid="ignore"):
        d2 = np.where(f1 != 0, np.abs((f2 - f1) / f1), 0.0)
    assert d2.max() <= 5e-4
    t1 = read_trajectory(str(tmp_path / "sphere_63_120_nomass_I_trajectory.out"))
    t2 = read_trajectory(str(tmp_path / "sphere_63_120_nomass_II_trajectory.out"))
    assert [f[0] for f in t2] == list(range(11)) == [f[0] for f in t1]
    for (_, x), (_, y) in zip(t1[:5], t2[:5]):
        assert np.array_equal(x[0], y[0])  # before the restarted step: the same run


def test_existing_outputs_are_moved_aside(tmp_path):
    path = sphere_script(tmp_path, "again", num_steps=1)
    for k in range(2):
        print(f"[case] world backups run {k}", flush=True)
        with World(path) as w:
            w.run()
    for name in ("again_trajectory.out", "again_measurement.out", "again.fcp"):
        assert (tmp_path / name).exists()
        assert (tmp_path / f"__{name}__bckp.1").exists(), name
    assert open(tmp_path / "again_trajectory.out").read() == open(tmp_path / "__again_trajectory.out__bckp.1").read()


def test_two_blob_world_against_oracle(oracle_mod, tmp_path):
    """two blobs share the stream in turn: each step draws blob 0's elements, then blob 1's"""
    path = sphere_script(tmp_path, "pair", centroids=((0.0, 0.0, 0.0), (3.0, 0.0, 0.0)), num_steps=4, check=2,
                         do_stokes=0)
    s = Script(path)
    specs = [s.blob_spec(b) for b in range(2)]
    orcs = [oracle_mod.Blob(sp.pos, sp.top, sp.A, sp.B, sp.G, sp.E, **sp.oracle_kwargs()) for sp in specs]
    seed = s.rng_state()
    per_step = sum(7 * sp.top.shape[0] for sp in specs)
    print("[case] world two blobs", flush=True)
    with World(path) as w:
        w.run()
        got = [w.blob(b).positions() for b in range(2)]
    frames = read_trajectory(str(tmp_path / "pair_trajectory.out"), num_blobs=2)
    _, meas = read_measurements(str(tmp_path / "pair_measurement.out"))
    dhead, det = read_detailed(str(tmp_path / "pair_measurement.fdm"), 2)
    blob_head = "%-14s%-14s%-14s%-14s%-14s" % ("StrainEnergy", "Centroid[0]", "Centroid[1]", "Centroid[2]", "RMSD")
    assert dhead == "%-14s" % "Time" + "| B0 " + blob_head + "| B1 " + blob_head, repr(dhead)
    assert det.shape == (3, 11)
    assert np.array_equal(det[:, 0], meas[:, 0])
    assert [f[0] for f in frames] == [0, 2, 4]
    pos0 = [sp.pos.copy() for sp in specs]
    scale = max(np.abs(p).max() for p in pos0) * units.LENGTH
    for n in range(5):
        for b, o in enumerate(orcs):
            o.set_rng_state(_native.advance_rng_state(seed, n * per_step + 7 * specs[0].top.shape[0] * b))
        if n % 2 == 0:
            ps = [o.positions() for o in orcs]
            for b in range(2):
                assert close(frames[n // 2][1][b][:, :3], ps[b] * units.LENGTH, PRINT_RTOL, 1e-12 * scale), (n, b)
            allp = np.concatenate(ps)
            strain = sum(o.strain_energy() for o in orcs) * units.ENERGY
            rmsd = np.sqrt(sum(((p - q) ** 2).sum() for p, q in zip(ps, pos0)) / allp.shape[0]) * units.LENGTH
            row = meas[n // 2]
            assert close(row[1], strain, 1e-5, 1e-9 * 4.11e-21), (n, row[1], strain)
            assert close(row[2:5], allp.mean(axis=0) * units.LENGTH, PRINT_RTOL, 1e-12 * scale)
            assert close(row[5], rmsd, 1e-5, 1e-12 * scale)
            for b in range(2):  # the detailed file: each blob's own strain energy, centroid, RMSD
                d = det[n // 2][1 + 5 * b:6 + 5 * b]
                e_b = orcs[b].strain_energy() * units.ENERGY
                r_b = np.sqrt(((ps[b] - pos0[b]) ** 2).sum(axis=1).mean()) * units.LENGTH
                assert close(d[0], e_b, 1e-5, 1e-9 * 4.11e-21), (n, b, d[0], e_b)
                assert close(d[1:4], ps[b].mean(axis=0) * units.LENGTH, PRINT_RTOL, 1e-12 * scale), (n, b)
                assert close(d[4], r_b, 1e-5, 1e-12 * scale), (n, b, d[4], r_b)
        for o in orcs:
            assert o.step() == 0
    for b in range(2):
        assert rel_err(got[b], orcs[b].positions()) <= 1e-10, b


def test_detailed_file_switch_and_backups(tmp_path):
    """the detailed file is written by default (with one blob its columns are the global
    measurement's), moved aside like the others, and not written with detailed=False (the
    reference's ffea -d); det_measurement_out_fname names it"""
    path = sphere_script(tmp_path, "det", num_steps=2)
    print("[case] world detailed file", flush=True)
    for k in range(2):
        with World(path) as w:
            w.run()
    assert (tmp_path / "det_measurement.fdm").exists() and (tmp_path / "__det_measurement.fdm__bckp.1").exists()
    _, meas = read_measurements(str(tmp_path / "det_measurement.out"))
    _, det = read_detailed(str(tmp_path / "det_measurement.fdm"), 1)
    assert det.shape == meas.shape and close(det, meas, PRINT_RTOL, 0.0)  # the same sums, printed
    off = sphere_script(tmp_path, "nodet", num_steps=1)
    with World(off, detailed=False) as w:
        w.run()
    assert (tmp_path / "nodet_measurement.out").exists() and not (tmp_path / "nodet_measurement.fdm").exists()
    named = sphere_script(tmp_path, "named", num_steps=1, extra=[("det_measurement_out_fname", "named.details")])
    with World(named) as w:
        w.run()
    assert (tmp_path / "named.details").exists() and not (tmp_path / "named_measurement.fdm").exists()


def test_restart_deleting_frames(tmp_path):
    """ffea -l 2 (delete_frames): a restart drops the last two complete frames and their
    measurement lines and resumes from the frame before them; with noise off the resumed run
    retraces the uninterrupted one (measurements within 5e-4, as the reference's restart
    check)"""
    full = sphere_script(tmp_path, "fullrun", num_steps=8, calc_noise=0, traj="full_traj.out", meas="full_meas.out")
    part = sphere_script(tmp_path, "part", num_steps=8, calc_noise=0, traj="part_traj.out", meas="part_meas.out")
    again = sphere_script(tmp_path, "again2", num_steps=8, calc_noise=0, restart=1, traj="part_traj.out",
                          meas="part_meas.out", extra=[("checkpoint_in", "part.fcp")])
    print("[case] world restart deleting 2 frames", flush=True)
    for path in (full, part):
        with World(path) as w:
            w.run()
    with World(again, delete_frames=2) as w:
        assert w.step == 6
        w.run()
        assert w.step == 9
    t1 = read_trajectory(str(tmp_path / "full_traj.out"))
    t2 = read_trajectory(str(tmp_path / "part_traj.out"))
    assert [f[0] for f in t2] == list(range(9)) == [f[0] for f in t1]
    _, m1 = read_measurements(str(tmp_path / "full_meas.out"))
    _, m2 = read_measurements(str(tmp_path / "part_meas.out"))
    _, d2 = read_detailed(str(tmp_path / "part_meas.fdm"), 1)
    assert m1.shape == m2.shape == d2.shape == (9, 6)
    assert np.array_equal(m1[:6], m2[:6])
    with np.errstate(divide="ignore", invalid="ignore"):
        d = np.where(m1 != 0, np.abs((m2 - m1) / m1), 0.0)
    assert d.max() <= 5e-4, d.max()
    assert close(d2, m2, PRINT_RTOL, 0.0)
    with pytest.raises(FFEAException):
        World(again, delete_frames=-1)
