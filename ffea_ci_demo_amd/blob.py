"""Python mirror of the reference's per-blob step interface over the C ABI.

`GpuBlob` exposes the pair World::run calls on a blob every step —
update_internal_forces (Blob.cpp:400-491) and update_positions (Blob.cpp:493-508) —
plus node state access, the noise stream state and the solver's right-hand side.
Failures raise `FFEAException` with the reference's messages (inversion,
CG non-convergence).
"""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np

from . import _native


class FFEAException(RuntimeError):
    """Raised with the reference's message (FFEA_return_codes.h:52-69)."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def _opt(a, dtype):
    if a is None:
        return None, None
    arr = np.ascontiguousarray(a, dtype)
    if arr.size == 0:
        return None, arr
    return arr.ctypes.data, arr


class _Handle:
    """what every handle (one blob or an ensemble) offers"""

    h = None

    def close(self):
        if getattr(self, "h", None):
            _native.lib().ffea_gpu_blob_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != _native.OK:
            msg = _native.lib().ffea_gpu_blob_last_error(self.h).decode()
            raise FFEAException(rc, msg)

    def step(self, nsteps=1):
        """nsteps x (update_internal_forces; update_positions) for every blob of the handle"""
        self._check(_native.lib().ffea_gpu_blob_step(self.h, int(nsteps)))

    def info(self):
        i = _native.Info()
        self._check(_native.lib().ffea_gpu_blob_get_info(self.h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in _native.Info._fields_}

    def strain_energies(self):
        """Blob::calculate_strain_energy (Blob.cpp:1052-1064) of every blob at its current
        positions, meso energy units"""
        out = np.zeros(self.info()["num_blobs"])
        self._check(_native.lib().ffea_gpu_batch_strain_energy(self.h, out))
        return out

    def set_row_blocks(self, blocks):
        self._check(_native.lib().ffea_gpu_blob_set_row_blocks(self.h, int(blocks)))

    def autotune(self, iterations=20):
        """time CG iterations at each candidate launch shape, keep the fastest (result-neutral)"""
        chosen = C.c_int()
        self._check(_native.lib().ffea_gpu_blob_autotune(self.h, int(iterations), C.byref(chosen)))
        return chosen.value

    def launch_shape(self):
        """{"row_blocks", "spmv_unroll", "assembly_lanes"}: the launch-shape knobs in force"""
        v = [C.c_int() for _ in range(3)]
        self._check(_native.lib().ffea_gpu_blob_get_launch_shape(self.h, *[C.byref(x) for x in v]))
        return {"row_blocks": v[0].value, "spmv_unroll": v[1].value, "assembly_lanes": v[2].value}

    def set_launch_shape(self, row_blocks=0, spmv_unroll=0, assembly_lanes=0):
        """sets the launch-shape knobs (0: unchanged); results do not depend on them.
        assembly_lanes: 1, 4 or 16 lanes per row, or -1 (shape records staged in LDS)"""
        self._check(_native.lib().ffea_gpu_blob_set_launch_shape(self.h, int(row_blocks), int(spmv_unroll),
                                                                 int(assembly_lanes)))

    def use_graph(self, enable=True, iterations=0):
        """run each step() as one captured HIP graph launch holding `iterations` CG
        iterations (0: a little over the last step's count)"""
        self._check(_native.lib().ffea_gpu_blob_use_graph(self.h, 1 if enable else 0, int(iterations)))

    def set_refinement(self, tolerance):
        """FP64 refinement of the mixed form's solves to |b - V v| <= tolerance |b| (0: off)"""
        self._check(_native.lib().ffea_gpu_blob_set_refinement(self.h, float(tolerance)))
        self._refine_tol = float(tolerance)

    def refinement_sweeps(self):
        """(correction solves of the last solve, in all)"""
        last, total = C.c_int(), C.c_longlong()
        self._check(_native.lib().ffea_gpu_blob_get_refinement(self.h, C.byref(last), C.byref(total)))
        return last.value, total.value

    def set_preconditioner(self, kind="jacobi"):
        """CG preconditioner: "jacobi" (the reference's) or "multigrid" (a smoothed-
        aggregation V-cycle built from the matrix at the current positions); returns the
        multigrid level info (empty for Jacobi)"""
        k = {"jacobi": _native.PRECOND_JACOBI, "multigrid": _native.PRECOND_MULTIGRID}[kind]
        self._check(_native.lib().ffea_gpu_blob_set_preconditioner(self.h, k))
        return self.multigrid_info()

    def multigrid_info(self):
        i = _native.MultigridInfo()
        self._check(_native.lib().ffea_gpu_blob_get_multigrid_info(self.h, C.byref(i)))
        return i.as_dict()

    def set_timing(self, phases=None):
        """time the named phases (default: all); set_timing(()) turns timing off"""
        names = _native.PHASES if phases is None else phases
        mask = 0
        for nm in names:
            mask |= 1 << _native.PHASES.index(nm)
        self._check(_native.lib().ffea_gpu_blob_set_timing(self.h, mask))

    def timing(self):
        ms = np.zeros(_native.NUM_PHASES)
        n = np.zeros(_native.NUM_PHASES, np.int64)
        self._check(_native.lib().ffea_gpu_blob_get_timing(self.h, ms, n))
        return {name: (float(ms[i]), int(n[i])) for i, name in enumerate(_native.PHASES)}

    def synchronize(self):
        self._check(_native.lib().ffea_gpu_blob_synchronize(self.h))

    def set_solver(self, kind, rho=None):
        """\"masslumped\" (MassLumpedSolver) or \"cg\" (ConjugateGradientSolver, consistent
        mass), both with rho per element in meso units, or \"nomass\" (the default mass-less CG)"""
        L = _native.lib()
        if kind in ("masslumped", "cg"):
            r = np.asarray(rho, np.float64)
            r = np.ascontiguousarray(np.broadcast_to(r, (self.ne,)) if r.ndim == 0 else r.ravel())
            code = _native.SOLVER_MASSLUMPED if kind == "masslumped" else _native.SOLVER_ITERATIVE
            self.run_generation = getattr(self, "run_generation", 0) + 1  # the library drops a kept run
            self._check(L.ffea_gpu_blob_set_solver(self.h, code, r.ctypes.data))
        elif kind == "nomass":
            self._check(L.ffea_gpu_blob_set_solver(self.h, _native.SOLVER_NOMASS_CG, None))
        else:
            raise ValueError(f"unknown solver {kind!r}")


def _node_vectors(a, nt):
    """one FP64 3-vector per node, flattened; the library reads exactly 3 nt values"""
    v = np.ascontiguousarray(a, np.float64).reshape(-1)
    if v.size != 3 * nt:
        raise ValueError(f"expected {nt} node 3-vectors ({3 * nt} values), got {v.size} values")
    return v


def _params(dt, kT, epsilon2, max_iterations_cg, calc_noise, calc_stokes, precision, device):
    prec = {"fp64": _native.FP64, "mixed": _native.MIXED}[precision]
    return _native.Params(float(dt), float(kT), float(epsilon2), int(max_iterations_cg), int(calc_noise),
                          int(calc_stokes), prec, int(device))


class GpuBlob(_Handle):
    run_generation = 0  # bumped by every checkpointed run and material change

    def __init__(self, pos, top, A, B, G, E, *, dt, kT, epsilon2, max_iterations_cg=1000, calc_noise=1,
                 calc_stokes=0, pinned=(), ctf_node=(), ctf_force=(), stokes_drag=None, rng_state=None,
                 surface_forces=None, rotational_forces=None, external_nodes=(), precision="fp64", device=0):
        """surface_forces: None or (face_nodes [nf][3], set_force [ns][3] meso, set_ptr [ns+1],
        set_faces): linear surface constant forces (Blob::load_ctforces, Blob.cpp:2730-2806);
        rotational_forces: None or (node [n], force [n] meso, axis [n][6] meso point and unit
        direction, type [n] of "pf" / "pt"): rotational constant forces (Blob.cpp:2650-2726);
        external_nodes: the nodes add_node_forces may load (World::apply_springs)"""
        L = _native.lib()
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        top = np.ascontiguousarray(top, np.int32).reshape(-1, 10)
        ne = top.shape[0]

        def per_elem(v):
            return np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (ne,)))

        mats = [per_elem(A), per_elem(B), per_elem(G), per_elem(E)]
        p = _params(dt, kT, epsilon2, max_iterations_cg, calc_noise, calc_stokes, precision, device)
        pin_p, pins = _opt(pinned, np.int32)
        cn_p, cn = _opt(ctf_node, np.int32)
        cf_p, cf = _opt(ctf_force, np.float64)
        drag_p, drag = _opt(stokes_drag, np.float64)
        npin = 0 if pins is None else pins.size
        nctf = 0 if cn is None else cn.size
        st = np.ascontiguousarray(rng_state if rng_state is not None else [12345] * 6, np.uint32)
        h = C.c_void_p()
        sf = None
        if surface_forces is not None:
            fn, sfo, sp, sfc = surface_forces
            keep = [np.ascontiguousarray(fn, np.int32).reshape(-1), np.ascontiguousarray(sfo, np.float64).reshape(-1),
                    np.ascontiguousarray(sp, np.int32), np.ascontiguousarray(sfc, np.int32)]
            sf = _native.SurfaceForces(keep[0].size // 3, keep[0].ctypes.data, keep[2].size - 1, keep[1].ctypes.data,
                                       keep[2].ctypes.data, keep[3].ctypes.data)
        rf = None
        if rotational_forces is not None:
            rn, rfo, rax, rty = rotational_forces
            rkeep = [np.ascontiguousarray(rn, np.int32).ravel(), np.ascontiguousarray(rfo, np.float64).ravel(),
                     np.ascontiguousarray(rax, np.float64).reshape(-1),
                     np.frombuffer("".join(str(t) for t in rty).encode("ascii"), np.uint8).copy()]
            n = rkeep[0].size
            if rkeep[1].size != n or rkeep[2].size != 6 * n or rkeep[3].size != 2 * n:
                raise FFEAException(_native.ERR_INVALID, "rotational forces: node, force, axis and type sizes differ")
            rf = _native.RotationalForces(n, rkeep[0].ctypes.data, rkeep[1].ctypes.data, rkeep[2].ctypes.data,
                                          rkeep[3].ctypes.data)
        ext_p, ext = _opt(external_nodes, np.int32)
        rc = L.ffea_gpu_blob_create_ex(C.byref(p), pos.shape[0], ne, pos, top, *mats, npin, pin_p, nctf, cn_p, cf_p,
                                       drag_p, st, None if sf is None else C.byref(sf),
                                       None if rf is None else C.byref(rf), 0 if ext is None else ext.size, ext_p,
                                       C.byref(h))
        if rc != _native.OK:
            raise FFEAException(rc, L.ffea_gpu_last_error().decode())
        self.h = h
        self.nt = pos.shape[0]
        self.ne = ne
        self.precision = precision

    @classmethod
    def from_spec(cls, spec, precision="fp64", **kw):
        g = cls._from_spec(spec, precision, **kw)
        if getattr(spec, "solver", "nomass") != "nomass":
            try:
                g.set_solver(spec.solver, spec.rho)
            except Exception:
                g.close()
                raise
        return g

    @classmethod
    def _from_spec(cls, spec, precision="fp64", **kw):
        return cls(spec.pos, spec.top, spec.A, spec.B, spec.G, spec.E, dt=spec.dt, kT=spec.kT,
                   epsilon2=spec.epsilon2, max_iterations_cg=spec.max_iterations_cg, calc_noise=spec.calc_noise,
                   calc_stokes=spec.calc_stokes, pinned=spec.pinned, ctf_node=spec.ctf_node,
                   ctf_force=spec.ctf_force, stokes_drag=spec.stokes_drag, rng_state=spec.rng_state,
                   surface_forces=getattr(spec, "surface_forces", None),
                   rotational_forces=getattr(spec, "rotational_forces", None),
                   external_nodes=getattr(spec, "external_nodes", ()), precision=precision, **kw)

    # -- the reference's per-step interface
    def update_internal_forces(self):
        self._check(_native.lib().ffea_gpu_blob_update_internal_forces(self.h))

    def update_positions(self):
        self._check(_native.lib().ffea_gpu_blob_update_positions(self.h))

    def strain_energy(self):
        return float(self.strain_energies()[0])

    # -- state
    def _get(self, fn, shape):
        out = np.zeros(shape, np.float64)
        self._check(getattr(_native.lib(), fn)(self.h, out.reshape(-1)))
        return out

    def positions(self):
        return self._get("ffea_gpu_blob_get_positions", (self.nt, 3))

    def kinetic_energy(self):
        e = C.c_double()
        self._check(_native.lib().ffea_gpu_blob_kinetic_energy(self.h, C.byref(e)))
        return e.value

    def set_velocities(self, v):
        self._check(_native.lib().ffea_gpu_blob_set_velocities(self.h, _node_vectors(v, self.nt)))

    def add_node_forces(self, nodes, forces):
        """Blob::add_force_to_node for the next step (meso units, list order)"""
        n = np.ascontiguousarray(nodes, np.int32).ravel()
        f = np.ascontiguousarray(forces, np.float64).reshape(-1)
        if f.size != 3 * n.size:
            raise ValueError("one 3-vector of force per node")
        self._check(_native.lib().ffea_gpu_blob_add_node_forces(self.h, n.size, n if n.size else np.zeros(1, np.int32),
                                                               f if f.size else np.zeros(3)))

    def node_positions(self, nodes):
        """Blob::get_node for each of the given nodes, [n][3]"""
        n = np.ascontiguousarray(nodes, np.int32).ravel()
        out = np.zeros((n.size, 3))
        if n.size:
            self._check(_native.lib().ffea_gpu_blob_get_node_positions(self.h, n.size, n, out.reshape(-1)))
        return out

    def set_positions(self, p):
        self._check(_native.lib().ffea_gpu_blob_set_positions(self.h, _node_vectors(p, self.nt)))

    def velocities(self):
        return self._get("ffea_gpu_blob_get_velocities", (self.nt, 3))

    def rhs(self):
        return self._get("ffea_gpu_blob_get_rhs", (self.nt, 3))

    def element_forces(self):
        return self._get("ffea_gpu_blob_get_element_forces", (self.ne, 4, 3))

    def cg_iterations(self):
        last = C.c_int()
        total = C.c_longlong()
        self._check(_native.lib().ffea_gpu_blob_get_cg_iterations(self.h, C.byref(last), C.byref(total)))
        return last.value, total.value

    def rng_state(self):
        out = np.zeros(6, np.uint32)
        self._check(_native.lib().ffea_gpu_blob_get_rng_state(self.h, out))
        return out

    def set_rng_state(self, s):
        self._check(_native.lib().ffea_gpu_blob_set_rng_state(self.h, np.ascontiguousarray(s, np.uint32)))

    def set_rng_stream(self, s, draws_per_step):
        self._check(_native.lib().ffea_gpu_blob_set_rng_stream(self.h, np.ascontiguousarray(s, np.uint32),
                                                               int(draws_per_step)))

    # -- differentiation (noise off, some nodes pinned)
    def set_material(self, A, B, G, E):
        """per-element A, B, G, E in meso units (scalars broadcast)"""
        arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (self.ne,))) for v in (A, B, G, E)]
        self.run_generation += 1  # the library drops a kept run made with the earlier material
        self._check(_native.lib().ffea_gpu_blob_set_material(self.h, *arrs))

    def run_checkpointed(self, nsteps, interval=20):
        """nsteps steps from the current state, positions kept every `interval` steps;
        `run_generation` names the run kept (it changes with every run and material)"""
        self.run_generation += 1
        self._check(_native.lib().ffea_gpu_blob_run_checkpointed(self.h, int(nsteps), int(interval)))

    def adjoint(self, dL_dpos):
        """reverse of the last checkpointed run: dL/dpos at its end -> (dL/dA, dL/dB, dL/dG,
        dL/dE per element, dL/dpos at its start)"""
        msg = mixed_adjoint_warning(self.precision, getattr(self, "_refine_tol", 0.0))
        if msg:
            warnings.warn(msg, RuntimeWarning, stacklevel=2)
        g = _node_vectors(dL_dpos, self.nt)
        outs = [np.zeros(self.ne) for _ in range(4)]
        d0 = np.zeros((self.nt, 3))
        self._check(_native.lib().ffea_gpu_blob_adjoint(self.h, g, *[o.ctypes.data for o in outs], d0.ctypes.data))
        return outs[0], outs[1], outs[2], outs[3], d0


def mixed_adjoint_warning(precision, refine_tol):
    """the warning GpuBlob.adjoint gives for an FP32-mixed run without FP64 refinement, or None:
    on a static, pinned problem the FP32 operator's rounding is amplified by the condition
    number (gradients 19-20% off FP64 on a 50^3 blob; with refinement 1e-5)"""
    if precision == "mixed" and not refine_tol:
        return ("adjoint of an FP32-mixed run without FP64 refinement: on static, pinned problems the gradients "
                "can be far from FP64's (19-20% on a 50^3 blob); call set_refinement(1e-10) before the run")
    return None


class GpuSolver(_Handle):
    """The mass-less viscosity solver alone, the reference's NoMassCGSolver behind its Solver
    interface (Solver.h:31-52): created from the mesh, the per-element viscosities A, B, the
    pins and (calc_stokes) the drag (NoMassCGSolver::init, NoMassCGSolver.cpp:56-153);
    solve(pos, x) is NoMassCGSolver::solve (NoMassCGSolver.cpp:156-188) with the viscosity
    matrix built at the node positions pos: x, the force vector after linearise_force and
    pin zeroing, in; the velocities out. The caller's element loop forms the forces."""

    def __init__(self, pos, top, A, B, *, dt, epsilon2, max_iterations_cg=1000, calc_stokes=0, pinned=(),
                 stokes_drag=None, precision="fp64", device=0, kT=0.0):
        L = _native.lib()
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        top = np.ascontiguousarray(top, np.int32).reshape(-1, 10)
        self.nt, self.ne = pos.shape[0], top.shape[0]
        A = np.ascontiguousarray(np.broadcast_to(np.asarray(A, np.float64), (self.ne,)))
        B = np.ascontiguousarray(np.broadcast_to(np.asarray(B, np.float64), (self.ne,)))
        pin_p, pin = _opt(pinned, np.int32)
        npin = 0 if pin is None else pin.size
        drag = None
        if calc_stokes:
            drag = np.ascontiguousarray(stokes_drag, np.float64).reshape(-1)
            if drag.size != self.nt:
                raise ValueError("stokes_drag needs one value per node")
        prm = _params(dt, kT, epsilon2, max_iterations_cg, 0, calc_stokes, precision, device)
        h = C.c_void_p()
        rc = L.ffea_gpu_solver_create(C.byref(prm), self.nt, self.ne, pos.reshape(-1), top.reshape(-1), A, B,
                                      npin, pin_p, drag.ctypes.data if drag is not None else None, C.byref(h))
        if rc != _native.OK:
            raise FFEAException(rc, L.ffea_gpu_last_error().decode())
        self.h = h

    @classmethod
    def from_spec(cls, spec, precision="fp64", **kw):
        return cls(spec.pos, spec.top, spec.A, spec.B, dt=spec.dt, epsilon2=spec.epsilon2,
                   max_iterations_cg=spec.max_iterations_cg, calc_stokes=spec.calc_stokes, pinned=spec.pinned,
                   stokes_drag=spec.stokes_drag, precision=precision, kT=spec.kT, **kw)

    def solve(self, pos, rhs):
        """the solution of V(pos) v = rhs, [num_nodes][3] (rhs is not modified)"""
        p = _node_vectors(pos, self.nt)
        x = _node_vectors(rhs, self.nt).copy()
        self._check(_native.lib().ffea_gpu_solver_solve(self.h, p, x))
        return x.reshape(self.nt, 3)

    def apply_matrix(self, pos, v):
        """V(pos) v over every row of the reference's matrix, [num_nodes][3]"""
        out = np.zeros(3 * self.nt)
        self._check(_native.lib().ffea_gpu_solver_apply_matrix(self.h, _node_vectors(pos, self.nt),
                                                               _node_vectors(v, self.nt), out))
        return out.reshape(self.nt, 3)

    def cg_iterations(self):
        last = C.c_int()
        total = C.c_longlong()
        self._check(_native.lib().ffea_gpu_blob_get_cg_iterations(self.h, C.byref(last), C.byref(total)))
        return last.value, total.value


class GpuBatch(_Handle):
    """An ensemble of independent blobs stepped together in one handle.

    `specs` are BlobSpec-like objects (pos, top, A, B, G, E, pinned, ctf_node, ctf_force,
    rng_state); run parameters (dt, kT, epsilon2, max_iterations_cg, calc_noise) are
    shared and taken from the first spec. The reference steps the blobs of a World one
    after another (World.cpp:2031-2059); every kernel here covers all of them.
    """

    def __init__(self, specs, precision="fp64", device=0):
        L = _native.lib()
        specs = list(specs)
        s0 = specs[0]
        for s in specs[1:]:
            if (s.dt, s.kT, s.epsilon2, s.max_iterations_cg, s.calc_noise, s.calc_stokes) != \
                    (s0.dt, s0.kT, s0.epsilon2, s0.max_iterations_cg, s0.calc_noise, s0.calc_stokes):
                raise ValueError("blobs of one batch share dt, kT, epsilon2, max_iterations_cg, calc_noise and "
                                 "calc_stokes")
        p = _params(s0.dt, s0.kT, s0.epsilon2, s0.max_iterations_cg, s0.calc_noise, s0.calc_stokes, precision, device)
        args, keep = _batch_arrays(specs)
        h = C.c_void_p()
        rc = L.ffea_gpu_batch_create(C.byref(p), *args, C.byref(h))
        del keep
        if rc != _native.OK:
            raise FFEAException(rc, L.ffea_gpu_last_error().decode())
        self.h = h
        self.num_blobs = len(specs)
        self.nt = [s.pos.reshape(-1, 3).shape[0] for s in specs]
        self.ne = [s.top.reshape(-1, 10).shape[0] for s in specs]
        self.precision = precision

    def _get(self, fn, b, shape):
        out = np.zeros(shape, np.float64)
        self._check(getattr(_native.lib(), fn)(self.h, int(b), out.reshape(-1)))
        return out

    def positions(self, b):
        return self._get("ffea_gpu_batch_get_positions", b, (self.nt[b], 3))

    def set_positions(self, b, p):
        if not 0 <= int(b) < len(self.nt):
            raise IndexError(f"blob {b} is not in this ensemble of {len(self.nt)}")
        self._check(_native.lib().ffea_gpu_batch_set_positions(self.h, int(b), _node_vectors(p, self.nt[b])))

    def velocities(self, b):
        return self._get("ffea_gpu_batch_get_velocities", b, (self.nt[b], 3))

    def rhs(self, b):
        return self._get("ffea_gpu_batch_get_rhs", b, (self.nt[b], 3))

    def element_forces(self, b):
        return self._get("ffea_gpu_batch_get_element_forces", b, (self.ne[b], 4, 3))

    def cg_iterations(self):
        last = np.zeros(self.num_blobs, np.int32)
        total = np.zeros(self.num_blobs, np.int64)
        self._check(_native.lib().ffea_gpu_batch_get_cg_iterations(self.h, last, total))
        return last, total

    def rng_state(self, b):
        out = np.zeros(6, np.uint32)
        self._check(_native.lib().ffea_gpu_batch_get_rng_state(self.h, int(b), out))
        return out

    def set_rng_stream(self, b, s, draws_per_step=-1):
        self._check(_native.lib().ffea_gpu_batch_set_rng_stream(self.h, int(b), np.ascontiguousarray(s, np.uint32),
                                                                int(draws_per_step)))


def _batch_arrays(specs):
    """per-blob arrays of ffea_gpu_batch_create (after the params), and the numpy/ctypes
    objects that must stay alive during the call"""
    nb = len(specs)
    keep = []

    def arr(a, dtype, shape=None):
        x = np.ascontiguousarray(a, dtype)
        if shape is not None:
            x = x.reshape(shape)
        keep.append(x)
        return x

    pos = [arr(s.pos, np.float64, (-1, 3)) for s in specs]
    top = [arr(s.top, np.int32, (-1, 10)) for s in specs]
    ne = [t.shape[0] for t in top]
    mats = {k: [arr(np.broadcast_to(np.asarray(getattr(s, k), np.float64), (n,)), np.float64)
                for s, n in zip(specs, ne)] for k in "ABGE"}
    pins = [arr(s.pinned if s.pinned is not None else [], np.int32) for s in specs]
    cn = [arr(s.ctf_node if s.ctf_node is not None else [], np.int32) for s in specs]
    cf = [arr(s.ctf_force if s.ctf_force is not None else np.zeros((0, 3)), np.float64, (-1, 3)) for s in specs]
    rng = arr(np.stack([np.asarray(s.rng_state, np.uint32) for s in specs]).reshape(-1), np.uint32)
    stokes = [arr(s.stokes_drag if s.stokes_drag is not None else [], np.float64) for s in specs]

    def ptrs(lst):
        a = (C.c_void_p * nb)(*[x.ctypes.data if x.size else None for x in lst])
        keep.append(a)
        return C.cast(a, C.POINTER(C.c_void_p))

    npin = arr([x.size for x in pins], np.int32)
    nctf = arr([x.size for x in cn], np.int32)
    args = (nb, arr([x.shape[0] for x in pos], np.int32), arr(ne, np.int32), ptrs(pos), ptrs(top), ptrs(mats["A"]),
            ptrs(mats["B"]), ptrs(mats["G"]), ptrs(mats["E"]), npin.ctypes.data, C.cast(ptrs(pins), C.c_void_p),
            nctf.ctypes.data, C.cast(ptrs(cn), C.c_void_p), C.cast(ptrs(cf), C.c_void_p),
            C.cast(ptrs(stokes), C.c_void_p) if any(x.size for x in stokes) else None, rng)
    return args, keep


def check_layout(specs, precision="fp64"):
    """Build the device tables of a batch on the host (no device) and check every kernel
    index against its buffer; returns the sizes (ffea_gpu_check_layout)."""
    L = _native.lib()
    specs = list(specs)
    s0 = specs[0]
    p = _params(s0.dt, s0.kT, s0.epsilon2, s0.max_iterations_cg, s0.calc_noise, s0.calc_stokes, precision, 0)
    args, keep = _batch_arrays(specs)
    info = _native.Info()
    rc = L.ffea_gpu_check_layout(C.byref(p), *args, C.byref(info))
    del keep
    if rc != _native.OK:
        raise FFEAException(rc, L.ffea_gpu_last_error().decode())
    return {f: getattr(info, f) for f, _ in _native.Info._fields_}
