"""ORACLE — test infrastructure only.

ctypes access to oracle/libffea_oracle.so (the C restatement of the reference's
blob step) and, where built, to oracle/_ref/libffea_ref.so (the reference's own
RNG and sparse-matrix translation units). Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FFEA_ORACLE_LIB names another build of the same source (the OpenMP timing build that
# oracle/cpu_baseline.py loads); the default is the single-threaded checker
LIB_PATH = os.environ.get("FFEA_ORACLE_LIB") or os.path.join(HERE, "libffea_oracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libffea_ref.so")
REF_TEST_RNG = os.path.join(HERE, "_ref", "testRngStream")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


class OrcRng(C.Structure):
    _fields_ = [("Cg", C.c_double * 6), ("Bg", C.c_double * 6), ("Ig", C.c_double * 6),
                ("anti", C.c_int), ("inc_prec", C.c_int)]


class OrcPackage(C.Structure):
    _fields_ = [("next", C.c_double * 6)]


class OrcParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("kT", C.c_double), ("epsilon2", C.c_double),
                ("max_iterations_cg", C.c_int), ("calc_noise", C.c_int), ("calc_stokes", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "libffea_oracle.so"], check=True)


def build_ref(ref_root="/root/reference"):
    """Build oracle/_ref from the reference sources when the tree is present."""
    if not os.path.isdir(ref_root):
        return False
    subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={ref_root}"], check=True)
    return True


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.orc_rng_package_default.argtypes = [P(OrcPackage)]
        L.orc_rng_package_set_seed.argtypes = [P(OrcPackage), _up]
        L.orc_rng_new_stream.argtypes = [P(OrcRng), P(OrcPackage)]
        L.orc_rng_set_seed.argtypes = [P(OrcRng), _up]
        L.orc_rng_get_state.argtypes = [P(OrcRng), _up]
        L.orc_rng_u01.argtypes = [P(OrcRng)]
        L.orc_rng_u01.restype = C.c_double
        L.orc_rng_rand_u01.argtypes = [P(OrcRng)]
        L.orc_rng_rand_u01.restype = C.c_double
        L.orc_rng_rand_int.argtypes = [P(OrcRng), C.c_int, C.c_int]
        L.orc_rng_fill_u01.argtypes = [P(OrcRng), C.c_int, _dp]
        L.orc_rng_advance_state.argtypes = [P(OrcRng), C.c_int32, C.c_int32]
        for f in ("reset_start_stream", "reset_start_substream", "reset_next_substream"):
            getattr(L, "orc_rng_" + f).argtypes = [P(OrcRng)]
        L.orc_rng_set_antithetic.argtypes = [P(OrcRng), C.c_int]
        L.orc_rng_increased_precis.argtypes = [P(OrcRng), C.c_int]
        L.orc_sixseed_from_rng_seed.argtypes = [C.c_int, _up]
        L.orc_blob_create.argtypes = [C.c_int, C.c_int, _dp, _ip, _dp, _dp, _dp, _dp, C.c_int, _ip,
                                      C.c_int, _ip, _dp, C.c_void_p, P(OrcParams), _up]
        L.orc_blob_create.restype = C.c_void_p
        L.orc_blob_destroy.argtypes = [C.c_void_p]
        L.orc_blob_set_surface_forces.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int, _dp, _ip, _ip]
        L.orc_blob_set_surface_forces.restype = C.c_int
        L.orc_blob_set_rotational_forces.argtypes = [C.c_void_p, C.c_int, _ip, _dp, _dp, _ip]
        L.orc_blob_set_rotational_forces.restype = C.c_int
        L.orc_rotational_force.argtypes = [_dp, _dp, C.c_double, C.c_int, _dp]
        L.orc_blob_set_mass_lumped.argtypes = [C.c_void_p, _dp]
        L.orc_blob_set_mass_lumped.restype = C.c_int
        L.orc_blob_set_mass_consistent.argtypes = [C.c_void_p, _dp]
        L.orc_blob_set_mass_consistent.restype = C.c_int
        L.orc_blob_kinetic_energy.argtypes = [C.c_void_p]
        L.orc_blob_kinetic_energy.restype = C.c_double
        L.orc_blob_add_node_forces.argtypes = [C.c_void_p, C.c_int, _ip, _dp]
        L.orc_blob_add_node_forces.restype = C.c_int
        L.orc_spring_forces.argtypes = [_dp, _dp, C.c_double, C.c_double, _dp, _dp]
        L.orc_spring_forces.restype = C.c_int
        L.orc_spring_energy.argtypes = [_dp, _dp, C.c_double, C.c_double]
        L.orc_spring_energy.restype = C.c_double
        for f in ("update_internal_forces", "update_positions", "step", "num_nodes", "num_elements",
                  "last_cg_iterations", "matrix_nnz"):
            getattr(L, "orc_blob_" + f).argtypes = [C.c_void_p]
            getattr(L, "orc_blob_" + f).restype = C.c_int
        for f in ("get_positions", "set_positions", "get_velocities", "set_velocities", "get_rhs", "get_element_forces",
                  "get_inverse_diagonal"):
            getattr(L, "orc_blob_" + f).argtypes = [C.c_void_p, _dp]
        L.orc_blob_get_rng_state.argtypes = [C.c_void_p, _up]
        L.orc_blob_set_rng_state.argtypes = [C.c_void_p, _up]
        L.orc_blob_strain_energy.argtypes = [C.c_void_p]
        L.orc_blob_strain_energy.restype = C.c_double
        L.orc_blob_get_rest_state.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_blob_build_matrix.argtypes = [C.c_void_p]
        L.orc_blob_get_matrix.argtypes = [C.c_void_p, _ip, _ip, _dp]
        L.orc_blob_apply_matrix.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_blob_solve.argtypes = [C.c_void_p, _dp]
        L.orc_blob_solve.restype = C.c_int
        L.orc_blob_get_element_matrices.argtypes = [C.c_void_p, _dp]
        L.orc_blob_get_pinned_flags.argtypes = [C.c_void_p, _ip]
        L.orc_blob_time_step.argtypes = [C.c_void_p, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.orc_blob_time_step.restype = C.c_int
        L.orc_element_elastic_force.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_element_shape.argtypes = [_dp, _dp, P(C.c_double)]
        L.orc_element_shape.restype = C.c_double
        L.orc_element_viscosity_matrix.argtypes = [_dp, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_element_noise_stress.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double,
                                               C.c_double, _dp, _dp]
        _lib = L
    return _lib


_ref = None


def ref_lib():
    """The reference's own RNG / sparse units (None when oracle/_ref is not built)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_LIB_PATH):
            return None
        L = C.CDLL(REF_LIB_PATH)
        L.ref_rng_stream_draws.argtypes = [_up, C.c_int, C.c_int, C.c_int, _dp, _up]
        L.ref_rng_advance_draws.argtypes = [_up, C.c_int, C.c_int, C.c_int, _dp, _up]
        L.ref_mat3_invert.argtypes = [_dp, _dp, C.POINTER(C.c_double)]
        L.ref_nomass_matrix.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_nomass_matrix.restype = C.c_int
        L.ref_rotational_force.argtypes = [_dp, _dp, C.c_double, C.c_int, _dp]
        L.ref_spring_forces.argtypes = [_dp, _dp, C.c_double, C.c_double, _dp, _dp]
        L.ref_spring_forces.restype = C.c_int
        L.ref_spring_energy.argtypes = [_dp, _dp, C.c_double, C.c_double]
        L.ref_spring_energy.restype = C.c_double
        _ref = L
    return _ref


def rotational_force(x, axis, F, torque, use_ref=False):
    """Force of one rotational constant force on a node at x (Blob.cpp:1693-1714): the
    oracle's restatement, or with use_ref the reference's own functions (oracle/_ref)."""
    L = ref_lib() if use_ref else lib()
    out = np.zeros(3)
    fn = L.ref_rotational_force if use_ref else L.orc_rotational_force
    fn(np.ascontiguousarray(x, np.float64), np.ascontiguousarray(axis, np.float64), float(F), int(torque), out)
    return out


def spring_forces(x0, x1, k, l, use_ref=False):
    """Forces (f0, f1) of one spring on its nodes at x0 and x1 (World.cpp:4070-4106), or
    None for coincident nodes: the oracle's restatement, or with use_ref the reference's
    own vector functions (oracle/_ref)."""
    L = ref_lib() if use_ref else lib()
    f0, f1 = np.zeros(3), np.zeros(3)
    fn = L.ref_spring_forces if use_ref else L.orc_spring_forces
    if fn(np.ascontiguousarray(x0, np.float64), np.ascontiguousarray(x1, np.float64), float(k), float(l), f0, f1):
        return None
    return f0, f1


def spring_energy(x0, x1, k, l, use_ref=False):
    """0.5 k (|x0 - x1| - l)^2 (World.cpp:4645-4648)"""
    L = ref_lib() if use_ref else lib()
    fn = L.ref_spring_energy if use_ref else L.orc_spring_energy
    return fn(np.ascontiguousarray(x0, np.float64), np.ascontiguousarray(x1, np.float64), float(k), float(l))


def apply_springs(blobs, springs):
    """World::apply_springs (World.cpp:4063-4108) over oracle blobs: springs is a list of
    (k, l, blob0, blob1, node0, node1, active) in file order; each active spring's forces go
    to its two blobs with Blob::add_force_to_node before their next step."""
    for k, l, b0, b1, n0, n1, active in springs:
        if not active:
            continue
        r = spring_forces(blobs[b0].positions()[n0], blobs[b1].positions()[n1], k, l)
        if r is None:
            raise ValueError("Unable to normalize a vector with length zero")
        blobs[b0].add_node_forces([n0], r[0])
        blobs[b1].add_node_forces([n1], r[1])


def springs_energy(blobs, springs):
    """World::make_measurements' spring energy: per (blob0, blob1) cell in file order, then
    the cells row by row (World.cpp:4632-4652, 4110-4123)"""
    nb = len(blobs)
    cell = [[0.0] * nb for _ in range(nb)]
    for k, l, b0, b1, n0, n1, _ in springs:
        cell[b0][b1] += spring_energy(blobs[b0].positions()[n0], blobs[b1].positions()[n1], k, l)
    e = 0.0
    for i in range(nb):
        for j in range(nb):
            e += cell[i][j]
    return e


class Rng:
    """One MRG32k3a stream of the oracle (RngStream semantics)."""

    def __init__(self, state=None, package: "Package" = None):
        self.s = OrcRng()
        if package is not None:
            lib().orc_rng_new_stream(C.byref(self.s), C.byref(package.p))
        elif state is not None:
            ok = lib().orc_rng_set_seed(C.byref(self.s), np.ascontiguousarray(state, np.uint32))
            if not ok:
                raise ValueError("invalid seed")

    def u01(self):
        return lib().orc_rng_rand_u01(C.byref(self.s))

    def fill(self, n):
        out = np.empty(n, np.float64)
        lib().orc_rng_fill_u01(C.byref(self.s), n, out)
        return out

    def rand_int(self, lo, hi):
        return lib().orc_rng_rand_int(C.byref(self.s), lo, hi)

    def advance(self, e, c):
        lib().orc_rng_advance_state(C.byref(self.s), e, c)

    def reset_start_stream(self):
        lib().orc_rng_reset_start_stream(C.byref(self.s))

    def reset_start_substream(self):
        lib().orc_rng_reset_start_substream(C.byref(self.s))

    def reset_next_substream(self):
        lib().orc_rng_reset_next_substream(C.byref(self.s))

    def set_antithetic(self, a):
        lib().orc_rng_set_antithetic(C.byref(self.s), int(a))

    def increased_precis(self, a):
        lib().orc_rng_increased_precis(C.byref(self.s), int(a))

    def state(self):
        out = np.zeros(6, np.uint32)
        lib().orc_rng_get_state(C.byref(self.s), out)
        return out


class Package:
    def __init__(self):
        self.p = OrcPackage()
        lib().orc_rng_package_default(C.byref(self.p))

    def set_seed(self, seed):
        return bool(lib().orc_rng_package_set_seed(C.byref(self.p), np.ascontiguousarray(seed, np.uint32)))


def sixseed(rng_seed: int) -> np.ndarray:
    out = np.zeros(6, np.uint32)
    lib().orc_sixseed_from_rng_seed(int(rng_seed), out)
    return out


class Blob:
    """The oracle's blob, created from arrays in meso units."""

    def __init__(self, pos, top, A, B, G, E, *, dt, kT, epsilon2, max_iterations_cg=10000, calc_noise=1,
                 calc_stokes=0, pinned=(), ctf_node=(), ctf_force=(), stokes_drag=None, rng_state=None,
                 surface_forces=None, rotational_forces=None):
        """surface_forces: None or (face_nodes [nf][3], set_force [ns][3] meso, set_ptr [ns+1],
        set_faces), the linear surface constant forces of Blob::load_ctforces;
        rotational_forces: None or (node [n], force [n] meso, axis [n][6] meso, type [n] of
        "pf" / "pt"), its rotational forces"""
        L = lib()
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        top = np.ascontiguousarray(top, np.int32).reshape(-1, 10)
        ne = top.shape[0]

        def per_elem(v):
            v = np.asarray(v, np.float64)
            return np.ascontiguousarray(np.broadcast_to(v, (ne,)))

        self._keep = [per_elem(A), per_elem(B), per_elem(G), per_elem(E)]
        pins = np.ascontiguousarray(pinned, np.int32).ravel()
        cn = np.ascontiguousarray(ctf_node, np.int32).ravel()
        cf = np.ascontiguousarray(ctf_force, np.float64).ravel()
        if cf.size == 0:
            cf = np.zeros(1)
        if pins.size == 0:
            pins = np.zeros(1, np.int32)
            npin = 0
        else:
            npin = pins.size
        if cn.size == 0:
            cn = np.zeros(1, np.int32)
            nctf = 0
        else:
            nctf = cn.size
        drag = None if stokes_drag is None else np.ascontiguousarray(stokes_drag, np.float64)
        self._drag = drag
        p = OrcParams(dt, kT, epsilon2, int(max_iterations_cg), int(calc_noise), int(calc_stokes))
        st = np.ascontiguousarray(rng_state if rng_state is not None else [12345] * 6, np.uint32)
        h = L.orc_blob_create(pos.shape[0], ne, pos, top, *self._keep, npin, pins, nctf, cn, cf,
                              None if drag is None else drag.ctypes.data, C.byref(p), st)
        if not h:
            raise ValueError("oracle blob creation failed")
        self.h = h
        self.nt = pos.shape[0]
        self.ne = ne
        self._top = top
        if surface_forces is not None:
            fn, sf, sp, sfc = surface_forces
            fn = np.ascontiguousarray(fn, np.int32).reshape(-1)
            sf = np.ascontiguousarray(sf, np.float64).reshape(-1)
            sp = np.ascontiguousarray(sp, np.int32)
            sfc = np.ascontiguousarray(sfc, np.int32)
            if L.orc_blob_set_surface_forces(self.h, fn.size // 3, fn if fn.size else np.zeros(3, np.int32),
                                             sp.size - 1, sf if sf.size else np.zeros(3), sp,
                                             sfc if sfc.size else np.zeros(1, np.int32)) != 0:
                raise ValueError("oracle surface forces rejected")
        if rotational_forces is not None:
            rn, rf, ra, rt = rotational_forces
            rn = np.ascontiguousarray(rn, np.int32).ravel()
            rf = np.ascontiguousarray(rf, np.float64).ravel()
            ra = np.ascontiguousarray(ra, np.float64).reshape(-1)
            tq = np.array([1 if str(t)[1:2] == "t" else (0 if str(t)[1:2] == "f" else 2) for t in rt], np.int32)
            if rn.size and (rf.size != rn.size or ra.size != 6 * rn.size or tq.size != rn.size
                            or any(str(t)[:1] != "p" for t in rt)):
                raise ValueError("oracle rotational forces rejected")
            if rn.size and L.orc_blob_set_rotational_forces(self.h, rn.size, rn, rf, ra, tq) != 0:
                raise ValueError("oracle rotational forces rejected")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_blob_destroy(self.h)
            self.h = None

    def step(self):
        return lib().orc_blob_step(self.h)

    def set_mass_lumped(self, rho):
        """MassLumpedSolver in place of the mass-less CG (rho per element, meso units)"""
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, np.float64), (self.ne,)))
        if lib().orc_blob_set_mass_lumped(self.h, r) != 0:
            raise ValueError("oracle densities rejected")

    def set_mass_consistent(self, rho):
        """ConjugateGradientSolver (consistent mass, CG each step) in place of the mass-less CG"""
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, np.float64), (self.ne,)))
        if lib().orc_blob_set_mass_consistent(self.h, r) != 0:
            raise ValueError("oracle densities rejected")

    def kinetic_energy(self):
        return lib().orc_blob_kinetic_energy(self.h)

    def add_node_forces(self, nodes, forces):
        """Blob::add_force_to_node for the next step (meso units, list order)"""
        n = np.ascontiguousarray(nodes, np.int32).ravel()
        f = np.ascontiguousarray(forces, np.float64).reshape(-1)
        if f.size != 3 * n.size or lib().orc_blob_add_node_forces(self.h, n.size, n, f) != 0:
            raise ValueError("oracle node forces rejected")

    def update_internal_forces(self):
        return lib().orc_blob_update_internal_forces(self.h)

    def update_positions(self):
        return lib().orc_blob_update_positions(self.h)

    def _get(self, fn, shape):
        out = np.zeros(shape, np.float64)
        getattr(lib(), fn)(self.h, out.reshape(-1))
        return out

    def positions(self):
        return self._get("orc_blob_get_positions", (self.nt, 3))

    def set_positions(self, p):
        lib().orc_blob_set_positions(self.h, np.ascontiguousarray(p, np.float64).reshape(-1))

    def set_velocities(self, v):
        lib().orc_blob_set_velocities(self.h, np.ascontiguousarray(v, np.float64).reshape(-1))

    def velocities(self):
        return self._get("orc_blob_get_velocities", (self.nt, 3))

    def rhs(self):
        return self._get("orc_blob_get_rhs", (self.nt, 3))

    def element_forces(self):
        return self._get("orc_blob_get_element_forces", (self.ne, 4, 3))

    def cg_iterations(self):
        return lib().orc_blob_last_cg_iterations(self.h)

    def rng_state(self):
        out = np.zeros(6, np.uint32)
        lib().orc_blob_get_rng_state(self.h, out)
        return out

    def set_rng_state(self, s):
        lib().orc_blob_set_rng_state(self.h, np.ascontiguousarray(s, np.uint32))

    def strain_energy(self):
        return lib().orc_blob_strain_energy(self.h)

    def rest_state(self):
        j = np.zeros((self.ne, 9))
        v = np.zeros(self.ne)
        lib().orc_blob_get_rest_state(self.h, j.reshape(-1), v)
        return j, v

    def build_matrix(self):
        lib().orc_blob_build_matrix(self.h)

    def matrix(self):
        nnz = lib().orc_blob_matrix_nnz(self.h)
        rp = np.zeros(3 * self.nt + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz, np.float64)
        lib().orc_blob_get_matrix(self.h, rp, col, val)
        return rp, col, val

    def inverse_diagonal(self):
        return self._get("orc_blob_get_inverse_diagonal", (3 * self.nt,))

    def apply_matrix(self, x):
        out = np.zeros(3 * self.nt)
        lib().orc_blob_apply_matrix(self.h, np.ascontiguousarray(x, np.float64).reshape(-1), out)
        return out.reshape(-1, 3)

    def element_matrices(self):
        return self._get("orc_blob_get_element_matrices", (self.ne, 12, 12))

    def pinned_flags(self):
        out = np.zeros(self.nt, np.int32)
        lib().orc_blob_get_pinned_flags(self.h, out)
        return out

    def time_step(self, cg_iters):
        """(t_element, t_gather_and_build, t_per_cg_iteration) in seconds for one step"""
        te, tb, ti = C.c_double(), C.c_double(), C.c_double()
        lib().orc_blob_time_step(self.h, int(cg_iters), C.byref(te), C.byref(tb), C.byref(ti))
        return te.value, tb.value, ti.value

    def solve(self, b):
        x = np.ascontiguousarray(b, np.float64).reshape(-1).copy()
        it = lib().orc_blob_solve(self.h, x)
        return x.reshape(-1, 3), it


def ref_matrix(blob: "Blob", x=None):
    """The reference's own fixed-pattern matrix (oracle/_ref) for the oracle blob's
    current element matrices: (rowptr, col, val, inv_diag, y=V x or None)."""
    R = ref_lib()
    if R is None:
        return None
    V = np.ascontiguousarray(blob.element_matrices())
    top = np.ascontiguousarray(blob._top, np.int32)
    pin = np.ascontiguousarray(blob.pinned_flags(), np.int32)
    drag = blob._drag
    dp = None if drag is None else drag.ctypes.data
    nnz = R.ref_nomass_matrix(blob.nt, blob.ne, top, pin, V, dp, None, None, None, None, None, None)
    rp = np.zeros(3 * blob.nt + 1, np.int32)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz, np.float64)
    inv = np.zeros(3 * blob.nt, np.float64)
    y = None
    xp = yp = None
    if x is not None:
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        y = np.zeros(3 * blob.nt, np.float64)
        xp, yp = x.ctypes.data, y.ctypes.data
    R.ref_nomass_matrix(blob.nt, blob.ne, top, pin, V, dp, rp.ctypes.data, col.ctypes.data, val.ctypes.data,
                        inv.ctypes.data, xp, yp)
    return rp, col, val, inv, (None if y is None else y.reshape(-1, 3))
