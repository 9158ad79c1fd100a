"""ctypes binding of libffea_gpu.so (the C ABI declared in include/ffea_gpu.h).

The library is loaded from this package directory; there is no fallback: a
missing library or HIP device raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libffea_gpu.so")

OK, ERR_INVERSION, ERR_CG_NO_CONVERGE, ERR_INVALID, ERR_HIP, ERR_NO_DEVICE, ERR_IO = range(7)
FP64, MIXED = 0, 1
WORLD_NO_DETAIL = 1  # ffea_world_create_ex: no detailed measurement file (FFEA_WORLD_NO_DETAIL)
NUM_PHASES = 6
PHASES = ("element", "assemble", "gather_cg_start", "cg_spmv", "cg_vector", "update")

EXPORTED = [
    "ffea_gpu_blob_create", "ffea_gpu_blob_destroy", "ffea_gpu_blob_update_internal_forces",
    "ffea_gpu_blob_update_positions", "ffea_gpu_blob_step", "ffea_gpu_blob_get_positions",
    "ffea_gpu_blob_set_positions", "ffea_gpu_blob_get_velocities", "ffea_gpu_blob_get_rhs",
    "ffea_gpu_blob_get_element_forces", "ffea_gpu_blob_get_cg_iterations", "ffea_gpu_blob_get_rng_state",
    "ffea_gpu_blob_set_rng_state", "ffea_gpu_blob_get_info", "ffea_gpu_blob_set_timing",
    "ffea_gpu_blob_get_timing", "ffea_gpu_copy_probe", "ffea_gpu_blob_synchronize", "ffea_gpu_blob_last_error",
    "ffea_gpu_last_error", "ffea_gpu_batch_create", "ffea_gpu_batch_get_positions", "ffea_gpu_batch_set_positions",
    "ffea_gpu_batch_get_velocities", "ffea_gpu_batch_get_rhs", "ffea_gpu_batch_get_element_forces",
    "ffea_gpu_batch_get_cg_iterations", "ffea_gpu_batch_get_rng_state", "ffea_gpu_batch_set_rng_stream",
    "ffea_gpu_rng_stream_state", "ffea_gpu_advance_rng_state", "ffea_gpu_blob_set_row_blocks",
    "ffea_gpu_blob_autotune", "ffea_gpu_blob_use_graph", "ffea_gpu_check_layout", "ffea_gpu_blob_set_material",
    "ffea_gpu_blob_run_checkpointed", "ffea_gpu_blob_adjoint", "ffea_gpu_batch_strain_energy",
    "ffea_world_create", "ffea_world_create_ex", "ffea_world_run", "ffea_world_current_step", "ffea_world_blob", "ffea_world_destroy",
    "ffea_world_last_error", "ffea_gpu_blob_set_preconditioner", "ffea_gpu_blob_get_multigrid_info",
    "ffea_gpu_multigrid_inspect", "ffea_gpu_blob_create_ex", "ffea_gpu_blob_add_node_forces",
    "ffea_gpu_blob_get_node_positions", "ffea_gpu_blob_set_solver", "ffea_gpu_blob_set_velocities",
    "ffea_gpu_blob_kinetic_energy", "ffea_gpu_solver_create", "ffea_gpu_solver_solve", "ffea_gpu_solver_apply_matrix",
    "ffea_gpu_blob_set_refinement", "ffea_gpu_blob_get_refinement", "ffea_gpu_batch_strain_energy_noalloc",
    "ffea_gpu_blob_get_launch_shape", "ffea_gpu_blob_set_launch_shape", "ffea_gpu_read_probe",
]

PRECOND_JACOBI, PRECOND_MULTIGRID = 0, 1
SOLVER_ITERATIVE, SOLVER_MASSLUMPED, SOLVER_NOMASS_CG = 1, 2, 3
MG_MAX_LEVELS = 8


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("kT", C.c_double), ("epsilon2", C.c_double),
                ("max_iterations_cg", C.c_int), ("calc_noise", C.c_int), ("calc_stokes", C.c_int),
                ("precision", C.c_int), ("device", C.c_int)]


class Info(C.Structure):
    _fields_ = [("num_nodes", C.c_int), ("num_elements", C.c_int), ("num_linear_nodes", C.c_int),
                ("num_blocks", C.c_longlong), ("num_block_slots", C.c_longlong), ("precision", C.c_int),
                ("device_bytes", C.c_longlong), ("num_blobs", C.c_int), ("num_chunks", C.c_int),
                ("row_blocks", C.c_int), ("graph_iterations", C.c_int),
                ("single_launch_cg", C.c_int)]


class SurfaceForces(C.Structure):
    _fields_ = [("num_faces", C.c_int), ("face_nodes", C.c_void_p), ("num_sets", C.c_int),
                ("set_force", C.c_void_p), ("set_ptr", C.c_void_p), ("set_faces", C.c_void_p)]


class RotationalForces(C.Structure):
    _fields_ = [("num", C.c_int), ("node", C.c_void_p), ("force", C.c_void_p), ("axis", C.c_void_p),
                ("type", C.c_void_p)]


class MultigridInfo(C.Structure):
    _fields_ = [("num_levels", C.c_int), ("rows", C.c_int * MG_MAX_LEVELS), ("block", C.c_int * MG_MAX_LEVELS),
                ("blocks", C.c_longlong * MG_MAX_LEVELS), ("transfer_blocks", C.c_longlong * MG_MAX_LEVELS),
                ("omega", C.c_double * MG_MAX_LEVELS), ("lmax", C.c_double * MG_MAX_LEVELS),
                ("coarsest_unknowns", C.c_int), ("setup_seconds", C.c_double)]

    def as_dict(self):
        L = self.num_levels
        return {"num_levels": L, "rows": list(self.rows[:L]), "block": list(self.block[:L]),
                "blocks": list(self.blocks[:L]), "transfer_blocks": list(self.transfer_blocks[:max(0, L - 1)]),
                "omega": list(self.omega[:L]), "lmax": list(self.lmax[:L]),
                "coarsest_unknowns": self.coarsest_unknowns, "setup_seconds": self.setup_seconds}


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                           "or python -m ffea_ci_demo_amd.build")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.ffea_gpu_blob_create.argtypes = [C.POINTER(Params), C.c_int, C.c_int, _dp, _ip, _dp, _dp, _dp, _dp,
                                       C.c_int, vp, C.c_int, vp, vp, vp, _up, C.POINTER(vp)]
    L.ffea_gpu_blob_create_ex.argtypes = [C.POINTER(Params), C.c_int, C.c_int, _dp, _ip, _dp, _dp, _dp, _dp,
                                          C.c_int, vp, C.c_int, vp, vp, vp, _up, C.POINTER(SurfaceForces),
                                          C.POINTER(RotationalForces), C.c_int, vp, C.POINTER(vp)]
    L.ffea_gpu_solver_create.argtypes = [C.POINTER(Params), C.c_int, C.c_int, _dp, _ip, _dp, _dp, C.c_int, vp, vp,
                                         C.POINTER(vp)]
    L.ffea_gpu_solver_solve.argtypes = [vp, _dp, _dp]
    L.ffea_gpu_solver_apply_matrix.argtypes = [vp, _dp, _dp, _dp]
    L.ffea_gpu_blob_set_refinement.argtypes = [vp, C.c_double]
    L.ffea_gpu_blob_get_refinement.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
    L.ffea_gpu_blob_add_node_forces.argtypes = [vp, C.c_int, _ip, _dp]
    L.ffea_gpu_blob_get_node_positions.argtypes = [vp, C.c_int, _ip, _dp]
    L.ffea_gpu_blob_set_solver.argtypes = [vp, C.c_int, vp]
    L.ffea_gpu_blob_set_velocities.argtypes = [vp, _dp]
    L.ffea_gpu_blob_kinetic_energy.argtypes = [vp, C.POINTER(C.c_double)]
    L.ffea_gpu_blob_destroy.argtypes = [vp]
    L.ffea_gpu_blob_destroy.restype = None
    for f in ("update_internal_forces", "update_positions", "synchronize"):
        getattr(L, "ffea_gpu_blob_" + f).argtypes = [vp]
    L.ffea_gpu_blob_step.argtypes = [vp, C.c_int]
    for f in ("get_positions", "set_positions", "get_velocities", "get_rhs", "get_element_forces"):
        getattr(L, "ffea_gpu_blob_" + f).argtypes = [vp, _dp]
    L.ffea_gpu_blob_get_cg_iterations.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
    L.ffea_gpu_blob_get_rng_state.argtypes = [vp, _up]
    L.ffea_gpu_blob_set_rng_state.argtypes = [vp, _up]
    L.ffea_gpu_blob_get_info.argtypes = [vp, C.POINTER(Info)]
    L.ffea_gpu_blob_set_timing.argtypes = [vp, C.c_uint]
    L.ffea_gpu_copy_probe.argtypes = [C.c_int, C.c_longlong, C.c_int, C.POINTER(C.c_double)]
    L.ffea_gpu_read_probe.argtypes = [C.c_int, C.c_longlong, C.c_int, C.POINTER(C.c_double)]
    L.ffea_gpu_blob_get_timing.argtypes = [vp, _dp, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
    pp = C.POINTER(C.c_void_p)
    L.ffea_gpu_batch_create.argtypes = [C.POINTER(Params), C.c_int, _ip, _ip, pp, pp, pp, pp, pp, pp, vp, vp, vp, vp,
                                        vp, vp, _up, C.POINTER(vp)]
    L.ffea_gpu_check_layout.argtypes = [C.POINTER(Params), C.c_int, _ip, _ip, pp, pp, pp, pp, pp, pp, vp, vp, vp, vp,
                                        vp, vp, _up, C.POINTER(Info)]
    for f in ("get_positions", "set_positions", "get_velocities", "get_rhs", "get_element_forces"):
        getattr(L, "ffea_gpu_batch_" + f).argtypes = [vp, C.c_int, _dp]
    L.ffea_gpu_batch_get_cg_iterations.argtypes = [vp, _ip, np.ctypeslib.ndpointer(dtype=np.int64,
                                                                                     flags="C_CONTIGUOUS")]
    L.ffea_gpu_batch_get_rng_state.argtypes = [vp, C.c_int, _up]
    L.ffea_gpu_batch_set_rng_stream.argtypes = [vp, C.c_int, _up, C.c_longlong]
    L.ffea_gpu_blob_set_rng_stream.argtypes = [vp, _up, C.c_longlong]
    L.ffea_gpu_rng_stream_state.argtypes = [_up, C.c_longlong, _up]
    L.ffea_gpu_advance_rng_state.argtypes = [_up, C.c_longlong, _up]
    L.ffea_gpu_blob_set_row_blocks.argtypes = [vp, C.c_int]
    L.ffea_gpu_blob_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.ffea_gpu_blob_get_launch_shape.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.ffea_gpu_blob_set_launch_shape.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.ffea_gpu_blob_use_graph.argtypes = [vp, C.c_int, C.c_int]
    L.ffea_gpu_blob_set_material.argtypes = [vp, _dp, _dp, _dp, _dp]
    L.ffea_gpu_blob_run_checkpointed.argtypes = [vp, C.c_int, C.c_int]
    L.ffea_gpu_blob_adjoint.argtypes = [vp, _dp, vp, vp, vp, vp, vp]
    L.ffea_gpu_batch_strain_energy.argtypes = [vp, _dp]
    L.ffea_gpu_batch_strain_energy_noalloc.argtypes = [vp, _dp]
    L.ffea_gpu_blob_set_preconditioner.argtypes = [vp, C.c_int]
    L.ffea_gpu_blob_get_multigrid_info.argtypes = [vp, C.POINTER(MultigridInfo)]
    L.ffea_gpu_multigrid_inspect.argtypes = [C.c_int, _ip, _ip, _dp, _dp, vp, C.c_int, C.POINTER(MultigridInfo),
                                             vp, vp, vp, vp, vp, vp]
    L.ffea_world_create.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(vp)]
    L.ffea_world_create_ex.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
    L.ffea_world_run.argtypes = [vp, C.c_longlong]
    L.ffea_world_current_step.argtypes = [vp]
    L.ffea_world_current_step.restype = C.c_longlong
    L.ffea_world_blob.argtypes = [vp, C.c_int, C.POINTER(vp)]
    L.ffea_world_destroy.argtypes = [vp]
    L.ffea_world_destroy.restype = None
    L.ffea_world_last_error.argtypes = [vp]
    L.ffea_world_last_error.restype = C.c_char_p
    L.ffea_gpu_blob_last_error.argtypes = [vp]
    L.ffea_gpu_blob_last_error.restype = C.c_char_p
    L.ffea_gpu_last_error.argtypes = []
    L.ffea_gpu_last_error.restype = C.c_char_p
    _lib = L
    return L


def multigrid_inspect(rowptr, col, val, pos, pinned=None, free_blob=True, with_matrices=False):
    """Host-only multigrid set-up of a 3x3-block CSR corner-node matrix (no device):
    returns the level info dict and, with with_matrices, the first prolongator P0 and
    coarse operator A1 as (rowptr, col, blocks) triples."""
    L = lib()
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64).reshape(-1)
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1)
    n = rowptr.size - 1
    pin = None if pinned is None else np.ascontiguousarray(pinned, np.uint8)
    pin_p = None if pin is None else pin.ctypes.data
    info = MultigridInfo()
    rc = L.ffea_gpu_multigrid_inspect(n, rowptr, col, val, pos, pin_p, int(bool(free_blob)), C.byref(info),
                                      None, None, None, None, None, None)
    if rc != OK:
        raise RuntimeError(L.ffea_gpu_last_error().decode())
    d = info.as_dict()
    if not with_matrices:
        return d
    npb, n1, nab = d["transfer_blocks"][0], d["rows"][1], d["blocks"][1]
    prp, pcol, pval = np.zeros(n + 1, np.int32), np.zeros(npb, np.int32), np.zeros((npb, 3, 6))
    arp, acol, aval = np.zeros(n1 + 1, np.int32), np.zeros(nab, np.int32), np.zeros((nab, 6, 6))
    rc = L.ffea_gpu_multigrid_inspect(n, rowptr, col, val, pos, pin_p, int(bool(free_blob)), C.byref(info),
                                      prp.ctypes.data, pcol.ctypes.data, pval.ctypes.data, arp.ctypes.data,
                                      acol.ctypes.data, aval.ctypes.data)
    if rc != OK:
        raise RuntimeError(L.ffea_gpu_last_error().decode())
    return d, (prp, pcol, pval), (arp, acol, aval)


def copy_probe(nbytes: int, iters: int = 20, device: int = 0) -> float:
    """ms per copy of nbytes (read nbytes + write nbytes)"""
    ms = C.c_double()
    rc = lib().ffea_gpu_copy_probe(device, int(nbytes), int(iters), C.byref(ms))
    if rc != OK:
        raise RuntimeError(lib().ffea_gpu_last_error().decode())
    return ms.value


def read_probe(nbytes: int, iters: int = 20, device: int = 0) -> float:
    """ms per streaming read of nbytes"""
    ms = C.c_double()
    rc = lib().ffea_gpu_read_probe(device, int(nbytes), int(iters), C.byref(ms))
    if rc != OK:
        raise RuntimeError(lib().ffea_gpu_last_error().decode())
    return ms.value


def rng_stream_state(package_seed, stream: int) -> np.ndarray:
    """initial state of the stream-th RngStream after SetPackageSeed(package_seed)"""
    out = np.zeros(6, np.uint32)
    rc = lib().ffea_gpu_rng_stream_state(np.ascontiguousarray(package_seed, np.uint32), int(stream), out)
    if rc != OK:
        raise ValueError("invalid MRG32k3a package seed or stream index")
    return out


def advance_rng_state(state, ndraws: int) -> np.ndarray:
    """state after ndraws draws (exact jump-ahead)"""
    out = np.zeros(6, np.uint32)
    rc = lib().ffea_gpu_advance_rng_state(np.ascontiguousarray(state, np.uint32), int(ndraws), out)
    if rc != OK:
        raise ValueError("invalid MRG32k3a state or draw count")
    return out
