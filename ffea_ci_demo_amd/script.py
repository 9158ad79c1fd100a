"""Reading FFEA input (.ffea + blob files) through the library's C++ reader
(include/ffea_script.h); no GPU needed. `Script.blob_spec(b)` gives the arrays of
blob b in meso units, `Script.gpu_blob(b)` the GPU blob built from them."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native, units
from .workloads import BlobSpec


class ScriptParams(C.Structure):
    _fields_ = [("num_blobs", C.c_int), ("num_steps", C.c_longlong), ("check", C.c_int), ("rng_seed", C.c_int),
                ("restart", C.c_int), ("calc_noise", C.c_int), ("calc_stokes", C.c_int),
                ("calc_ctforces", C.c_int), ("calc_ssint", C.c_int), ("calc_steric", C.c_int),
                ("calc_es", C.c_int), ("calc_kinetics", C.c_int), ("calc_springs", C.c_int),
                ("calc_preComp", C.c_int), ("move_into_box", C.c_int), ("max_iterations_cg", C.c_int),
                ("dt", C.c_double), ("kT", C.c_double), ("epsilon2", C.c_double), ("stokes_visc", C.c_double)]


def _bind(L):
    if getattr(L, "_script_bound", False):
        return L
    vp = C.c_void_p
    L.ffea_script_load.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.ffea_script_free.argtypes = [vp]
    L.ffea_script_free.restype = None
    L.ffea_script_get_params.argtypes = [vp, C.POINTER(ScriptParams)]
    ip = C.POINTER(C.c_int)
    L.ffea_script_blob_sizes.argtypes = [vp, C.c_int, ip, ip, ip, ip, ip]
    L.ffea_script_blob_arrays.argtypes = [vp, C.c_int] + [vp] * 11
    L.ffea_script_rng_state.argtypes = [vp, np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")]
    L.ffea_script_blob_surface_forces.argtypes = [vp, C.c_int, ip, ip, ip, vp, vp, vp, vp]
    L.ffea_script_blob_rotational_forces.argtypes = [vp, C.c_int, ip, vp, vp, vp, vp]
    L.ffea_script_springs.argtypes = [vp, ip, vp, vp, vp, vp, vp, vp]
    L.ffea_script_blob_solver.argtypes = [vp, C.c_int, ip]
    L.ffea_gpu_blob_create_from_script.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
    L._script_bound = True
    return L


class ScriptError(RuntimeError):
    pass


class Script:
    def __init__(self, path: str):
        L = _bind(_native.lib())
        h = C.c_void_p()
        rc = L.ffea_script_load(path.encode(), C.byref(h))
        if rc != _native.OK:
            raise ScriptError(L.ffea_gpu_last_error().decode())
        self.h = h
        self.path = path

    def close(self):
        if getattr(self, "h", None):
            _native.lib().ffea_script_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> dict:
        p = ScriptParams()
        _native.lib().ffea_script_get_params(self.h, C.byref(p))
        return {f: getattr(p, f) for f, _ in ScriptParams._fields_}

    def rng_state(self):
        out = np.zeros(6, np.uint32)
        _native.lib().ffea_script_rng_state(self.h, out)
        return out

    def blob_arrays(self, b: int = 0) -> dict:
        L = _native.lib()
        nn, ne, npin, nctf, nf = (C.c_int() for _ in range(5))
        rc = L.ffea_script_blob_sizes(self.h, b, C.byref(nn), C.byref(ne), C.byref(npin), C.byref(nctf), C.byref(nf))
        if rc != _native.OK:
            raise ScriptError(f"no blob {b}")
        a = {"pos": np.zeros((nn.value, 3)), "top": np.zeros((ne.value, 10), np.int32),
             "A": np.zeros(ne.value), "B": np.zeros(ne.value), "G": np.zeros(ne.value), "E": np.zeros(ne.value),
             "rho": np.zeros(ne.value), "pinned": np.zeros(npin.value, np.int32),
             "ctf_node": np.zeros(nctf.value, np.int32), "ctf_force": np.zeros((nctf.value, 3)),
             "stokes_radius": np.zeros(nn.value)}
        order = ["pos", "top", "A", "B", "G", "E", "rho", "pinned", "ctf_node", "ctf_force", "stokes_radius"]
        L.ffea_script_blob_arrays(self.h, b, *[a[k].ctypes.data if a[k].size else None for k in order])
        a["num_surface_faces"] = nf.value
        return a

    def surface_forces(self, b: int = 0):
        """blob b's linear surface constant forces as (face_nodes, set_force, set_ptr,
        set_faces), or None"""
        L = _native.lib()
        nf, ns, nent = C.c_int(), C.c_int(), C.c_int()
        if L.ffea_script_blob_surface_forces(self.h, b, C.byref(nf), C.byref(ns), C.byref(nent), None, None, None,
                                             None) != _native.OK:
            raise ScriptError(f"no blob {b}")
        if ns.value <= 0:
            return None
        fn = np.zeros((nf.value, 3), np.int32)
        sf = np.zeros((ns.value, 3))
        sp = np.zeros(ns.value + 1, np.int32)
        sfc = np.zeros(max(1, nent.value), np.int32)
        L.ffea_script_blob_surface_forces(self.h, b, None, None, None, fn.ctypes.data, sf.ctypes.data, sp.ctypes.data,
                                          sfc.ctypes.data)
        return fn, sf, sp, sfc[:nent.value]

    def rotational_forces(self, b: int = 0):
        """blob b's rotational constant forces as (node, force, axis, type), or None"""
        L = _native.lib()
        n = C.c_int()
        if L.ffea_script_blob_rotational_forces(self.h, b, C.byref(n), None, None, None, None) != _native.OK:
            raise ScriptError(f"no blob {b}")
        if n.value <= 0:
            return None
        node = np.zeros(n.value, np.int32)
        force = np.zeros(n.value)
        axis = np.zeros((n.value, 6))
        typ = C.create_string_buffer(2 * n.value)
        L.ffea_script_blob_rotational_forces(self.h, b, None, node.ctypes.data, force.ctypes.data, axis.ctypes.data,
                                             typ)
        raw = typ.raw.decode("ascii")
        return node, force, axis, [raw[2 * i:2 * i + 2] for i in range(n.value)]

    def springs(self):
        """the script's springs, file order: (k, l, blob0, blob1, node0, node1, active), k and l
        in meso units"""
        L = _native.lib()
        n = C.c_int()
        L.ffea_script_springs(self.h, C.byref(n), None, None, None, None, None, None)
        m = n.value
        k, l = np.zeros(max(1, m)), np.zeros(max(1, m))
        blob, conf, node = (np.zeros(max(2, 2 * m), np.int32) for _ in range(3))
        act = np.zeros(max(1, m), np.int32)
        L.ffea_script_springs(self.h, None, k.ctypes.data, l.ctypes.data, blob.ctypes.data, conf.ctypes.data,
                              node.ctypes.data, act.ctypes.data)
        return [(float(k[i]), float(l[i]), int(blob[2 * i]), int(blob[2 * i + 1]), int(node[2 * i]),
                 int(node[2 * i + 1]), bool(act[i])) for i in range(m)]

    def blob_spec(self, b: int = 0) -> BlobSpec:
        a = self.blob_arrays(b)
        p = self.params()
        drag = None
        if p["calc_stokes"]:  # Blob::init (Blob.cpp:264-268)
            drag = units.stokes_drag(a["stokes_radius"], p["stokes_visc"])
        return BlobSpec(pos=a["pos"], top=a["top"], A=a["A"], B=a["B"], G=a["G"], E=a["E"], dt=p["dt"], kT=p["kT"],
                        epsilon2=p["epsilon2"], max_iterations_cg=p["max_iterations_cg"], calc_noise=p["calc_noise"],
                        rng_state=self.rng_state(), pinned=a["pinned"], ctf_node=a["ctf_node"],
                        ctf_force=a["ctf_force"], calc_stokes=p["calc_stokes"], stokes_drag=drag, name=self.path,
                        surface_forces=self.surface_forces(b), rotational_forces=self.rotational_forces(b),
                        external_nodes=tuple(n for sp in self.springs() if sp[6]
                                             for bb, n in ((sp[2], sp[4]), (sp[3], sp[5])) if bb == b),
                        solver=self.solver(b), rho=a["rho"])

    def solver(self, b: int = 0) -> str:
        L = _native.lib()
        k = C.c_int()
        if L.ffea_script_blob_solver(self.h, b, C.byref(k)) != _native.OK:
            raise ScriptError(f"no blob {b}")
        return {_native.SOLVER_MASSLUMPED: "masslumped", _native.SOLVER_ITERATIVE: "cg"}.get(k.value, "nomass")

    def gpu_blob(self, b: int = 0, precision: str = "fp64", device: int = 0):
        from .blob import FFEAException, GpuBlob
        L = _native.lib()
        h = C.c_void_p()
        prec = {"fp64": _native.FP64, "mixed": _native.MIXED}[precision]
        rc = L.ffea_gpu_blob_create_from_script(self.h, b, prec, device, C.byref(h))
        if rc != _native.OK:
            raise FFEAException(rc, L.ffea_gpu_last_error().decode())
        g = GpuBlob.__new__(GpuBlob)
        g.h = h
        a = self.blob_arrays(b)
        g.nt = a["pos"].shape[0]
        g.ne = a["top"].shape[0]
        g.precision = precision
        return g
