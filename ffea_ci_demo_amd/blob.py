"""Python mirror of the reference's per-blob step interface over the C ABI.

`GpuBlob` exposes the pair World::run calls on a blob every step —
update_internal_forces (Blob.cpp:400-491) and update_positions (Blob.cpp:493-508) —
plus node state access, the noise stream state and the solver's right-hand side.
Failures raise `FFEAException` with the reference's messages (inversion,
CG non-convergence).
"""
from __future__ import annotations

import ctypes as C

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


class GpuBlob:
    def __init__(self, pos, top, A, B, G, E, *, dt, kT, epsilon2, max_iterations_cg=1000, calc_noise=1,
                 calc_stokes=0, pinned=(), ctf_node=(), ctf_force=(), stokes_drag=None, rng_state=None,
                 precision="fp64", device=0):
        L = _native.lib()
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        top = np.ascontiguousarray(top, np.int32).reshape(-1, 10)
        ne = top.shape[0]

        def per_elem(v):
            return np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (ne,)))

        mats = [per_elem(A), per_elem(B), per_elem(G), per_elem(E)]
        prec = {"fp64": _native.FP64, "mixed": _native.MIXED}[precision]
        p = _native.Params(float(dt), float(kT), float(epsilon2), int(max_iterations_cg), int(calc_noise),
                           int(calc_stokes), prec, int(device))
        pin_p, pins = _opt(pinned, np.int32)
        cn_p, cn = _opt(ctf_node, np.int32)
        cf_p, cf = _opt(ctf_force, np.float64)
        drag_p, drag = _opt(stokes_drag, np.float64)
        npin = 0 if pins is None else pins.size
        nctf = 0 if cn is None else cn.size
        st = np.ascontiguousarray(rng_state if rng_state is not None else [12345] * 6, np.uint32)
        h = C.c_void_p()
        rc = L.ffea_gpu_blob_create(C.byref(p), pos.shape[0], ne, pos, top, *mats, npin, pin_p, nctf, cn_p, cf_p,
                                    drag_p, st, C.byref(h))
        if rc != _native.OK:
            raise FFEAException(rc, L.ffea_gpu_last_error().decode())
        self.h = h
        self.nt = pos.shape[0]
        self.ne = ne
        self.precision = precision

    @classmethod
    def from_spec(cls, spec, precision="fp64", **kw):
        return cls(spec.pos, spec.top, spec.A, spec.B, spec.G, spec.E, dt=spec.dt, kT=spec.kT,
                   epsilon2=spec.epsilon2, max_iterations_cg=spec.max_iterations_cg, calc_noise=spec.calc_noise,
                   calc_stokes=spec.calc_stokes, pinned=spec.pinned, ctf_node=spec.ctf_node,
                   ctf_force=spec.ctf_force, stokes_drag=spec.stokes_drag, rng_state=spec.rng_state,
                   precision=precision, **kw)

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

    # -- the reference's per-step interface
    def update_internal_forces(self):
        self._check(_native.lib().ffea_gpu_blob_update_internal_forces(self.h))

    def update_positions(self):
        self._check(_native.lib().ffea_gpu_blob_update_positions(self.h))

    def step(self, nsteps=1):
        self._check(_native.lib().ffea_gpu_blob_step(self.h, int(nsteps)))

    # -- state
    def _get(self, fn, shape):
        out = np.zeros(shape, np.float64)
        self._check(getattr(_native.lib(), fn)(self.h, out.reshape(-1)))
        return out

    def positions(self):
        return self._get("ffea_gpu_blob_get_positions", (self.nt, 3))

    def set_positions(self, p):
        self._check(_native.lib().ffea_gpu_blob_set_positions(self.h, np.ascontiguousarray(p, np.float64).reshape(-1)))

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

    def info(self):
        i = _native.Info()
        self._check(_native.lib().ffea_gpu_blob_get_info(self.h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in _native.Info._fields_}

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
