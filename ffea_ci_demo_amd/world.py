"""The World loop of an FFEA script on the GPU blob path (include/ffea_world.h).

`World(path)` is World::init (World.cpp:155-930) for the blob path: every blob of the
script on the GPU, the trajectory, measurement, detailed measurement (.fdm; not with
`detailed=False`, the reference's -d) and checkpoint files opened with the reference's
headers, or, with `restart = 1`, the noise stream read back from
`checkpoint_in` and the run resumed from the last complete trajectory frame.
`World.run()` is World::run (World.cpp:1747-2060). `python -m ffea_ci_demo_amd.world
script.ffea` runs a script as the reference's `ffea script.ffea` does.
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys

from . import _native
from .blob import FFEAException, GpuBlob


class _BlobView(GpuBlob):
    """a blob owned by a World: closing the view leaves the handle to the world"""

    def __init__(self, h, nt, ne, precision, owner):
        self.h = h
        self.nt = nt
        self.ne = ne
        self.precision = precision
        self._owner = owner

    def close(self):
        self.h = None


class World:
    def __init__(self, script_path: str, precision: str = "fp64", device: int = 0, detailed: bool = True,
                 delete_frames: int = 0):
        L = _native.lib()
        prec = {"fp64": _native.FP64, "mixed": _native.MIXED}[precision]
        h = C.c_void_p()
        flags = 0 if detailed else _native.WORLD_NO_DETAIL
        rc = L.ffea_world_create_ex(script_path.encode(), prec, int(device), flags, int(delete_frames), C.byref(h))
        if rc != _native.OK:
            raise FFEAException(rc, L.ffea_gpu_last_error().decode())
        self.h = h
        self.precision = precision
        self.path = script_path

    def close(self):
        if getattr(self, "h", None):
            _native.lib().ffea_world_destroy(self.h)
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

    def run(self, max_steps: int = -1):
        """steps from the current step to num_steps (inclusive), or at most max_steps of them"""
        rc = _native.lib().ffea_world_run(self.h, int(max_steps))
        if rc != _native.OK:
            raise FFEAException(rc, _native.lib().ffea_world_last_error(self.h).decode())

    @property
    def step(self) -> int:
        return int(_native.lib().ffea_world_current_step(self.h))

    def blob(self, index: int) -> GpuBlob:
        b = C.c_void_p()
        rc = _native.lib().ffea_world_blob(self.h, int(index), C.byref(b))
        if rc != _native.OK:
            raise IndexError(f"no blob {index}")
        v = _BlobView(b, 0, 0, self.precision, self)
        info = v.info()
        v.nt, v.ne = info["num_nodes"], info["num_elements"]
        return v


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="run an FFEA script's blobs on the GPU")
    ap.add_argument("script")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("-d", "--no-detail", action="store_true",
                    help="write no detailed measurement file (.fdm), as the reference's ffea -d")
    ap.add_argument("-l", "--delete-frames", type=int, default=0,
                    help="when restarting, drop this many final frames first, as the reference's ffea -l")
    a = ap.parse_args(argv)
    try:
        with World(a.script, a.precision, a.device, detailed=not a.no_detail, delete_frames=a.delete_frames) as w:
            w.run()
    except FFEAException as e:
        print(f"FFEA: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
