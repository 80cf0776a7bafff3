"""TEST INFRASTRUCTURE: ctypes driver of the CPU oracle (oracle/liboracle.so).

The oracle restates the reference's FLIP substep on the CPU (see
flip_oracle.cpp).  Only tests/, the smoke check and bench.py's CPU-baseline leg
use it, as the checker; it is never on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from splashtank_amd.capi import StDims, StError, StOpts, StParams, _Base

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_PROTOS = {
    "orc_last_error": (C.c_char_p, []),
    "orc_create": (C.c_int, [C.POINTER(StOpts), C.c_int, C.POINTER(C.c_void_p)]),
    "orc_destroy": (C.c_int, [C.c_void_p]),
    "orc_grow": (C.c_int, [C.c_void_p, C.c_uint32]),
    "orc_get_dims": (C.c_int, [C.c_void_p, C.POINTER(StDims)]),
    "orc_upload_particles": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "orc_set_count": (C.c_int, [C.c_void_p, C.c_uint32]),
    "orc_set_colliders": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t, C.c_uint64]),
    "orc_begin_step": (C.c_int, [C.c_void_p, C.POINTER(StParams)]),
    "orc_stage": (C.c_int, [C.c_void_p, C.c_int32]),
    "orc_step": (C.c_int, [C.c_void_p, C.POINTER(StParams)]),
    "orc_read_ctl": (C.c_int, [C.c_void_p, C.c_void_p]),
    "orc_dump_size": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_size_t)]),
    "orc_dump": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t]),
    "orc_threads": (C.c_int, []),
    "orc_set_spray_cap": (C.c_int, [C.c_void_p, C.c_uint32]),
    "orc_set_drips": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "orc_spray_put": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    "orc_read_air": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
    "orc_sincosf": (None, [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "orc_read_spray": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "orc_read_deep": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "orc_expf": (C.c_float, [C.c_float]),
    "orc_ocean_eval": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "orc_vnz": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_int]),
    "orc_hsh3": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_int]),
    "orc_poke": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t]),
    "orc_kernel": (C.c_int, [C.c_void_p, C.c_int32]),
}

# test hook: single narrow-band kernels of the oracle (Oracle.kernel)
KERNELS = {"phiInitAll": 0, "nbPhi": 1, "nbJacA": 2, "nbJacB": 3, "nbVel": 4, "saveVel": 5, "nbKill": 6,
           "reseed": 7, "narrow band": 8, "extrapolate": 9}

_lib = None


def sincosf(x: float):
    """the shared f32 sin/cos definition (include/st_math.h)"""
    s, c = C.c_float(0), C.c_float(0)
    load().orc_sincosf(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def load(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(path):
        raise RuntimeError(f"oracle not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in _PROTOS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class Oracle(_Base):
    """CPU restatement of one FlipSim; fp64=True runs the same algorithm in f64."""

    prefix = "orc"

    def __init__(self, opts: StOpts, fp64: bool = False):
        self.lib = load()
        self.fdtype = np.float64 if fp64 else np.float32
        h = C.c_void_p()
        rc = self.lib.orc_create(C.byref(opts), 1 if fp64 else 0, C.byref(h))
        if rc != 0:
            raise StError(f"orc_create -> {rc}: {self._err()}")
        self.h = h
        self.dims = StDims()
        self._call("get_dims", self.h, C.byref(self.dims))

    def _err(self):
        return self.lib.orc_last_error().decode()

    def step_n(self, prm, n):
        for _ in range(n):
            self.step(prm)

    def grow(self, slots):
        """grow (sim.js:137-152): more slots, the old ones unchanged, the new ones on the free stack."""
        self._call("grow", self.h, int(slots))
        self._call("get_dims", self.h, C.byref(self.dims))

    def read_particles(self, n=None):
        return self.dump("pos")

    def set_spray_cap(self, n):
        self._call("set_spray_cap", self.h, int(n))

    def poke(self, what, a):
        """test hook: overwrite a per-cell field in the canonical order of dump()"""
        from splashtank_amd.capi import DUMP, DUMP_KIND
        kind, _ = DUMP_KIND[what]
        dt = self.fdtype if kind == "f" else np.dtype(kind)
        a = np.ascontiguousarray(a, dtype=dt)
        self._call("poke", self.h, DUMP[what], a.ctypes.data, a.nbytes)

    def kernel(self, name):
        """test hook: run one kernel (KERNELS: the narrow band's, or the opt-in
        extrapolation with the last step's settings) on the current state"""
        self._call("kernel", self.h, KERNELS[name])

    def set_drips(self, samples):
        a = np.ascontiguousarray(samples, np.float32).reshape(-1, 8)
        self._call("set_drips", self.h, a.ctypes.data if a.size else None, a.shape[0])

    def spray_put(self, p4, v4, handed=None):
        """test hook: whitewater points into slots 0..n-1, ring counter = handed (default n)"""
        p4 = np.ascontiguousarray(p4, np.float64).reshape(-1, 4)
        v4 = np.ascontiguousarray(v4, np.float64).reshape(-1, 4)
        n = p4.shape[0]
        self._call("spray_put", self.h, p4.ctypes.data, v4.ctypes.data, n, n if handed is None else int(handed))

    def read_air(self, room=1 << 20):
        """the air field, (cells, 4) and the cell size"""
        a = np.zeros((room, 4), np.float64)
        n = C.c_uint32(0)
        cell = C.c_double(0)
        self._call("read_air", self.h, a.ctypes.data, room, C.byref(n), C.byref(cell))
        return a[:n.value], cell.value

    def read_deep(self, room=200000):
        """Deep water points (x, y, z, 4096 + speed) in active-cell order, as readDeep."""
        p = np.zeros((max(room, 1), 4), np.float64)
        n = C.c_uint32(0)
        self._call("read_deep", self.h, p.ctypes.data, room, C.byref(n))
        return p[:n.value].astype(self.fdtype)

    def read_spray(self, room=1 << 22):
        """(positions with life in w, velocities with kind in w) of the slots handed out, f64 arrays."""
        p = np.zeros((room, 4), np.float64)
        v = np.zeros((room, 4), np.float64)
        n = C.c_uint32(0)
        self._call("read_spray", self.h, p.ctypes.data, v.ctypes.data, room, C.byref(n))
        return p[:n.value].astype(self.fdtype), v[:n.value].astype(self.fdtype)

    def ocean_eval(self, pts):
        """test hook: the ocean helpers of the begun step at points (n, 3) ->
        (n, 5): oceanEta(x, z), oceanVel (3), layerW"""
        p = np.zeros((len(pts), 4), np.float64)
        p[:, :3] = np.asarray(pts, np.float64).reshape(-1, 3)
        out = np.zeros((len(pts), 5), np.float64)
        self._call("ocean_eval", self.h, p.ctypes.data, len(pts), out.ctypes.data)
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def threads() -> int:
    return int(load().orc_threads())


def expf(x: float) -> float:
    return float(load().orc_expf(float(x)))


def vnz(x, y, z, fp64=False) -> float:
    """the oracle's gust noise (wgsl-solver.js:202-207)"""
    return float(load().orc_vnz(float(x), float(y), float(z), 1 if fp64 else 0))


def hsh3(x, y, z, fp64=False) -> float:
    """the oracle's lattice hash of the gust noise (wgsl-solver.js:198-201)"""
    return float(load().orc_hsh3(float(x), float(y), float(z), 1 if fp64 else 0))
