"""ctypes binding of the C ABI in include/splashtank.h (libsplashtank_hip.so).

`FlipSimC` mirrors the members of the reference's JS `FlipSim`
(src/flip/sim.js:46-547) that the C ABI backs: construction options, particle
upload, colliders, step / stage, readback, control words and canonical dumps.
Loading fails loudly when the HIP library is not built: there is no fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsplashtank_hip.so")

# Handles still open when the interpreter exits are closed first thing at exit,
# while the HIP runtime is still up, dependents before what they depend on:
# adjoints and snapshots, then batches, then sims.
_LIVE = {kind: weakref.WeakSet() for kind in ("adjoint", "snapshot", "batch", "sim")}


def close_all():
    for kind in ("adjoint", "snapshot", "batch", "sim"):
        for obj in list(_LIVE[kind]):
            try:
                obj.close()
            except Exception:
                pass


atexit.register(close_all)

ST_OK, ST_ERR_ARG, ST_ERR_HIP, ST_ERR_STATE, ST_ERR_NOMEM = 0, -1, -2, -3, -4
ST_OPEN_X0, ST_OPEN_X1, ST_OPEN_Y0, ST_OPEN_Y1, ST_OPEN_Z0, ST_OPEN_Z1 = 1, 2, 4, 8, 16, 32

STAGES = ["sort", "blocks", "clear", "p2g", "narrow band", "classify", "normFaces",
          "divergence", "pressure", "project", "g2p", "spray", "reseed"]
(ST_STAGE_SORT, ST_STAGE_BLOCKS, ST_STAGE_CLEAR, ST_STAGE_P2G, ST_STAGE_NARROWBAND,
 ST_STAGE_CLASSIFY, ST_STAGE_NORMFACES, ST_STAGE_DIVERGENCE, ST_STAGE_PRESSURE,
 ST_STAGE_PROJECT, ST_STAGE_G2P, ST_STAGE_SPRAY, ST_STAGE_RESEED) = range(13)
ST_STAGE_COUNT = 13

DUMP = {"tiles": 0, "cnt": 1, "ctype": 2, "pres": 3, "rhs": 4, "accv": 5, "accw": 6,
        "facev": 7, "faceo": 8, "velg": 9, "pos": 10, "vel": 11, "tmap": 12}
# element type and values per entry (per cell, per face triple or per particle)
DUMP_KIND = {"tiles": ("u4", 1), "cnt": ("i4", 1), "ctype": ("u4", 1), "pres": ("f", 1),
             "rhs": ("f", 1), "accv": ("i4", 3), "accw": ("i4", 3), "facev": ("f", 3),
             "faceo": ("f", 3), "velg": ("f", 3), "pos": ("f", 4), "vel": ("f", 4),
             "tmap": ("u4", 1)}


class StOpts(C.Structure):
    _fields_ = [("d", C.c_double), ("ratio", C.c_double), ("size", C.c_double * 3),
                ("open", C.c_uint32), ("max_n", C.c_uint32), ("slots", C.c_uint32),
                ("device", C.c_int32)]


class StDims(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("d", "h", "ratio", "rest", "x0", "y0", "z0", "x1", "y1", "z1")] + \
               [(n, C.c_int32) for n in ("nx", "ny", "nz", "tx", "ty", "tz")] + \
               [(n, C.c_uint32) for n in ("ntiles", "scap", "scap_max", "max_n", "ncell")]


class StParams(C.Structure):
    _fields_ = [("dt", C.c_float), ("g", C.c_float), ("flip", C.c_float), ("comp", C.c_float),
                ("omega", C.c_float), ("vmax", C.c_float), ("tear", C.c_float), ("iters", C.c_int32),
                ("wind", C.c_float * 3), ("drag", C.c_float), ("foam_tau", C.c_float),
                ("foam_sens", C.c_float), ("gust", C.c_float), ("no_foam", C.c_uint32),
                ("no_cull", C.c_uint32), ("fast_air", C.c_uint32), ("no_solid_cache", C.c_uint32),
                ("nb_on", C.c_uint32), ("nb_band", C.c_float), ("only_tiles", C.c_uint32),
                ("nb_init", C.c_uint32), ("nfill", C.c_uint32), ("nfv", C.c_uint32),
                ("fills", C.c_float * 64), ("sor_mode", C.c_uint32),
                ("spr_on", C.c_uint32), ("spr_rate", C.c_float), ("spr_min", C.c_float), ("spr_life", C.c_float),
                ("spr_mist", C.c_float), ("spr_drag", C.c_float), ("spr_mdrag", C.c_float), ("vs", C.c_float),
                ("ww_foam_life", C.c_float), ("ww_rise", C.c_float), ("ww_crest", C.c_float), ("ww_impact", C.c_float),
                ("air_on", C.c_uint32), ("air_couple", C.c_float), ("air_decay", C.c_float), ("air_blur", C.c_float),
                ("drip_on", C.c_uint32), ("drip_rate", C.c_float), ("drip_dry", C.c_float),
                ("ocean_on", C.c_uint32), ("nwave", C.c_uint32), ("ocean_time", C.c_float),
                ("ocean_level", C.c_float), ("ocean_layer", C.c_float), ("waves", C.c_float * 256),
                ("rk2", C.c_uint32), ("extrap_layers", C.c_uint32)]


ST_OCEAN_MAX_WAVES = 32


class StOcean(C.Structure):
    _fields_ = [("height", C.c_double), ("length", C.c_double), ("dir", C.c_double), ("spread", C.c_double),
                ("seed", C.c_int32)]


def ocean_waves(height, length, dir=0.0, spread=0.0, seed=1, g=9.81, depth=1.0):
    """oceanWaves (src/shared.js:769-791) through st_ocean_waves: the wave list
    as an (n, 8) f32 array (kx, kz, amplitude, angular frequency, phase,
    wavenumber, 0, 0)."""
    lib = load()
    oc = StOcean(float(height), float(length), float(dir), float(spread), int(seed))
    w = np.zeros((ST_OCEAN_MAX_WAVES, 8), np.float32)
    n = C.c_uint32(0)
    rc = lib.st_ocean_waves(C.byref(oc), float(g), float(depth), w.ctypes.data, ST_OCEAN_MAX_WAVES, C.byref(n))
    if rc != 0:
        raise StError(f"st_ocean_waves -> {rc}")
    return w[: n.value].copy()


class StSdfGrid(C.Structure):
    _fields_ = [("dims", C.c_int32 * 3), ("origin", C.c_double * 3), ("cell", C.c_double)]


def build_sdf(verts, tris, cell=0.025, margin=0.16):
    """buildSDF (src/shared.js:98-181) through st_build_sdf: verts (n, 3) doubles,
    tris (m, 3) vertex indices -> dict(dims, origin, cell, data (f32, x fastest))."""
    lib = load()
    v = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    g = StSdfGrid()
    rc = lib.st_build_sdf(v.ctypes.data, v.shape[0], t.ctypes.data, t.shape[0], float(cell), float(margin), C.byref(g), None, 0)
    if rc != 0:
        raise StError(f"st_build_sdf -> {rc}")
    n = g.dims[0] * g.dims[1] * g.dims[2]
    data = np.zeros(n, np.float32)
    rc = lib.st_build_sdf(v.ctypes.data, v.shape[0], t.ctypes.data, t.shape[0], float(cell), float(margin), C.byref(g),
                          data.ctypes.data, n)
    if rc != 0:
        raise StError(f"st_build_sdf -> {rc}")
    return {"dims": list(g.dims), "origin": list(g.origin), "cell": g.cell, "data": data}


class StViewLayout(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("nW", "nS", "nD", "words", "oWp", "oWs", "oWf", "oSp", "oSk", "oDp", "oDs",
                                          "stride", "from_pos")]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class StNorms(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("res_l2", "res_max", "rhs_l2", "div_l2", "div_max",
                                          "err_l2", "err_max", "fluid")]


ST_SOR_TUNED = 3
ST_TUNE_MAX = 8


class StSorTune(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("width", C.c_uint32), ("n", C.c_uint32),
                ("cand_mode", C.c_uint32 * ST_TUNE_MAX), ("cand_width", C.c_uint32 * ST_TUNE_MAX),
                ("cand_ms", C.c_float * ST_TUNE_MAX)]

    def as_dict(self):
        return {"mode": self.mode, "width": self.width,
                "candidates": [{"mode": self.cand_mode[k], "width": self.cand_width[k], "ms": self.cand_ms[k]}
                               for k in range(self.n)]}


def make_opts(d, size=(3.0, 2.6, 3.0), ratio=2.0, open_=0, max_n=1, slots=0, device=0) -> StOpts:
    o = StOpts()
    o.d = float(d)
    o.ratio = float(ratio)
    for k in range(3):
        o.size[k] = float(size[k])
    o.open = int(open_)
    o.max_n = int(max_n)
    o.slots = int(slots)
    o.device = int(device)
    return o


def make_params(**kw) -> StParams:
    """The reference's step() defaults (sim.js:197-199) overridden by keywords."""
    p = StParams()
    p.dt = 1.0 / 60.0
    p.g = 9.81
    p.flip = 0.95
    p.comp = 6.0
    p.omega = 0.0
    p.vmax = 14.0
    p.tear = 0.0
    p.iters = 30
    p.foam_tau = 1.2
    p.foam_sens = 1.0
    p.nb_band = 3.0
    p.ww_crest = 1.0
    p.ww_impact = 0.5
    p.air_blur = 0.35
    for k, v in kw.items():
        if k == "wind":
            for q in range(3):
                p.wind[q] = float(v[q])
        elif k == "fills":
            for q, x in enumerate(v):
                p.fills[q] = float(x)
        elif k == "waves":
            w = np.asarray(v, np.float32).reshape(-1)
            if w.size > 8 * ST_OCEAN_MAX_WAVES:
                raise ValueError("at most 32 waves")
            for q, x in enumerate(w):
                p.waves[q] = float(x)
            p.nwave = w.size // 8
        else:
            setattr(p, k, v)
    return p


PROTOS = {
    "st_last_error": (C.c_char_p, []),
    "st_version": (C.c_char_p, []),
    "st_params_default": (None, [C.POINTER(StParams)]),
    "st_create": (C.c_int, [C.POINTER(StOpts), C.POINTER(C.c_void_p)]),
    "st_destroy": (C.c_int, [C.c_void_p]),
    "st_get_dims": (C.c_int, [C.c_void_p, C.POINTER(StDims)]),
    "st_upload_particles": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "st_set_count": (C.c_int, [C.c_void_p, C.c_uint32]),
    "st_clear": (C.c_int, [C.c_void_p]),
    "st_get_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "st_set_colliders": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t, C.c_uint64]),
    "st_step": (C.c_int, [C.c_void_p, C.POINTER(StParams)]),
    "st_step_n": (C.c_int, [C.c_void_p, C.POINTER(StParams), C.c_int32, C.c_int32]),
    "st_begin_step": (C.c_int, [C.c_void_p, C.POINTER(StParams)]),
    "st_stage": (C.c_int, [C.c_void_p, C.c_int32]),
    "st_time_stages": (C.c_int, [C.c_void_p, C.POINTER(StParams), C.c_void_p]),
    "st_read_particles": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "st_read_ctl": (C.c_int, [C.c_void_p, C.c_void_p]),
    "st_grow": (C.c_int, [C.c_void_p, C.c_uint32]),
    "st_dump_size": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_size_t)]),
    "st_dump": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t]),
    "st_sync": (C.c_int, [C.c_void_p]),
    "st_mem_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "st_sor_path": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "st_solve_norms": (C.c_int, [C.c_void_p, C.POINTER(StNorms)]),
    "st_solve_norms_static": (C.c_int, [C.c_void_p, C.POINTER(StNorms)]),
    "st_set_spray_cap": (C.c_int, [C.c_void_p, C.c_uint32]),
    "st_set_drips": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "st_pack_view": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t, C.POINTER(StViewLayout)]),
    "st_ocean_waves": (C.c_int, [C.POINTER(StOcean), C.c_double, C.c_double, C.c_void_p, C.c_uint32,
                                 C.POINTER(C.c_uint32)]),
    "st_build_sdf": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_double, C.c_double, C.POINTER(StSdfGrid),
                               C.c_void_p, C.c_size_t]),
    "st_read_spray": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "st_read_deep": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "st_read_begin": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "st_read_wait": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "st_read_free": (C.c_int, [C.c_void_p]),
    "st_tune_sor": (C.c_int, [C.c_void_p, C.POINTER(StParams), C.c_int32, C.POINTER(StSorTune)]),
    "st_batch_tune_sor": (C.c_int, [C.c_void_p, C.POINTER(StParams), C.c_uint32, C.c_int32, C.POINTER(StSorTune)]),
    "st_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(C.c_void_p)]),
    "st_batch_destroy": (C.c_int, [C.c_void_p]),
    "st_batch_step": (C.c_int, [C.c_void_p, C.POINTER(StParams), C.c_uint32]),
    "st_batch_step_n": (C.c_int, [C.c_void_p, C.POINTER(StParams), C.c_uint32, C.c_int32, C.c_int32]),
    "st_batch_sync": (C.c_int, [C.c_void_p]),
    "st_batch_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "st_snapshot_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "st_snapshot_restore": (C.c_int, [C.c_void_p, C.c_void_p]),
    "st_snapshot_destroy": (C.c_int, [C.c_void_p]),
    "st_adj_create": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "st_adj_destroy": (C.c_int, [C.c_void_p]),
    "st_adj_record": (C.c_int, [C.c_void_p, C.POINTER(StParams)]),
    "st_adj_seed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "st_adj_backward": (C.c_int, [C.c_void_p]),
    "st_adj_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "st_adj_col_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "st_adj_set_sor_path": (C.c_int, [C.c_void_p, C.c_uint32]),
    "st_batch_sor_grid": (C.c_int, [C.c_void_p, C.c_uint32]),
    "st_batch_sor_path": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "st_adj_sor_path": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library; raises when it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(path):
        raise RuntimeError(f"HIP library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in PROTOS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class StError(RuntimeError):
    pass


def _f32c(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class _Base:
    """Shared driver of the C ABI and of the oracle's mirror of it."""

    prefix = "st"
    fdtype = np.float32

    def _call(self, name, *args):
        rc = getattr(self.lib, f"{self.prefix}_{name}")(*args)
        if rc != 0:
            raise StError(f"{self.prefix}_{name} -> {rc}: {self._err()}")
        return rc

    def _err(self):
        return self.lib.st_last_error().decode()

    # --- particles and colliders ---
    def upload(self, pos4, vel4, first=0):
        pos4 = _f32c(pos4).reshape(-1, 4)
        vel4 = _f32c(vel4).reshape(-1, 4)
        assert pos4.shape == vel4.shape
        self._call("upload_particles", self.h, first, pos4.shape[0], pos4.ctypes.data, vel4.ctypes.data)

    def set_colliders(self, cf32, sdf, sig):
        cf32 = _f32c(cf32).reshape(-1)
        sdf = _f32c(sdf).reshape(-1)
        nc = cf32.size // 32
        self._call("set_colliders", self.h, cf32.ctypes.data if nc else None, nc,
                   sdf.ctypes.data if sdf.size else None, sdf.size, int(sig))
        self._keep = (cf32, sdf)

    # --- stepping ---
    def step(self, prm: StParams):
        self._call("step", self.h, C.byref(prm))

    def begin_step(self, prm: StParams):
        self._call("begin_step", self.h, C.byref(prm))

    def stage(self, k: int):
        self._call("stage", self.h, int(k))

    def stages(self, upto: int):
        """Run stages 0..upto inclusive of a begun step."""
        for k in range(upto + 1):
            self.stage(k)

    # --- readback ---
    def ctl(self):
        a = np.zeros(16, np.uint32)
        self._call("read_ctl", self.h, a.ctypes.data)
        return a

    def dump(self, what: str):
        sz = C.c_size_t(0)
        self._call("dump_size", self.h, DUMP[what], C.byref(sz))
        kind, per = DUMP_KIND[what]
        dt = self.fdtype if kind == "f" else np.dtype(kind)
        n = sz.value // np.dtype(dt).itemsize
        a = np.zeros(max(n, 1), dtype=dt)
        self._call("dump", self.h, DUMP[what], a.ctypes.data, sz.value)
        a = a[:n]
        if per > 1:
            a = a.reshape(-1, per)
        return a


class FlipSimC(_Base):
    """One tank on one GPU through libsplashtank_hip.so."""

    prefix = "st"

    def __init__(self, opts: StOpts):
        self.lib = load()
        h = C.c_void_p()
        self._call_create(opts, h)
        self.h = h
        _LIVE["sim"].add(self)
        self.dims = StDims()
        self._call("get_dims", self.h, C.byref(self.dims))

    def _call_create(self, opts, h):
        rc = self.lib.st_create(C.byref(opts), C.byref(h))
        if rc != 0:
            raise StError(f"st_create -> {rc}: {self.lib.st_last_error().decode()}")

    def step_n(self, prm: StParams, n: int, graph: bool = False):
        """n substeps with the same settings; graph=True replays captured HIP graphs."""
        self._call("step_n", self.h, C.byref(prm), int(n), 1 if graph else 0)

    def time_stages(self, prm: StParams):
        """One substep with HIP events around each stage: {stage name: ms}."""
        ms = np.zeros(ST_STAGE_COUNT, np.float32)
        self._call("time_stages", self.h, C.byref(prm), ms.ctypes.data)
        return {STAGES[k]: float(ms[k]) for k in range(ST_STAGE_COUNT)}

    def count(self):
        cnt = C.c_uint32(0)
        self._call("get_count", self.h, C.byref(cnt))
        return cnt.value

    def read_particles(self, n=None):
        cnt = C.c_uint32(0)
        self._call("get_count", self.h, C.byref(cnt))
        n = cnt.value if n is None else n
        a = np.zeros((max(n, 1), 4), np.float32)
        self._call("read_particles", self.h, a.ctypes.data, n)
        return a[:n]

    def sync(self):
        self._call("sync", self.h)

    def read_particles_async(self, n=None):
        """Start reading the particles (id order) without waiting: returns a
        PendingRead whose wait() gives the (n, 4) array of the state at this call."""
        cnt = C.c_uint32(0)
        self._call("get_count", self.h, C.byref(cnt))
        n = cnt.value if n is None else min(int(n), cnt.value)
        return PendingRead(self.lib, self.h, ST_READ_PARTICLES, n)

    def grow(self, slots):
        self._call("grow", self.h, int(slots))
        self._call("get_dims", self.h, C.byref(self.dims))

    def mem_bytes(self):
        b = C.c_uint64(0)
        self._call("mem_bytes", self.h, C.byref(b))
        return b.value

    def sor_path(self):
        """path of the last pressure solve: 0 per-iteration launches, 1 resident, 2 streaming"""
        v = C.c_uint32(0)
        self._call("sor_path", self.h, C.byref(v))
        return v.value

    def set_spray_cap(self, n):
        """setSprayCap: room for n whitewater points (0 turns the spray stage off)."""
        self._call("set_spray_cap", self.h, int(n))

    def set_drips(self, samples):
        """dripUpdate: drip samples, 8 floats each (scenes.drip_samples gives them)"""
        a = np.ascontiguousarray(samples, np.float32).reshape(-1, 8)
        self._call("set_drips", self.h, a.ctypes.data if a.size else None, a.shape[0])

    def read_deep(self, room=200000):
        """readDeep: deep water points (x, y, z, 4096 + speed) of the narrow band look,
        in active-cell order; empty when the narrow band is off."""
        p = np.zeros((max(room, 1), 4), np.float32)
        n = C.c_uint32(0)
        self._call("read_deep", self.h, p.ctypes.data, room, C.byref(n))
        return p[:n.value]

    def read_spray(self, room=1 << 22):
        """readSpray: (positions with life in w, velocities with kind in w) of the slots handed out."""
        p = np.zeros((room, 4), np.float32)
        v = np.zeros((room, 4), np.float32)
        n = C.c_uint32(0)
        self._call("read_spray", self.h, p.ctypes.data, v.ctypes.data, room, C.byref(n))
        return p[:n.value], v[:n.value]

    def pack_view(self, disp_max=0):
        """the live view packed on the GPU (st_pack_view): (layout dict, uint32 words)"""
        L = StViewLayout()
        self._call("pack_view", self.h, int(disp_max), None, 0, C.byref(L))
        out = np.zeros(max(L.words, 1), np.uint32)
        self._call("pack_view", self.h, int(disp_max), out.ctypes.data, out.size, C.byref(L))
        return L.as_dict(), out[:L.words]

    def solve_norms(self):
        """FP64 norms of the last substep's solve and projection (st_solve_norms), as a dict."""
        n = StNorms()
        self._call("solve_norms", self.h, C.byref(n))
        return {f: float(getattr(n, f)) for f, _ in StNorms._fields_}

    def solve_norms_static(self):
        """The same norms through st_solve_norms_static (no device allocation)."""
        n = StNorms()
        self._call("solve_norms_static", self.h, C.byref(n))
        return {f: float(getattr(n, f)) for f, _ in StNorms._fields_}

    def tune_sor(self, prm: StParams, reps: int = 3):
        """Time the pressure-solve candidates from the current state and keep the
        fastest for steps with sor_mode = ST_SOR_TUNED (st_tune_sor); the state is
        left as it was."""
        t = StSorTune()
        self._call("tune_sor", self.h, C.byref(prm), int(reps), C.byref(t))
        return t.as_dict()

    def close(self):
        # a tank still in a batch is refused; it stays open until the batch closes
        if getattr(self, "h", None) and self.lib.st_destroy(self.h) == 0:
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


ST_READ_PARTICLES, ST_READ_CTL, ST_READ_SPRAY, ST_READ_DEEP = 0, 1, 2, 3


class PendingRead:
    """One asynchronous read (st_read_begin): the copy is enqueued at construction;
    wait() blocks until it has landed and returns the result (any thread)."""

    def __init__(self, lib, sim_h, what, n):
        self.lib, self.what, self.n = lib, what, n
        t = C.c_void_p()
        if lib.st_read_begin(sim_h, what, n, C.byref(t)) != 0:
            raise StError(f"st_read_begin: {lib.st_last_error().decode()}")
        self.t = t

    def wait(self):
        room = self.n if self.what == ST_READ_PARTICLES else 16
        out = np.zeros((max(room, 1), 4) if self.what == ST_READ_PARTICLES else 16,
                       np.float32 if self.what == ST_READ_PARTICLES else np.uint32)
        got = C.c_uint32(0)
        try:
            if self.what not in (ST_READ_PARTICLES, ST_READ_CTL):
                raise StError("PendingRead.wait: particles and control words only")
            if self.lib.st_read_wait(self.t, out.ctypes.data, None, room, C.byref(got)) != 0:
                raise StError(f"st_read_wait: {self.lib.st_last_error().decode()}")
        finally:
            self.close()
        return out[:got.value]

    def close(self):
        if getattr(self, "t", None):
            self.lib.st_read_free(self.t)
            self.t = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FlipBatchC:
    """Many tanks (FlipSimC objects) advanced together, one launch per kernel
    for all of them (st_batch_*).  The tanks stay usable on their own (upload,
    read_particles, dump ...); while batched they share the batch's stream."""

    def __init__(self, tanks):
        self.lib = load()
        self.tanks = list(tanks)
        arr = (C.c_void_p * len(self.tanks))(*[t.h for t in self.tanks])
        h = C.c_void_p()
        rc = self.lib.st_batch_create(arr, len(self.tanks), C.byref(h))
        if rc != 0:
            raise StError(f"st_batch_create -> {rc}: {self.lib.st_last_error().decode()}")
        self.h = h
        _LIVE["batch"].add(self)

    def _call(self, name, *args):
        rc = getattr(self.lib, f"st_batch_{name}")(*args)
        if rc != 0:
            raise StError(f"st_batch_{name} -> {rc}: {self.lib.st_last_error().decode()}")

    def step(self, prm):
        """One substep of every tank: `prm` is one StParams for all or a list, one per tank."""
        if isinstance(prm, StParams):
            self._call("step", self.h, C.byref(prm), 1)
        else:
            arr = (StParams * len(prm))(*prm)
            self._call("step", self.h, arr, len(prm))

    def step_n(self, prm, n: int, graph: bool = False):
        """n batched substeps; graph=True replays captured HIP graphs."""
        if isinstance(prm, StParams):
            self._call("step_n", self.h, C.byref(prm), 1, int(n), 1 if graph else 0)
        else:
            arr = (StParams * len(prm))(*prm)
            self._call("step_n", self.h, arr, len(prm), int(n), 1 if graph else 0)

    def sync(self):
        self._call("sync", self.h)

    def sor_grid(self, workgroups: int):
        """Cap the one-launch batched SOR at `workgroups` resident workgroups
        (0: one launch per iteration)."""
        self._call("sor_grid", self.h, int(workgroups))

    def sor_path(self) -> bool:
        """Whether the last batched substep's SOR ran in the one launch."""
        v = C.c_uint32()
        self._call("sor_path", self.h, C.byref(v))
        return bool(v.value)

    def tune_sor(self, prm, reps: int = 3):
        """Time whole batched substeps at several SOR launch widths and keep the
        fastest (st_batch_tune_sor); the tanks are left as they were."""
        t = StSorTune()
        if isinstance(prm, StParams):
            rc = self.lib.st_batch_tune_sor(self.h, C.byref(prm), 1, int(reps), C.byref(t))
        else:
            arr = (StParams * len(prm))(*prm)
            rc = self.lib.st_batch_tune_sor(self.h, arr, len(prm), int(reps), C.byref(t))
        if rc != 0:
            raise StError(f"st_batch_tune_sor -> {rc}: {self.lib.st_last_error().decode()}")
        return t.as_dict()

    def close(self):
        if getattr(self, "h", None):
            self.lib.st_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SnapshotC:
    """A copy of a sim's state to restart substeps from (st_snapshot_*)."""

    def __init__(self, sim):
        self.lib = load()
        self.sim = sim
        h = C.c_void_p()
        if self.lib.st_snapshot_create(sim.h, C.byref(h)) != 0:
            raise StError(f"st_snapshot_create: {self.lib.st_last_error().decode()}")
        self.h = h
        _LIVE["snapshot"].add(self)

    def restore(self):
        if self.lib.st_snapshot_restore(self.sim.h, self.h) != 0:
            raise StError(f"st_snapshot_restore: {self.lib.st_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.st_snapshot_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AdjointC:
    """Adjoint of the substep for one sim (st_adj_*): record substeps, seed the
    gradient of the state after them, reverse them one at a time, read the
    gradient of the state before.  Gradients are (N, 4) float32 in particle-id
    order (w ignored)."""

    def __init__(self, sim, max_steps):
        self.lib = load()
        self.sim = sim
        h = C.c_void_p()
        if self.lib.st_adj_create(sim.h, int(max_steps), C.byref(h)) != 0:
            raise StError(f"st_adj_create: {self.lib.st_last_error().decode()}")
        self.h = h
        _LIVE["adjoint"].add(self)

    def _call(self, name, *args):
        if getattr(self.lib, f"st_adj_{name}")(self.h, *args) != 0:
            raise StError(f"st_adj_{name}: {self.lib.st_last_error().decode()}")

    def record(self, prm):
        self._call("record", C.byref(prm))

    def seed(self, g_pos, g_vel):
        n = self.sim.count()
        gp = np.zeros((n, 4), np.float32)
        gv = np.zeros((n, 4), np.float32)
        gp[:, :3] = np.asarray(g_pos, np.float32).reshape(n, -1)[:, :3]
        gv[:, :3] = np.asarray(g_vel, np.float32).reshape(n, -1)[:, :3]
        self._call("seed", gp.ctypes.data, gv.ctypes.data)

    def backward(self):
        self._call("backward")

    def grad(self):
        n = self.sim.count()
        gp = np.zeros((max(n, 1), 4), np.float32)
        gv = np.zeros((max(n, 1), 4), np.float32)
        self._call("grad", gp.ctypes.data, gv.ctypes.data)
        return gp[:n, :3], gv[:n, :3]

    def col_grad(self, ncols):
        """(ncols, 9) float32 of the substep reversed last: dL/d(centre),
        dL/d(linear velocity), dL/d(angular velocity) per collider."""
        g = np.zeros((max(int(ncols), 1), 9), np.float32)
        self._call("col_grad", g.ctypes.data, int(ncols))
        return g[: int(ncols)]

    def set_sor_path(self, one_launch):
        """True: the transposed solve in one cooperative launch where the
        device allows it; False (the default): one launch per sweep."""
        self._call("set_sor_path", 1 if one_launch else 0)

    def sor_path(self):
        """Whether the last backward() ran its transposed solve as one launch."""
        v = C.c_uint32()
        self._call("sor_path", C.byref(v))
        return bool(v.value)

    def close(self):
        if getattr(self, "h", None):
            self.lib.st_adj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

