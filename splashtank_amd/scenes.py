"""Synthetic scenes: particle lattices and collider records, seeded.

`fill_box` / `fill_sphere` follow GpuSim.fillBox / fillSphere
(src/shared.js:285-297): a lattice at spacing d starting d/2 in, coordinates
accumulated in doubles exactly as the JS loops do, and a jitter of
+-0.005*d/0.12 -- drawn from a seeded generator instead of Math.random so runs
repeat.  `box_collider` packs one collider record the way packColliders does
(src/gpu.js:42-70) around an analytic box distance grid.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def _axis(a0: float, a1: float, d: float) -> np.ndarray:
    """for (x = a0 + d/2; x < a1; x += d) -- JS double accumulation."""
    out = []
    x = a0 + d / 2
    while x < a1:
        out.append(x)
        x += d
    return np.array(out, dtype=np.float64)


def _axis_closed(r: float, d: float) -> np.ndarray:
    """for (x = -R; x <= R; x += d)"""
    out = []
    x = -r
    while x <= r:
        out.append(x)
        x += d
    return np.array(out, dtype=np.float64)


def fill_box(d, x0, x1, y0, y1, z0, z1, rng, vel=(0.0, 0.0, 0.0)):
    """Lattice in y (outer), x, z (inner) order, as shared.js:285-291."""
    ys, xs, zs = _axis(y0, y1, d), _axis(x0, x1, d), _axis(z0, z1, d)
    Y, X, Z = np.meshgrid(ys, xs, zs, indexing="ij")
    j = 0.01 * d / 0.12
    n = X.size
    jit = (rng.random((n, 3)) - 0.5) * j
    pos = np.empty((n, 4), np.float32)
    pos[:, 0] = X.reshape(-1) + jit[:, 0]
    pos[:, 1] = Y.reshape(-1) + jit[:, 1]
    pos[:, 2] = Z.reshape(-1) + jit[:, 2]
    vx, vy, vz = vel
    pos[:, 3] = math.sqrt(vx * vx + vy * vy + vz * vz)
    v = np.zeros((n, 4), np.float32)
    v[:, 0], v[:, 1], v[:, 2] = vx, vy, vz
    return pos, v


def fill_sphere(d, cx, cy, cz, r, rng, vel=(0.0, 0.0, 0.0)):
    """Lattice points of [-R, R]^3 inside the ball, y outer (shared.js:292-297)."""
    a = _axis_closed(r, d)
    Y, X, Z = np.meshgrid(a, a, a, indexing="ij")
    m = (X * X + Y * Y + Z * Z) <= r * r
    X, Y, Z = X[m], Y[m], Z[m]
    n = X.size
    j = 0.01 * d / 0.12
    jit = (rng.random((n, 3)) - 0.5) * j
    pos = np.empty((n, 4), np.float32)
    pos[:, 0] = cx + X + jit[:, 0]
    pos[:, 1] = cy + Y + jit[:, 1]
    pos[:, 2] = cz + Z + jit[:, 2]
    vx, vy, vz = vel
    pos[:, 3] = math.sqrt(vx * vx + vy * vy + vz * vz)
    v = np.zeros((n, 4), np.float32)
    v[:, 0], v[:, 1], v[:, 2] = vx, vy, vz
    return pos, v


def box_sdf(half, cell, margin):
    """Distance grid of an axis-aligned box of half extents `half` around the
    origin, on a lattice of spacing `cell` reaching `margin` past the box
    (the grid layout of buildSDF, shared.js:98-107: x fastest)."""
    half = np.asarray(half, np.float64)
    lo = -half - margin
    hi = half + margin
    dims = [int(math.ceil((hi[k] - lo[k]) / cell)) + 1 for k in range(3)]
    zs, ys, xs = [lo[k] + np.arange(dims[k]) * cell for k in (2, 1, 0)]
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    q = np.stack([np.abs(X) - half[0], np.abs(Y) - half[1], np.abs(Z) - half[2]], -1)
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(q.max(-1), 0.0)
    data = (outside + inside).astype(np.float32).reshape(-1)
    return dims, lo, data


def sphere_sdf(r, cell, margin):
    lo = np.full(3, -r - margin)
    dims = [int(math.ceil(2 * (r + margin) / cell)) + 1] * 3
    a = lo[0] + np.arange(dims[0]) * cell
    Z, Y, X = np.meshgrid(a, a, a, indexing="ij")
    data = (np.sqrt(X * X + Y * Y + Z * Z) - r).astype(np.float32).reshape(-1)
    return dims, lo, data


def collider_record(off, dims, origin, cell, pos, R=None, vel=(0, 0, 0), w=(0, 0, 0)):
    """32 floats of one collider, as packColliders (gpu.js:57-67)."""
    R = np.eye(3) if R is None else np.asarray(R, np.float64).reshape(3, 3)
    Rf = R.reshape(-1)
    cf = np.zeros(32, np.float32)
    cf[0] = off
    cf[1:4] = dims
    cf[4:7] = origin
    cf[7] = cell
    cf[8:11] = [Rf[0], Rf[3], Rf[6]]
    cf[12:15] = [Rf[1], Rf[4], Rf[7]]
    cf[16:19] = [Rf[2], Rf[5], Rf[8]]
    cf[20:23] = pos
    cf[24:27] = vel
    cf[28:31] = w
    return cf


@dataclass
class Scene:
    """Options, particles, step settings and (optional) moving colliders."""
    name: str
    d: float
    size: tuple
    open_: int
    pos: np.ndarray
    vel: np.ndarray
    params: dict
    colliders: list = field(default_factory=list)  # dicts: dims, origin, cell, data, path(t)->pos
    slots: int = 0
    ratio: float = 2.0

    @property
    def n(self):
        return self.pos.shape[0]

    def collider_state(self, t: float, dt: float):
        """Packed records and distance grids at time t; velocity from the pose
        change over the last dt as SceneRunner.step does (shared.js:531-537)."""
        if not self.colliders:
            return np.zeros(0, np.float32), np.zeros(0, np.float32), 0
        cfs, grids, off = [], [], 0
        for c in self.colliders:
            p1 = np.asarray(c["path"](t), np.float64)
            p0 = np.asarray(c["path"](t - dt), np.float64)
            v = (p1 - p0) / dt
            cfs.append(collider_record(off, c["dims"], c["origin"], c["cell"], p1, vel=v))
            grids.append(c["data"])
            off += c["data"].size
        return np.concatenate(cfs), np.concatenate(grids), 1


def dam_break(d=1.0 / 128, seed=1, size=(1.0, 1.0, 1.0), frac=(0.4, 0.6, 1.0), iters=40, dt=1.0 / 240):
    """Dam break: fluid box x in [x0, x0 + frac_x W], y in [0, frac_y H], z
    fraction about the centre; all walls closed.  At d = 1/128 in a unit box this
    is the 64^3 case: 51*77*128 = 502,656 particles."""
    W, H, D = size
    x0 = -W / 2
    z0 = -D / 2
    rng = np.random.default_rng(seed)
    pos, vel = fill_box(d, x0, x0 + frac[0] * W, 0.0, frac[1] * H, z0, z0 + frac[2] * D, rng)
    prm = dict(dt=dt, g=9.81, iters=iters, omega=1.7, flip=0.95, comp=6.0, tear=0.0, vmax=14.0)
    return Scene("dam_break", d, size, 0, pos, vel, prm)


def sparse_splash(d=1.0 / 512, seed=2, pool=0.25, sphere_r=0.12, sphere_y=0.7, sphere_v=-2.0,
                  box_half=0.1, box_speed=0.5, box_start=(-0.3, 0.2, 0.0), iters=80, dt=1.0 / 480):
    """Pool of depth `pool` across a unit box, a ball of fluid falling into it,
    and a box collider moving along x at box_speed.  At d = 1/512 this is the
    256^3 case (about 34.5M particles at 8 per cell)."""
    rng = np.random.default_rng(seed)
    p1, v1 = fill_box(d, -0.5, 0.5, 0.0, pool, -0.5, 0.5, rng)
    p2, v2 = fill_sphere(d, 0.0, sphere_y, 0.0, sphere_r, rng, vel=(0.0, sphere_v, 0.0))
    pos = np.concatenate([p1, p2])
    vel = np.concatenate([v1, v2])
    h = 2 * d
    dims, lo, data = box_sdf((box_half,) * 3, h, 4 * h)
    bs = np.asarray(box_start, np.float64)
    col = dict(dims=dims, origin=lo, cell=h, data=data,
               path=lambda t, bs=bs: bs + np.array([box_speed * t, 0.0, 0.0]))
    prm = dict(dt=dt, g=9.81, iters=iters, omega=1.7, flip=0.95, comp=6.0, tear=0.0, vmax=14.0)
    return Scene("sparse_splash", d, (1.0, 1.0, 1.0), 0, pos, vel, prm, colliders=[col])
