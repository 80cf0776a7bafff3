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

    def params_at(self, t: float, **kw) -> dict:
        """Step settings at simulated time t (the ocean's waves move with o.time)."""
        p = dict(self.params)
        if p.get("ocean_on"):
            p["ocean_time"] = t
        p.update(kw)
        return p

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
            # optional rotation: R(t) (local -> world) and a constant angular velocity
            R = c["rot"](t) if "rot" in c else None
            w = c.get("omega", (0.0, 0.0, 0.0))
            cfs.append(collider_record(off, c["dims"], c["origin"], c["cell"], p1, R=R, vel=v, w=w))
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


def axis_rotation(axis, angle):
    """Rotation matrix (local -> world) by `angle` radians about the unit `axis`
    (Rodrigues' formula)."""
    k = np.asarray(axis, np.float64)
    k = k / np.linalg.norm(k)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * (K @ K)


def spinning_colliders(d=1.0 / 64, seed=5, pool=0.3, iters=20, dt=1.0 / 240, spin=3.0):
    """A pool stirred by two overlapping colliders: a flat box tilted off the axes
    and spinning about a tilted axis at `spin` rad/s while it moves down and
    sideways, and a sphere grid (turned 40 degrees about x) sinking through the
    box's end.  Exercises the collider-local transform, the spin term of the
    surface velocity, the first-hit order of the cell lookup and the sequential
    pushes of two colliders."""
    rng = np.random.default_rng(seed)
    pos, vel = fill_box(d, -0.5, 0.5, 0.0, pool, -0.5, 0.5, rng)
    h = 2 * d
    half = np.array([0.16, 0.05, 0.1])
    dims, lo, data = box_sdf(half, h / 2, 4 * h / 2)
    ax = np.array([0.3, 1.0, 0.2])
    ax = ax / np.linalg.norm(ax)
    R0 = axis_rotation((0, 0, 1), math.radians(25))
    box = dict(dims=dims, origin=lo, cell=h / 2, data=data,
               path=lambda t: np.array([-0.05 + 0.4 * t, 0.34 - 0.6 * t, 0.02]),
               rot=lambda t: axis_rotation(ax, spin * t) @ R0, omega=tuple(spin * ax))
    sd, slo, sdata = sphere_sdf(0.09, h / 2, 4 * h / 2)
    Rs = axis_rotation((1, 0, 0), math.radians(40))
    sph = dict(dims=sd, origin=slo, cell=h / 2, data=sdata,
               path=lambda t: np.array([0.12 + 0.4 * t, 0.36 - 0.9 * t, 0.04]),
               rot=lambda t: Rs, omega=(0.0, 0.0, 1.5))
    prm = dict(dt=dt, g=9.81, iters=iters, omega=1.7, flip=0.95, comp=6.0, tear=0.0, vmax=14.0)
    return Scene("spinning_colliders", d, (1.0, 1.0, 1.0), 0, pos, vel, prm, colliders=[box, sph])


def nb_seed(d, h, size, fills, band=3, rng=None, cap=None):
    """Narrow-band start (nbSeed, shared.js:805-853) for box and sphere fills:
    particles only within `band` cells of each shape's free surface (box faces
    on the tank walls are not free surfaces), then spare parked particles the
    solver refills the band from.  Returns pos, vel and the packed fill shapes
    (8 floats each: box 1,x0,y0,z0,x1,y1,z1,0 / sphere 2,x,y,z,r,0,0,0) with
    wall-touching box faces pushed out to +-1e4 as nbSeed does."""
    rng = rng or np.random.default_rng(0)
    W, H, D = size
    b = dict(x0=-W / 2, x1=W / 2, y0=0.0, y1=H, z0=-D / 2, z1=D / 2)
    bandw = max(2, band) * h
    near = 0.6 * h
    jit = 0.01 * d / 0.12
    ps, packed = [], []
    for f in fills:
        if f["t"] == "box":
            e = dict(x0=-1e4 if f["x0"] <= b["x0"] + near else f["x0"], x1=1e4 if f["x1"] >= b["x1"] - near else f["x1"],
                     y0=-1e4 if f["y0"] <= b["y0"] + near else f["y0"], y1=1e4 if f["y1"] >= b["y1"] - near else f["y1"],
                     z0=-1e4 if f["z0"] <= b["z0"] + near else f["z0"], z1=1e4 if f["z1"] >= b["z1"] - near else f["z1"])
            ys, xs, zs = _axis(f["y0"], f["y1"], d), _axis(f["x0"], f["x1"], d), _axis(f["z0"], f["z1"], d)
            Y, X, Z = np.meshgrid(ys, xs, zs, indexing="ij")
            X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
            dist = np.minimum.reduce([X - e["x0"], e["x1"] - X, Y - e["y0"], e["y1"] - Y, Z - e["z0"], e["z1"] - Z])
            m = dist <= bandw
            ps.append(np.stack([X[m], Y[m], Z[m]], 1))
            packed.append([1, e["x0"], e["y0"], e["z0"], e["x1"], e["y1"], e["z1"], 0])
        else:
            r = f["r"]
            a = _axis_closed(r, d)
            Y, X, Z = np.meshgrid(a, a, a, indexing="ij")
            X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
            dist = r - np.sqrt(X * X + Y * Y + Z * Z)
            m = (dist >= 0) & (dist <= bandw)
            ps.append(np.stack([f["x"] + X[m], f["y"] + Y[m], f["z"] + Z[m]], 1))
            packed.append([2, f["x"], f["y"], f["z"], r, 0, 0, 0])
    P = np.concatenate(ps) if ps else np.zeros((0, 3))
    n = P.shape[0]
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = P + (rng.random((n, 3)) - 0.5) * jit
    spare = int(np.ceil(0.35 * n)) + 4000
    if cap is not None:
        spare = max(0, min(spare, cap - n))
    park = np.zeros((spare, 4), np.float32)
    park[:, 1] = -1e6
    pos = np.concatenate([pos, park])
    vel = np.zeros_like(pos)
    return pos, vel, np.asarray(packed, np.float32).reshape(-1)


def ocean_eta_vel(W, level, y0, x, y, z, t):
    """Surface height and water velocity of the wave list W (n, 8) in f64 over
    arrays of points: the sums of oceanEtaJS / oceanVelJS (src/shared.js:792-803),
    a Wheeler-stretched Airy velocity under the surface."""
    W = np.asarray(W, np.float64)
    x, y, z = (np.asarray(q, np.float64) for q in (x, y, z))
    kx, kz, a, w, ph, k = (W[:, q].reshape((-1,) + (1,) * x.ndim) for q in range(6))
    th = kx * x + kz * z - w * t + ph
    eta = level + (a * np.cos(th)).sum(0)
    Dp = max(level - y0, 1e-3)
    ys = np.clip((y - y0) * Dp / np.maximum(Dp + eta - level, 0.1 * Dp), 0.0, Dp)
    kk = np.maximum(k, 1e-4)
    e1, e2 = np.exp(kk * (ys - Dp)), np.exp(-kk * (ys + Dp))
    dn = np.maximum(1.0 - np.exp(-2.0 * kk * Dp), 1e-4)
    hc = a * w * (e1 + e2) / dn * np.cos(th) / kk
    vel = np.stack([(hc * kx).sum(0), (a * w * (e1 - e2) / dn * np.sin(th)).sum(0), (hc * kz).sum(0)], -1)
    return eta, vel


def ocean_seed(d, h, size, oc, W, t=0.0, band=3, rng=None):
    """The ocean part of nbSeed (shared.js:840-848): on every lattice column of
    the box outside the outer half of the boundary layer, particles from a band
    below the surface (eta capped a spacing under the lid) up to it, moving with
    the waves.  y runs from y0 + d/2 + max(0, floor((eta - band - y0) / d)) d by
    d below eta; the points are computed as start + k d rather than the JS loop's
    running sum.  Jitter +-0.005 d / 0.12 on positions only."""
    rng = rng or np.random.default_rng(0)
    Wd, Hd, Dd = size
    x0, x1, y0, y1, z0, z1 = -Wd / 2, Wd / 2, 0.0, Hd, -Dd / 2, Dd / 2
    bandw = max(2, band) * h
    xs, zs = _axis(x0, x1, d), _axis(z0, z1, d)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    X, Z = X.ravel(), Z.ravel()
    dw = np.minimum(np.minimum(X - x0, x1 - X), np.minimum(Z - z0, z1 - Z))
    keep = 1.0 - dw / max(oc["layer"], 1e-4) < 0.5
    X, Z = X[keep], Z[keep]
    eta, _ = ocean_eta_vel(W, oc["level"], y0, X, np.zeros_like(X), Z, t)
    eta = np.minimum(eta, y1 - d)
    ystart = y0 + d / 2 + np.maximum(0.0, np.floor((eta - bandw - y0) / d)) * d
    m = int(math.ceil((bandw + d) / d)) + 2
    Y = ystart[:, None] + np.arange(m)[None, :] * d
    inside = Y < eta[:, None]
    Xc = np.broadcast_to(X[:, None], Y.shape)[inside]
    Zc = np.broadcast_to(Z[:, None], Y.shape)[inside]
    Yc = Y[inside]
    _, v = ocean_eta_vel(W, oc["level"], y0, Xc, Yc, Zc, t)
    jit = 0.01 * d / 0.12
    n = Xc.size
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = np.stack([Xc, Yc, Zc], 1) + (rng.random((n, 3)) - 0.5) * jit
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = v
    pos[:, 3] = np.minimum(np.linalg.norm(v, axis=1), 63.0)
    return pos, vel


def stress(d=1.0 / 96, seed=9, level=0.35, layer=0.12, band=3, iters=30, dt=1.0 / 240,
           ocean=(0.04, 0.45, 30.0, 40.0, 7), drop=(0.12, 0.72, -0.05, 0.09), spray_frac=0.25):
    """Every feature at once: the ocean side boundary (24 Airy waves, the
    boundary layer `layer` wide) over a deep pool of depth `level` in
    narrow-band mode (band 3: only the band under the surface has particles), a
    ball of water dropped into it, whitewater with the air field, and two SDF
    colliders -- a box grid tilted and spinning as it ploughs along x through
    the surface, and a sphere grid sinking and rising through it.  At d = 1/1024
    this is the 512^3 full-feature stress case.  Step settings carry the ocean
    and the whitewater; `ocean_time` must follow the simulated time
    (Scene.params_at)."""
    from splashtank_amd import capi
    rng = np.random.default_rng(seed)
    h = 2 * d
    height, length, odir, spread, oseed = ocean
    oc = dict(height=height, length=length, dir=odir, spread=spread, seed=oseed, level=level, layer=layer)
    W = capi.ocean_waves(height, length, odir, spread, oseed, g=9.81, depth=level)
    p1, v1 = ocean_seed(d, h, (1.0, 1.0, 1.0), oc, W, 0.0, band, rng)
    fills = [dict(t="sphere", x=drop[0], y=drop[1], z=drop[2], r=drop[3])]
    p2, v2, packed = nb_seed(d, h, (1.0, 1.0, 1.0), fills, band=band, rng=rng)
    live2 = p2[:, 1] > -1e5
    ocean_n = p1.shape[0]
    spare = int(math.ceil(0.35 * (ocean_n + int(live2.sum())))) + 4000
    park = np.zeros((spare, 4), np.float32)
    park[:, 1] = -1e6
    pos = np.concatenate([p1, p2[live2], park])
    vel = np.concatenate([v1, v2[live2], np.zeros_like(park)])
    bh = np.array([0.09, 0.05, 0.12])
    dims, lo, data = box_sdf(bh, h / 2, 4 * h / 2)
    ax = np.array([0.2, 1.0, -0.3])
    ax = ax / np.linalg.norm(ax)
    R0 = axis_rotation((0, 0, 1), math.radians(20))
    box = dict(dims=dims, origin=lo, cell=h / 2, data=data,
               path=lambda t: np.array([-0.28 + 0.5 * t, level + 0.02 - 0.2 * t, 0.1]),
               rot=lambda t: axis_rotation(ax, 2.0 * t) @ R0, omega=tuple(2.0 * ax))
    sd, slo, sdata = sphere_sdf(0.08, h / 2, 4 * h / 2)
    sph = dict(dims=sd, origin=slo, cell=h / 2, data=sdata,
               path=lambda t: np.array([-0.1, level + 0.12 - 0.8 * t + 2.0 * t * t, -0.18]),
               rot=lambda t: axis_rotation((1, 0, 0), math.radians(30)), omega=(0.0, 0.0, 0.0))
    prm = dict(dt=dt, g=9.81, iters=iters, omega=1.7, flip=0.95, comp=6.0, tear=0.0, vmax=14.0,
               nb_on=1, nb_band=float(band), ocean_on=1, waves=W, ocean_level=level, ocean_layer=layer,
               spr_on=1, spr_rate=40.0, spr_min=1.0, spr_life=1.2, spr_mist=0.3, spr_drag=1.0, spr_mdrag=5.0,
               air_on=1, air_couple=3.0, air_decay=0.6, air_blur=0.35)
    sc = Scene("stress", d, (1.0, 1.0, 1.0), 0, pos, vel, prm, colliders=[box, sph])
    sc.nb_fills = packed
    sc.spray_cap = max(1024, int(spray_frac * ocean_n))
    sc.ocean = oc
    return sc


def nb_pool(d=1.0 / 64, depth=0.5, band=3, seed=4, iters=30, dt=1.0 / 240, drop=None):
    """A deep pool in narrow-band mode (only the band under the surface has
    particles), optionally with a ball of water dropped into it."""
    h = 2 * d
    fills = [dict(t="box", x0=-0.5, x1=0.5, y0=0.0, y1=depth, z0=-0.5, z1=0.5)]
    if drop:
        fills.append(dict(t="sphere", x=drop[0], y=drop[1], z=drop[2], r=drop[3]))
    pos, vel, packed = nb_seed(d, h, (1.0, 1.0, 1.0), fills, band=band, rng=np.random.default_rng(seed))
    prm = dict(dt=dt, g=9.81, iters=iters, omega=1.7, flip=0.95, comp=6.0, tear=0.0, vmax=14.0,
               nb_on=1, nb_band=float(band))
    sc = Scene("nb_pool", d, (1.0, 1.0, 1.0), 0, pos, vel, prm)
    sc.nb_fills = packed
    return sc


def batched_tanks(count=64, seed=3, nmin=32, nmax=128, iters=40, dt=1.0 / 240):
    """Independent tanks for batched stepping: per tank a unit box of n^3 cells
    with n log-uniform in [nmin, nmax] rounded to a multiple of 8, a cell/spacing
    ratio in [1.59, 2] (4 to 8 particles per cell) and fluid fractions fx, fy
    from U(0.2, 0.6).  Even tanks are dam breaks (fluid x in [-0.5, -0.5 + fx],
    y in [0, fy], full z); odd tanks drop a ball of radius fx / 3 from y = 0.7
    into a pool of depth fy / 2.  All walls closed, zero initial velocity."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax))) / 8.0)) * 8
        n = max(8, n)
        ratio = float(rng.uniform(1.59, 2.0))
        fx, fy = (float(v) for v in rng.uniform(0.2, 0.6, 2))
        h = 1.0 / n
        d = h / ratio
        prng = np.random.default_rng(int(rng.integers(1 << 31)))
        if k % 2 == 0:
            pos, vel = fill_box(d, -0.5, -0.5 + fx, 0.0, fy, -0.5, 0.5, prng)
            name = f"tank{k}_dam_n{n}"
        else:
            p1, v1 = fill_box(d, -0.5, 0.5, 0.0, 0.5 * fy, -0.5, 0.5, prng)
            p2, v2 = fill_sphere(d, 0.0, 0.7, 0.0, fx / 3.0, prng)
            pos, vel = np.concatenate([p1, p2]), np.concatenate([v1, v2])
            name = f"tank{k}_drop_n{n}"
        prm = dict(dt=dt, g=9.81, iters=iters, omega=1.7, flip=0.95, comp=6.0, tear=0.0, vmax=14.0)
        out.append(Scene(name, d, (1.0, 1.0, 1.0), 0, pos, vel, prm, ratio=ratio))
    return out


def _js_hypot3(a, b, c):
    """Math.hypot of three finite numbers as JS engines evaluate it: scaled by the
    largest magnitude, squares summed with Kahan compensation, root, rescale."""
    m = max(abs(a), abs(b), abs(c))
    if m == 0.0:
        return 0.0
    s = comp = 0.0
    for x in (a, b, c):
        n = x / m
        summand = n * n - comp
        prelim = s + summand
        comp = (prelim - s) - summand
        s = prelim
    return math.sqrt(s) * m


def drip_samples(sdf, h, k):
    """dripSamples (src/flip/bake.js:19-40): points over a collider's surface from
    its distance grid -- grid points within 0.6 cells of the surface moved onto it
    along the gradient, thinned to about one per solver cell of surface (100 to
    6000) by a seeded stride -- as 8 f32 each: local position, collider number k,
    local normal, wetness 0.  `sdf` = dict(id, dims, origin, cell, data)."""
    nx, ny, nz = (int(x) for x in sdf["dims"])
    c, o = float(sdf["cell"]), [float(x) for x in sdf["origin"]]
    d = np.asarray(sdf["data"], np.float32).reshape(-1)

    def at(i, j, q):
        return float(d[min(nx - 1, max(0, i)) + nx * (min(ny - 1, max(0, j)) + ny * min(nz - 1, max(0, q)))])

    pts = []
    for q in range(1, nz - 1):
        for j in range(1, ny - 1):
            for i in range(1, nx - 1):
                v = float(d[i + nx * (j + ny * q)])
                if abs(v) > 0.6 * c:
                    continue
                gx, gy, gz = at(i + 1, j, q) - at(i - 1, j, q), at(i, j + 1, q) - at(i, j - 1, q), at(i, j, q + 1) - at(i, j, q - 1)
                gl = _js_hypot3(gx, gy, gz)
                if gl < 1e-9:
                    continue
                gx, gy, gz = gx / gl, gy / gl, gz / gl
                pts.append((o[0] + i * c - gx * v, o[1] + j * c - gy * v, o[2] + q * c - gz * v, gx, gy, gz))
    nsurf = len(pts)
    want = max(100, min(6000, math.floor(nsurf * c * c / (h * h) + 0.5)))
    step = max(1.0, nsurf / want)
    sd = 7 + int(sdf["id"])
    out = []
    sd = (sd * 16807) % 2147483647
    t = sd / 2147483647 * step
    while t < nsurf:
        p = pts[math.floor(t)]
        out.append((p[0], p[1], p[2], float(k), p[3], p[4], p[5], 0.0))
        t += step
    return np.array(out, np.float32).reshape(-1, 8)
