"""Differentiable rollouts of the FLIP substep through the hand-written adjoint
kernels (st_adj_*), as a torch.autograd.Function, plus the particle-density loss
of the inverse problem (match a target density after K substeps).

The gradient is the adjoint of the substep at fixed topology (the tile set,
cell types, particle counts and branches of the forward pass), with a snapshot
every `every` substeps: the backward pass restores each snapshot in turn (last
first), re-runs its substeps recording their tapes and reverses them, so tape
memory is bounded by `every` substeps.

Keyframed colliders (KeyedCollider) are parameters too: their centre at each
substep is the keyframe curve of keyPose (shared.js:608-621), linear in the key
positions, and their velocity the pose difference over the substep
(shared.js:531-537); the per-substep collider gradients of st_adj_col_grad go
back to the key positions through those two linear maps.
"""
from __future__ import annotations

import numpy as np
import torch

from . import capi, scenes


def _f4(a: torch.Tensor) -> np.ndarray:
    x = a.detach().to("cpu", torch.float32).numpy()
    out = np.zeros((x.shape[0], 4), np.float32)
    out[:, :3] = x[:, :3]
    return out


def pose_weights(times, t: float, smooth: bool = True) -> np.ndarray:
    """keyPose (shared.js:608-621) as weights over the keys: the pose at time t
    is weights @ key values (it is linear in them).  Before the first key and
    after the last the pose holds; between keys i and i+1 it is linear, or with
    `smooth` the Catmull-Rom curve through keys i-1..i+2 (the end keys repeated)."""
    times = [float(x) for x in times]
    n = len(times)
    w = np.zeros(n)
    if t <= times[0]:
        w[0] = 1.0
        return w
    if t >= times[-1]:
        w[-1] = 1.0
        return w
    i = 0
    while i < n - 2 and t >= times[i + 1]:
        i += 1
    a, b = i, i + 1
    u = (t - times[a]) / max(times[b] - times[a], 1e-6)
    if not smooth:
        w[a] += 1.0 - u
        w[b] += u
        return w
    p0 = i - 1 if i > 0 else a
    p3 = i + 2 if i + 2 < n else b
    u2, u3 = u * u, u * u * u
    w[a] += 0.5 * (2.0 - 5.0 * u2 + 3.0 * u3)
    w[b] += 0.5 * (u + 4.0 * u2 - 3.0 * u3)
    w[p0] += 0.5 * (-u + 2.0 * u2 - u3)
    w[p3] += 0.5 * (-u2 + u3)
    return w


class KeyedCollider:
    """A rigid collider (distance grid dims/origin/cell/data, fixed rotation R)
    whose centre follows keyframes at `times`: at substep time t its centre is
    pose_weights(times, t) @ keys and its velocity (centre(t) - centre(t - dt)) /
    dt, or zero when that is faster than `lim` (SceneRunner's jump rule,
    shared.js:528-535)."""

    def __init__(self, dims, origin, cell, data, times, smooth=True, R=None, lim=20.0):
        self.dims = [int(x) for x in dims]
        self.origin = np.asarray(origin, np.float64)
        self.cell = float(cell)
        self.data = np.asarray(data, np.float32).reshape(-1)
        self.times = [float(x) for x in times]
        self.smooth = bool(smooth)
        self.R = None if R is None else np.asarray(R, np.float64).reshape(3, 3)
        self.lim = float(lim)

    def at(self, keys: np.ndarray, t: float, dt: float):
        """(centre, velocity, d centre / d keys, d velocity / d keys) at time t."""
        wc = pose_weights(self.times, t, self.smooth)
        wv = (wc - pose_weights(self.times, t - dt, self.smooth)) / dt
        c, v = wc @ keys, wv @ keys
        if float(v @ v) >= self.lim * self.lim:
            v, wv = np.zeros(3), np.zeros_like(wv)
        return c, v, wc, wv


class Rollout:
    """One tank run for `nsteps` substeps from given particle states.  Calling it
    with (pos0, vel0) tensors, (N, 3), particle-id order, and one (nkeys, 3) key
    position tensor per keyed collider, returns the final (pos, vel),
    differentiable with respect to all of them.  Substep s runs at time t0 + s dt."""

    def __init__(self, opts: capi.StOpts, prm: capi.StParams, nsteps: int, every: int = 10, cols=(), t0: float = 0.0):
        self.sim = capi.FlipSimC(opts)
        self.base = capi.SnapshotC(self.sim)  # the empty tank: every rollout starts here
        self.prm = prm
        self.nsteps = int(nsteps)
        self.every = max(1, int(every))
        self.cols = list(cols)
        self.t0 = float(t0)
        self.keys = [None] * len(self.cols)
        self._sdf = np.concatenate([c.data for c in self.cols]) if self.cols else np.zeros(0, np.float32)

    def __call__(self, pos0: torch.Tensor, vel0: torch.Tensor, *keys: torch.Tensor):
        return _RolloutFn.apply(pos0, vel0, self, *keys)

    def set_keys(self, keys):
        if len(keys) != len(self.cols):
            raise ValueError(f"{len(self.cols)} keyed colliders, {len(keys)} key sets")
        self.keys = [np.asarray(k, np.float64).reshape(len(c.times), 3) for k, c in zip(keys, self.cols)]

    def _colliders(self, s: int):
        """Upload the collider records of substep s; returns the weights that map
        the keys to their centres and velocities."""
        if not self.cols:
            return []
        dt = float(self.prm.dt)
        t = self.t0 + s * dt
        recs, ws, off = [], [], 0
        for c, k in zip(self.cols, self.keys):
            cen, vel, wc, wv = c.at(k, t, dt)
            recs.append(scenes.collider_record(off, c.dims, c.origin, c.cell, cen, R=c.R, vel=vel))
            ws.append((wc, wv))
            off += c.data.size
        self.sim.set_colliders(np.concatenate(recs), self._sdf, 1)
        return ws

    def forward_raw(self, pos0: np.ndarray, vel0: np.ndarray, keep_snaps: bool = False):
        """Run the substeps; return final (pos, vel) as (N, 3) float32 arrays."""
        self.base.restore()
        p4 = np.zeros((pos0.shape[0], 4), np.float32)
        v4 = np.zeros_like(p4)
        p4[:, :3] = pos0[:, :3]
        v4[:, :3] = vel0[:, :3]
        self.sim.upload(p4, v4)
        snaps = []
        for s in range(self.nsteps):
            if keep_snaps and s % self.every == 0:
                snaps.append(capi.SnapshotC(self.sim))
            self._colliders(s)
            self.sim.step(self.prm)
        pos = self.sim.dump("pos")[:, :3].copy()
        vel = self.sim.dump("vel")[:, :3].copy()
        return pos, vel, snaps

    def backward_raw(self, snaps, g_pos: np.ndarray, g_vel: np.ndarray):
        """Gradient of the initial (pos, vel), and of the key positions of each
        keyed collider, from that of the final state."""
        adj = capi.AdjointC(self.sim, self.every)
        gkeys = [np.zeros((len(c.times), 3)) for c in self.cols]
        try:
            adj.seed(g_pos, g_vel)
            for k in reversed(range(len(snaps))):
                snaps[k].restore()
                first = k * self.every
                seg = min(self.every, self.nsteps - first)
                ws = []
                for s in range(first, first + seg):
                    ws.append(self._colliders(s))
                    adj.record(self.prm)
                for s in reversed(range(seg)):
                    adj.backward()
                    if self.cols:
                        g = adj.col_grad(len(self.cols)).astype(np.float64)
                        for q, (wc, wv) in enumerate(ws[s]):
                            gkeys[q] += np.outer(wc, g[q, 0:3]) + np.outer(wv, g[q, 3:6])
            gp, gv = adj.grad()
            return gp, gv, gkeys
        finally:
            adj.close()

    def close(self):
        self.base.close()
        self.sim.close()


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos0, vel0, ro: Rollout, *keys):
        ro.set_keys([k.detach().to("cpu", torch.float64).numpy() for k in keys])
        pos, vel, snaps = ro.forward_raw(_f4(pos0)[:, :3], _f4(vel0)[:, :3], keep_snaps=True)
        ctx.ro = ro
        ctx.snaps = snaps
        ctx.keys = [k.detach().to("cpu", torch.float64).numpy() for k in keys]
        ctx.kmeta = [(k.device, k.dtype) for k in keys]
        ctx.dev = pos0.device
        ctx.dtype = pos0.dtype
        return (torch.from_numpy(pos).to(pos0.device, pos0.dtype), torch.from_numpy(vel).to(vel0.device, vel0.dtype))

    @staticmethod
    def backward(ctx, g_pos, g_vel):
        n = g_pos.shape[0]
        gp = np.zeros((n, 3), np.float32) if g_pos is None else g_pos.detach().to("cpu", torch.float32).numpy()
        gv = np.zeros((n, 3), np.float32) if g_vel is None else g_vel.detach().to("cpu", torch.float32).numpy()
        ctx.ro.set_keys(ctx.keys)
        dp, dv, dk = ctx.ro.backward_raw(ctx.snaps, gp, gv)
        for q in ctx.snaps:
            q.close()
        ctx.snaps = []
        gk = [torch.from_numpy(g).to(dev, dt) for g, (dev, dt) in zip(dk, ctx.kmeta)]
        return (torch.from_numpy(dp).to(ctx.dev, ctx.dtype), torch.from_numpy(dv).to(ctx.dev, ctx.dtype), None, *gk)


def splat_density(pos: torch.Tensor, origin, h: float, dims) -> torch.Tensor:
    """Particles per cell by trilinear splatting onto cell centres (a dims-shaped
    tensor, x fastest in memory order (i, j, k) -> [i, j, k])."""
    nx, ny, nz = dims
    g = (pos - torch.as_tensor(origin, dtype=pos.dtype, device=pos.device)) / h - 0.5
    f0 = torch.floor(g)
    fr = g - f0
    i0 = f0.to(torch.int64)
    out = torch.zeros(nx * ny * nz, dtype=pos.dtype, device=pos.device)
    for c in range(8):
        d = [(c >> a) & 1 for a in range(3)]
        w = torch.ones_like(fr[:, 0])
        idx = []
        ok = torch.ones_like(w, dtype=torch.bool)
        for a, n in enumerate((nx, ny, nz)):
            w = w * (fr[:, a] if d[a] else 1.0 - fr[:, a])
            ia = i0[:, a] + d[a]
            ok = ok & (ia >= 0) & (ia < n)
            idx.append(ia.clamp(0, n - 1))
        flat = idx[0] * (ny * nz) + idx[1] * nz + idx[2]
        out = out.index_add(0, flat, torch.where(ok, w, torch.zeros_like(w)))
    return out.view(nx, ny, nz)


def gaussian_target(dims, origin, h, centres, sigma, amp):
    """Sum of Gaussian blobs sampled at cell centres."""
    nx, ny, nz = dims
    ax = [origin[a] + (np.arange(n) + 0.5) * h for a, n in enumerate((nx, ny, nz))]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    t = np.zeros(dims)
    for c in centres:
        t += amp * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * sigma * sigma))
    return t
