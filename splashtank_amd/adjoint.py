"""Differentiable rollouts of the FLIP substep through the hand-written adjoint
kernels (st_adj_*), as a torch.autograd.Function, plus the particle-density loss
of the inverse problem (match a target density after K substeps).

The gradient is the adjoint of the substep at fixed topology (the tile set,
cell types, particle counts and branches of the forward pass), with a snapshot
every `every` substeps: the backward pass restores each snapshot in turn (last
first), re-runs its substeps recording their tapes and reverses them, so tape
memory is bounded by `every` substeps.
"""
from __future__ import annotations

import numpy as np
import torch

from . import capi


def _f4(a: torch.Tensor) -> np.ndarray:
    x = a.detach().to("cpu", torch.float32).numpy()
    out = np.zeros((x.shape[0], 4), np.float32)
    out[:, :3] = x[:, :3]
    return out


class Rollout:
    """One tank run for `nsteps` substeps from given particle states.  Calling it
    with (pos0, vel0) tensors, (N, 3), particle-id order, returns the final
    (pos, vel), differentiable with respect to both."""

    def __init__(self, opts: capi.StOpts, prm: capi.StParams, nsteps: int, every: int = 10):
        self.sim = capi.FlipSimC(opts)
        self.base = capi.SnapshotC(self.sim)  # the empty tank: every rollout starts here
        self.prm = prm
        self.nsteps = int(nsteps)
        self.every = max(1, int(every))

    def __call__(self, pos0: torch.Tensor, vel0: torch.Tensor):
        return _RolloutFn.apply(pos0, vel0, self)

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
            self.sim.step(self.prm)
        pos = self.sim.dump("pos")[:, :3].copy()
        vel = self.sim.dump("vel")[:, :3].copy()
        return pos, vel, snaps

    def backward_raw(self, snaps, g_pos: np.ndarray, g_vel: np.ndarray):
        """Gradient of the initial (pos, vel) from that of the final state."""
        adj = capi.AdjointC(self.sim, self.every)
        try:
            adj.seed(g_pos, g_vel)
            for k in reversed(range(len(snaps))):
                snaps[k].restore()
                seg = min(self.every, self.nsteps - k * self.every)
                for _ in range(seg):
                    adj.record(self.prm)
                for _ in range(seg):
                    adj.backward()
            return adj.grad()
        finally:
            adj.close()

    def close(self):
        self.base.close()
        self.sim.close()


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos0, vel0, ro: Rollout):
        pos, vel, snaps = ro.forward_raw(_f4(pos0)[:, :3], _f4(vel0)[:, :3], keep_snaps=True)
        ctx.ro = ro
        ctx.snaps = snaps
        ctx.dev = pos0.device
        ctx.dtype = pos0.dtype
        return (torch.from_numpy(pos).to(pos0.device, pos0.dtype), torch.from_numpy(vel).to(vel0.device, vel0.dtype))

    @staticmethod
    def backward(ctx, g_pos, g_vel):
        n = g_pos.shape[0]
        gp = np.zeros((n, 3), np.float32) if g_pos is None else g_pos.detach().to("cpu", torch.float32).numpy()
        gv = np.zeros((n, 3), np.float32) if g_vel is None else g_vel.detach().to("cpu", torch.float32).numpy()
        dp, dv = ctx.ro.backward_raw(ctx.snaps, gp, gv)
        for q in ctx.snaps:
            q.close()
        ctx.snaps = []
        return (torch.from_numpy(dp).to(ctx.dev, ctx.dtype), torch.from_numpy(dv).to(ctx.dev, ctx.dtype), None)


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
