"""TEST INFRASTRUCTURE: an FP64 restatement of the FLIP substep on a dense grid,
in torch, for the adjoint tests.

It is the continuous map the GPU substep computes: the same stages and
formulas as the oracle (oracle/flip_oracle.cpp: stageP2G, classify, normFaces,
divergence, relax, project, stageG2P, collideP), with the fixed-point p2g sums
replaced by the real sums they round and every discrete choice (cell types,
counts, tiles kept for the warm start, face weight test, speed cap, box clamp,
wall contact and its normal) either taken from the state or, for finite
differences, frozen from a base run.  Closed box, no colliders, drag 0, tear 0
(the adjoint's domain).  torch autograd of it is the FP64 reference gradient;
central differences of it pin that gradient.
"""
from __future__ import annotations

import math

import numpy as np
import torch

T64 = torch.float64


class RefSim:
    def __init__(self, dims, prm):
        """dims: st_dims of the tank (from the oracle); prm: the step keywords."""
        self.nx, self.ny, self.nz = int(dims.nx), int(dims.ny), int(dims.nz)
        self.h = float(dims.h)
        self.ih = 1.0 / float(dims.h)
        self.x0 = torch.tensor([dims.x0, dims.y0, dims.z0], dtype=T64)
        self.x1 = torch.tensor([dims.x1, dims.y1, dims.z1], dtype=T64)
        self.margin = 0.3 * float(dims.d)
        self.rest = float(dims.rest)
        self.t = (int(dims.tx), int(dims.ty), int(dims.tz))
        self.dt = float(prm["dt"])
        self.g = float(prm.get("g", 9.81))
        self.flip = float(prm.get("flip", 0.95))
        self.comp = float(prm.get("comp", 6.0))
        om = float(prm.get("omega", 0.0))
        n = max(self.nx, self.ny, self.nz)
        self.omega = om if om != 0.0 else min(1.9, 2.0 / (1.0 + math.sin(math.pi / n)))
        self.vmax = float(prm.get("vmax", 14.0)) or 14.0
        self.iters = max(1, int(prm.get("iters", 30)))
        self.shape = (self.nx, self.ny, self.nz)
        I, J, K = np.meshgrid(np.arange(self.nx), np.arange(self.ny), np.arange(self.nz), indexing="ij")
        self.parity = torch.from_numpy(((I + J + K) & 1).astype(np.int64))
        # non-wall neighbours of every cell (walls only at the box faces here)
        nns = np.zeros(self.shape)
        for a, n_a in enumerate(self.shape):
            idx = (I, J, K)[a]
            nns += (idx > 0).astype(float) + (idx < n_a - 1).astype(float)
        self.nns = torch.from_numpy(nns)

    # ---- discrete state ----
    def cells(self, pos):
        c = torch.floor((pos.detach() - self.x0) * self.ih).to(torch.int64)
        return torch.stack([c[:, a].clamp(0, n - 1) for a, n in enumerate(self.shape)], 1)

    def need(self, cells):
        """tiles within two cells of a particle's cell (bin_mark)"""
        need = torch.zeros(self.t, dtype=torch.bool)
        lo = [(cells[:, a] - 2).clamp(min=0) >> 3 for a in range(3)]
        hi = [(cells[:, a] + 2).clamp(max=self.t[a] * 8 - 1) >> 3 for a in range(3)]
        for dx in range(2):
            for dy in range(2):
                for dz in range(2):
                    x = torch.minimum(lo[0] + dx, hi[0])
                    y = torch.minimum(lo[1] + dy, hi[1])
                    z = torch.minimum(lo[2] + dz, hi[2])
                    need[x, y, z] = True
        return need

    def tile_cells(self, tiles):
        """per-cell view of a per-tile mask"""
        t = tiles.repeat_interleave(8, 0).repeat_interleave(8, 1).repeat_interleave(8, 2)
        return t[: self.nx, : self.ny, : self.nz]

    def freeze(self, pos, vel, pres_prev, need_prev):
        """the discrete choices of one substep from its state"""
        cells = self.cells(pos)
        cnt = torch.zeros(self.shape, dtype=T64)
        cnt.index_put_((cells[:, 0], cells[:, 1], cells[:, 2]), torch.ones(len(pos), dtype=T64), accumulate=True)
        need = self.need(cells)
        kept = self.tile_cells(need & need_prev) if need_prev is not None else torch.zeros(self.shape, dtype=torch.bool)
        return dict(cnt=cnt, fluid=cnt > 0, need=need, kept=kept)

    # ---- stages ----
    def _stencil(self, pos, axis):
        g = (pos - self.x0) * self.ih
        sub = torch.full((3,), 0.5, dtype=T64)
        sub[axis] = 0.0
        g = g - sub
        f0 = torch.floor(g.detach())
        fr = g - f0
        i0 = f0.to(torch.int64)
        lim = [n - 1 for n in self.shape]
        lim[axis] += 1
        fshape = list(self.shape)
        fshape[axis] += 1
        out = []
        for c in range(8):
            d = [(c >> a) & 1 for a in range(3)]
            w = torch.ones(len(pos), dtype=T64)
            idx = []
            for a in range(3):
                w = w * (fr[:, a] if d[a] else 1.0 - fr[:, a])
                idx.append((i0[:, a] + d[a]).clamp(0, lim[a]))
            flat = (idx[0] * fshape[1] + idx[1]) * fshape[2] + idx[2]
            out.append((w, flat))
        return out, fshape

    def substep(self, pos, vel, pres_prev, fz, br=None):
        """One substep at the discrete state fz; br: frozen g2p branches (else
        taken here and returned)."""
        fluid, cnt = fz["fluid"], fz["cnt"]
        fl = fluid.to(T64)
        # p2g + normFaces
        VN, VO = [], []
        for a in range(3):
            st, fshape = self._stencil(pos, a)
            nf = fshape[0] * fshape[1] * fshape[2]
            V = torch.zeros(nf, dtype=T64)
            W = torch.zeros(nf, dtype=T64)
            for w, flat in st:
                V = V.index_add(0, flat, w * vel[:, a])
                W = W.index_add(0, flat, w)
            ok = (W.detach() * 65536.0 >= 0.5) if "wok" not in fz or fz["wok"][a] is None else fz["wok"][a]
            fz.setdefault("wok", [None, None, None])[a] = ok
            v = torch.where(ok, V / torch.where(ok, W, torch.ones_like(W)), torch.zeros_like(V)).view(fshape)
            old = v
            new = v - self.g * self.dt if a == 1 else v
            # faces on the box walls (own-axis index 0 and n) hold 0
            wall = torch.zeros(fshape, dtype=torch.bool)
            sl0 = [slice(None)] * 3
            sl1 = [slice(None)] * 3
            sl0[a] = 0
            sl1[a] = fshape[a] - 1
            wall[tuple(sl0)] = True
            wall[tuple(sl1)] = True
            VN.append(torch.where(wall, torch.zeros_like(new), new))
            VO.append(torch.where(wall, torch.zeros_like(old), old))
        # divergence
        div = ((VN[0][1:] - VN[0][:-1]) + (VN[1][:, 1:] - VN[1][:, :-1]) + (VN[2][:, :, 1:] - VN[2][:, :, :-1])) * self.ih
        air = ~fluid
        nb_air = torch.zeros(self.shape, dtype=torch.bool)
        nb_air[1:] |= air[:-1]
        nb_air[:-1] |= air[1:]
        nb_air[:, 1:] |= air[:, :-1]
        nb_air[:, :-1] |= air[:, 1:]
        nb_air[:, :, 1:] |= air[:, :, :-1]
        nb_air[:, :, :-1] |= air[:, :, 1:]
        dens = cnt / self.rest - 1.0
        tgt = self.comp * torch.where(nb_air, dens.clamp(min=0.0), dens)
        rhs = torch.where(fluid, (tgt - div) * self.h * self.h / self.dt, torch.zeros_like(div))
        # red-black SOR from the warm start
        p = pres_prev * (fz["kept"] & fluid).to(T64)
        act = fluid & (self.nns > 0)
        for _ in range(self.iters):
            for colour in (0, 1):
                q = p * fl
                s = torch.zeros_like(p)
                s[1:] = s[1:] + q[:-1]
                s[:-1] = s[:-1] + q[1:]
                s[:, 1:] = s[:, 1:] + q[:, :-1]
                s[:, :-1] = s[:, :-1] + q[:, 1:]
                s[:, :, 1:] = s[:, :, 1:] + q[:, :, :-1]
                s[:, :, :-1] = s[:, :, :-1] + q[:, :, 1:]
                pg = (rhs + s) / torch.where(act, self.nns, torch.ones_like(self.nns))
                upd = act & (self.parity == colour)
                p = torch.where(upd, p + self.omega * (pg - p), p)
        p = torch.where(act, p, torch.zeros_like(p))
        # project (interior faces with a fluid side)
        pf = p * fl
        k = self.dt * self.ih
        V0 = VN[0].clone()
        m = (fluid[1:] | fluid[:-1]).to(T64)
        V0[1:-1] = VN[0][1:-1] - k * m * (pf[1:] - pf[:-1])
        V1 = VN[1].clone()
        m = (fluid[:, 1:] | fluid[:, :-1]).to(T64)
        V1[:, 1:-1] = VN[1][:, 1:-1] - k * m * (pf[:, 1:] - pf[:, :-1])
        V2 = VN[2].clone()
        m = (fluid[:, :, 1:] | fluid[:, :, :-1]).to(T64)
        V2[:, :, 1:-1] = VN[2][:, :, 1:-1] - k * m * (pf[:, :, 1:] - pf[:, :, :-1])
        VN = [V0, V1, V2]
        # g2p
        vn, vo = [], []
        for a in range(3):
            st, fshape = self._stencil(pos, a)
            fn, fo = VN[a].reshape(-1), VO[a].reshape(-1)
            sn = sum(w * fn[flat] for w, flat in st)
            so = sum(w * fo[flat] for w, flat in st)
            vn.append(sn)
            vo.append(so)
        vn = torch.stack(vn, 1)
        vo = torch.stack(vo, 1)
        vm = vn + self.flip * (vel - vo)
        sp = torch.linalg.norm(vm, dim=1)
        br = {} if br is None else dict(br)
        capped = br["capped"] if "capped" in br else sp.detach() > self.vmax
        vc = torch.where(capped[:, None], vm * (self.vmax / sp)[:, None], vm)
        pt = pos + vc * self.dt
        lo = self.x0 + self.margin
        hi = self.x1 - self.margin
        keep = br["keep"] if "keep" in br else (pt.detach() > lo) & (pt.detach() < hi)
        low = br["low"] if "low" in br else pt.detach() <= lo
        pc = torch.where(keep, pt, torch.where(low, lo.expand_as(pt), hi.expand_as(pt)))
        if "nrm" not in br:
            dsp = (pc - pt).detach()
            dd = (dsp * dsp).sum(1)
            contact = dd > 1e-12
            nrm = torch.where(contact[:, None], dsp / dd.clamp(min=1e-300).sqrt()[:, None], torch.zeros_like(dsp))
            br["nrm"] = nrm
            br["wall"] = contact & ((vc.detach() * nrm).sum(1) < 0)
        br.update(capped=capped, keep=keep, low=low)
        nrm, wall = br["nrm"], br["wall"]
        vd = (vc * nrm).sum(1)
        vout = torch.where(wall[:, None], vc - nrm * vd[:, None], vc)
        return pc, vout, p, br

    def rollout(self, pos, vel, nsteps, frozen=None, clamps=None):
        """K substeps from an empty pool; returns final (pos, vel) and the frozen
        choices of every substep (pass them back to repeat the same branches).
        clamps: per substep, g2p's box-clamp branches to take (wall_branches)."""
        pres = torch.zeros(self.shape, dtype=T64)
        need_prev = None
        rec = [] if frozen is None else frozen
        for s in range(nsteps):
            if frozen is None:
                fz = self.freeze(pos, vel, pres, need_prev)
                pos, vel, pres, br = self.substep(pos, vel, pres, fz, clamps[s] if clamps else None)
                rec.append((fz, br))
                need_prev = fz["need"]
            else:
                fz, br = frozen[s]
                pos, vel, pres, _ = self.substep(pos, vel, pres, fz, br)
        return pos, vel, rec

    def wall_branches(self, dims, pos_after):
        """The box-clamp branches an f32 run took, read from its positions after
        the substep: a clamped coordinate equals the f32 wall bound exactly (the
        bound x0 + margin evaluated in f32 as the kernel does), a free one lies
        strictly inside.  The contact normal is the clamped axes' inward unit
        (both f32 sides of a corner weigh the same; corners are rare)."""
        m = np.float32(0.3 * float(dims.d))
        lo = np.array([np.float32(dims.x0) + m, np.float32(dims.y0) + m, np.float32(dims.z0) + m], np.float32)
        hi = np.array([np.float32(dims.x1) - m, np.float32(dims.y1) - m, np.float32(dims.z1) - m], np.float32)
        p = np.asarray(pos_after, np.float32)[:, :3]
        low = p == lo
        high = p == hi
        keep = ~(low | high)
        nrm = low.astype(np.float64) - high.astype(np.float64)
        nn = np.linalg.norm(nrm, axis=1)
        contact = nn > 0
        nrm[contact] /= nn[contact, None]
        return dict(keep=torch.from_numpy(keep), low=torch.from_numpy(low), nrm=torch.from_numpy(nrm),
                    wall=torch.from_numpy(contact))
