"""TEST INFRASTRUCTURE: FP64 norms of a substep's pressure solve and projection,
restated with numpy from canonical dumps (of the oracle or of the GPU sim).

Only tests/ use it, as the checker of st_solve_norms; never on the product path.
For every fluid cell (type 1), in f64 from the stored f32 (or f64) values:
  r   = rhs + sum of fluid-neighbour pressures - nns * p, the residual of the
        system relax() iterates on (src/flip/wgsl-solver.js:124-140; nns counts
        non-solid neighbours; out-of-box neighbours are walls by the open bits,
        ctypeOf at src/flip/wgsl-core.js:104-112);
  div = sum over axes of (v_hi - v_lo) / h of the projected faces (the
        divergence stencil, wgsl-solver.js:106-113);
  tgt = comp * (interior ? dens : max(dens, 0)), dens = cnt / rest - 1, 0 for
        deep narrow-band cells (wgsl-solver.js:114-121);
  e   = div - tgt.
The projection (wgsl-solver.js:165-180) makes e = -(dt / h^2) r in exact
arithmetic, which the tests use to pin this restatement.
"""
from __future__ import annotations

import numpy as np

TYPE_AIR, TYPE_FLUID, TYPE_SOLID = 0, 1, 2


def _dense(tiles, vals, dims, per):
    """canonical per-tile values -> array [k, j, i, per] over the whole tile range
    (zeros where a tile holds no storage)"""
    tx, ty, tz = int(dims.tx), int(dims.ty), int(dims.tz)
    out = np.zeros((tz * 8, ty * 8, tx * 8, per), dtype=np.float64 if vals.dtype.kind == "f" else np.int64)
    v = vals.reshape(len(tiles), 8, 8, 8, per)  # local index i + 8 (j + 8 k) -> [k, j, i]
    for n, t in enumerate(np.asarray(tiles, dtype=np.int64)):
        ti, tj, tk = t % tx, (t // tx) % ty, t // (tx * ty)
        out[tk * 8:tk * 8 + 8, tj * 8:tj * 8 + 8, ti * 8:ti * 8 + 8] = v[n]
    return out


def cell_terms(sim, comp: float, open_: int = 0, nb: bool = False):
    """Per fluid cell (r, div, tgt, e) from `sim`'s dumps (`open_` = the box's
    ST_OPEN_* bits); returns a dict of 1-D f64 arrays over the fluid cells (in
    [k, j, i] order)."""
    d = sim.dims
    nx, ny, nz = int(d.nx), int(d.ny), int(d.nz)
    tiles = sim.dump("tiles")
    ctype = _dense(tiles, sim.dump("ctype").astype(np.int64), d, 1)[..., 0]
    cnt = _dense(tiles, sim.dump("cnt").astype(np.int64), d, 1)[..., 0]
    pres = _dense(tiles, sim.dump("pres").astype(np.float64), d, 1)[..., 0]
    rhs = _dense(tiles, sim.dump("rhs").astype(np.float64), d, 1)[..., 0]
    face = _dense(tiles, sim.dump("facev").astype(np.float64).reshape(-1), d, 3)

    # cell types with a one-cell border of walls (open bit -> air, else solid)
    ty = np.zeros((nz + 2, ny + 2, nx + 2), np.int64)
    ty[1:-1, 1:-1, 1:-1] = ctype[:nz, :ny, :nx] & 0xFF
    wall = lambda bit: TYPE_AIR if (open_ & bit) else TYPE_SOLID  # noqa: E731
    ty[:, :, 0] = wall(1)
    ty[:, :, -1] = wall(2)
    ty[:, 0, 1:-1] = wall(4)
    ty[:, -1, 1:-1] = wall(8)
    ty[0, 1:-1, 1:-1] = wall(16)
    ty[-1, 1:-1, 1:-1] = wall(32)
    pp = np.zeros((nz + 2, ny + 2, nx + 2))
    pp[1:-1, 1:-1, 1:-1] = pres[:nz, :ny, :nx]

    c = (slice(1, -1), slice(1, -1), slice(1, -1))
    nbrs = [(slice(1, -1), slice(1, -1), slice(0, -2)), (slice(1, -1), slice(1, -1), slice(2, None)),
            (slice(1, -1), slice(0, -2), slice(1, -1)), (slice(1, -1), slice(2, None), slice(1, -1)),
            (slice(0, -2), slice(1, -1), slice(1, -1)), (slice(2, None), slice(1, -1), slice(1, -1))]
    fluid = ty[c] == TYPE_FLUID
    nns = np.zeros(fluid.shape, np.int64)
    air = np.zeros(fluid.shape, bool)
    s = np.zeros(fluid.shape)
    for q in nbrs:
        tq = ty[q]
        nns += tq != TYPE_SOLID
        air |= tq == TYPE_AIR
        s += np.where(tq == TYPE_FLUID, pp[q], 0.0)
    p = pres[:nz, :ny, :nx]
    r = np.where(nns == 0, 0.0, rhs[:nz, :ny, :nx] + s - nns * p)
    h = float(np.float32(d.h))
    div = ((face[:nz, :ny, 1:nx + 1, 0] - face[:nz, :ny, :nx, 0]) + (face[:nz, 1:ny + 1, :nx, 1] - face[:nz, :ny, :nx, 1]) +
           (face[1:nz + 1, :ny, :nx, 2] - face[:nz, :ny, :nx, 2])) / h
    rest = float(np.float32(d.rest))
    dens = cnt[:nz, :ny, :nx] / rest - 1.0
    tgt = float(np.float32(comp)) * np.where(air, np.maximum(dens, 0.0), dens)
    if nb:
        phi = 8.0 - (ctype[:nz, :ny, :nx] >> 16).astype(np.float32) / np.float32(4096.0)
        tgt = np.where(phi < -1.0, 0.0, tgt)
    f = fluid
    return {"r": r[f], "rhs": rhs[:nz, :ny, :nx][f], "div": div[f], "tgt": tgt[f], "e": (div - tgt)[f]}


def norms(sim, comp: float, open_: int = 0, nb: bool = False):
    """The st_norms fields from `sim`'s dumps (f64)."""
    t = cell_terms(sim, comp, open_, nb)
    mx = lambda a: float(np.max(np.abs(a))) if a.size else 0.0  # noqa: E731
    return {"res_l2": float(np.sqrt(np.sum(t["r"] ** 2))), "res_max": mx(t["r"]),
            "rhs_l2": float(np.sqrt(np.sum(t["rhs"] ** 2))),
            "div_l2": float(np.sqrt(np.sum(t["div"] ** 2))), "div_max": mx(t["div"]),
            "err_l2": float(np.sqrt(np.sum(t["e"] ** 2))), "err_max": mx(t["e"]),
            "fluid": float(t["r"].size)}
