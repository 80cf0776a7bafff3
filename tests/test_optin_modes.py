"""Opt-in modes with no counterpart in the reference (st_params.rk2,
st_params.extrap_layers), pinned on the CPU oracle by what the mathematics
fixes, and on the GPU against the oracle.  Off (0) is the reference's path; the
parity tests of that path run with both off.

RK2: in a rigid rotation field (linear, so the MAC grid's trilinear sampling
returns it exactly) one particle advected over one substep by the Euler step
and by the midpoint rule is compared with the exact circular motion: halving dt
divides the Euler error by about 4 (second order per step) and the midpoint
rule's by about 8 (third order), and the midpoint rule's error is far smaller.

Extrapolation: on a pool whose faces were overwritten by hand, the faces a layer
sets are exactly the breadth-first layers of air-air faces around the faces next
to the fluid (computed here on the dumped cell types); a constant field stays
constant; values stay within the range of the known ones; faces beyond the last
layer keep their values."""
import math

import numpy as np
import pytest

from splashtank_amd import capi, scenes
from oracle.oracle import Oracle
from tests.helpers import opts_for

STAGES_TO_PROJECT = [capi.ST_STAGE_BLOCKS, capi.ST_STAGE_CLEAR, capi.ST_STAGE_P2G, capi.ST_STAGE_CLASSIFY,
                     capi.ST_STAGE_NORMFACES, capi.ST_STAGE_DIVERGENCE, capi.ST_STAGE_PRESSURE, capi.ST_STAGE_PROJECT]


def _cells(o):
    """global (i, j, k) of every cell of the canonical dumps, shape (ncells, 3)"""
    d = o.dims
    tiles = o.dump("tiles").astype(np.int64)
    ti, tj, tk = tiles % d.tx, (tiles // d.tx) % d.ty, tiles // (d.tx * d.ty)
    l = np.arange(512)
    lx, ly, lz = l % 8, (l // 8) % 8, l // 64
    i = (ti[:, None] * 8 + lx[None, :]).reshape(-1)
    j = (tj[:, None] * 8 + ly[None, :]).reshape(-1)
    k = (tk[:, None] * 8 + lz[None, :]).reshape(-1)
    return np.stack([i, j, k], axis=1)


def _rotation_step(dt, rk2):
    """one particle at radius 0.2 about the box centre's vertical axis, in the
    field u = w x (p - c), w = (0, 1, 0); returns (position after g2p, exact)"""
    sc = scenes.dam_break(d=1.0 / 32)
    opts = opts_for(sc, max_n=4)
    o = Oracle(opts, fp64=True)
    d = o.dims
    c = np.array([0.0, 0.5, 0.0])
    p0 = c + np.array([0.2, 0.013, 0.07])
    o.upload(np.array([[*p0, 0.0]], np.float32), np.zeros((1, 4), np.float32))
    prm = capi.make_params(dt=dt, g=0.0, flip=0.0, iters=1, vmax=100.0, rk2=1 if rk2 else 0)
    o.begin_step(prm)
    for st in STAGES_TO_PROJECT:
        o.stage(st)
    ijk = _cells(o).astype(np.float64)
    h = float(d.h)
    base = np.array([d.x0, d.y0, d.z0], np.float64)
    f = np.zeros((ijk.shape[0], 3))
    # face of axis a at the cell's low side: coordinate a at i_a h, the others at the cell centre
    for a in range(3):
        pos = base + (ijk + 0.5) * h
        pos[:, a] = base[a] + ijk[:, a] * h
        u = np.stack([pos[:, 2] - c[2], np.zeros(len(pos)), -(pos[:, 0] - c[0])], axis=1)
        f[:, a] = u[:, a]
    o.poke("facev", f)
    o.poke("faceo", f)
    o.stage(capi.ST_STAGE_G2P)
    p1 = o.dump("pos")[0, :3].astype(np.float64)
    th = dt  # angle turned: |w| dt
    r = p0 - c
    exact = c + np.array([r[0] * math.cos(th) + r[2] * math.sin(th), r[1], -r[0] * math.sin(th) + r[2] * math.cos(th)])
    o.close()
    return p1, exact


def test_rk2_is_third_order_per_step_in_a_rotation():
    errs = {}
    for rk2 in (False, True):
        for dt in (0.02, 0.01):
            p1, ex = _rotation_step(dt, rk2)
            errs[(rk2, dt)] = float(np.linalg.norm(p1 - ex))
    e_ratio = errs[(False, 0.02)] / errs[(False, 0.01)]
    m_ratio = errs[(True, 0.02)] / errs[(True, 0.01)]
    print("errors", errs, "ratios", e_ratio, m_ratio)
    assert 3.5 < e_ratio < 4.5, "the Euler step's one-step error should scale as dt^2"
    assert 7.0 < m_ratio < 9.0, "the midpoint rule's one-step error should scale as dt^3"
    assert errs[(True, 0.01)] < errs[(False, 0.01)] / 20


def _pool_with_faces(layers, fp64=True):
    """a dam break stepped up to the projection with the extrapolation off, then
    every face overwritten: faces next to fluid (known) with a value, the others
    with a sentinel; returns the oracle and its cell table"""
    sc = scenes.dam_break(d=1.0 / 32, frac=(0.4, 0.5, 0.6))
    o = Oracle(opts_for(sc), fp64=fp64)
    o.upload(sc.pos, sc.vel)
    o.begin_step(capi.make_params(**dict(sc.params, extrap_layers=layers)))
    for st in STAGES_TO_PROJECT:
        o.stage(st)
    return o


def _known_and_settable(o):
    """per face (cells x 3): known (fluid on one side) and settable (air on both
    sides), from the dumped cell types; a cell not in the dumps counts as air
    inside the box and as a wall outside it (closed box)"""
    d = o.dims
    ijk = _cells(o)
    ty = (o.dump("ctype") & 0xFF).astype(np.int64)
    lut = {tuple(x): t for x, t in zip(map(tuple, ijk), ty)}

    def tof(i, j, k):
        if i < 0 or j < 0 or k < 0 or i >= d.nx or j >= d.ny or k >= d.nz:
            return 2
        return lut.get((i, j, k), 0)

    known = np.zeros((len(ijk), 3), bool)
    settable = np.zeros((len(ijk), 3), bool)
    for n, (i, j, k) in enumerate(ijk):
        th = tof(i, j, k)
        for a in range(3):
            lo = [i, j, k]
            lo[a] -= 1
            tl = tof(*lo)
            if tl == 1 or th == 1:
                known[n, a] = True
            elif tl == 0 and th == 0:
                settable[n, a] = True
    return ijk, known, settable


def _bfs_layers(ijk, known, settable, layers):
    """layer (1..layers) at which each settable face gets its value, 0 if never"""
    index = {tuple(x): n for n, x in enumerate(map(tuple, ijk))}
    got = np.zeros(known.shape, np.int64)
    cur = known.copy()
    for layer in range(1, layers + 1):
        nxt = cur.copy()
        for n, (i, j, k) in enumerate(ijk):
            for a in range(3):
                if cur[n, a] or not settable[n, a]:
                    continue
                for q in range(6):
                    r = [i, j, k]
                    r[q // 2] += 1 if q % 2 else -1
                    m = index.get(tuple(r))
                    if m is not None and cur[m, a]:
                        nxt[n, a] = True
                        got[n, a] = layer
                        break
        cur = nxt
    return got


@pytest.mark.parametrize("layers", [1, 2])
def test_extrapolation_layers_constant_and_bounds(layers):
    o = _pool_with_faces(layers)
    ijk, known, settable = _known_and_settable(o)
    got = _bfs_layers(ijk, known, settable, layers)
    sentinel = 12345.0
    # a constant field on the known faces: every face a layer reaches takes it
    f = np.where(known, 0.625, sentinel)
    o.poke("facev", f)
    o.poke("faceo", -f)
    o.kernel("extrapolate")
    fv, fo = o.dump("facev"), o.dump("faceo")
    reached = got > 0
    assert reached.any()
    assert np.all(fv[reached] == 0.625) and np.all(fo[reached] == -0.625)
    assert np.all(fv[known] == 0.625)
    assert np.all(fv[~known & ~reached] == sentinel), "faces beyond the last layer (or not settable) must keep their values"
    # random known values: what a layer sets lies within their range
    rng = np.random.default_rng(7)
    vals = rng.uniform(-2.0, 3.0, known.shape)
    o2 = _pool_with_faces(layers)
    o2.poke("facev", np.where(known, vals, sentinel))
    o2.poke("faceo", np.where(known, -vals, sentinel))
    o2.kernel("extrapolate")
    fv2 = o2.dump("facev")
    lo, hi = vals[known].min(), vals[known].max()
    assert np.all((fv2[reached] >= lo) & (fv2[reached] <= hi))
    assert np.array_equal(fv2 != sentinel, known | reached)
    o.close()
    o2.close()


def test_extrapolation_off_changes_nothing():
    o = _pool_with_faces(0)
    ijk, known, settable = _known_and_settable(o)
    f = np.where(known, 0.5, 7.0)
    o.poke("facev", f)
    o.kernel("extrapolate")
    assert np.array_equal(o.dump("facev"), f)
    o.close()


@pytest.mark.gpu
def test_optin_modes_gpu_match_oracle():
    """the GPU with rk2 and two extrapolation layers on, stage by stage against
    the oracle (the dam break with a moving box collider)"""
    from tests.test_gpu_parity import run_stagewise
    sc = scenes.sparse_splash(d=1.0 / 64, pool=0.25, sphere_r=0.12, box_half=0.1, iters=20, dt=1.0 / 240)
    sc.params.update(rk2=1, extrap_layers=2)
    run_stagewise(sc, 3)
