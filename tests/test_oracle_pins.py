"""Pins of the CPU oracle (oracle/flip_oracle.cpp) by what the mathematics fixes.

The reference ships no golden vectors or known-answer tests for the FLIP path
(its only assertions are qualitative page smoke checks), so each function of the
oracle is pinned here by a closed form, an invariant, brute force on a tiny
input, or a textbook solve -- each chosen so a plausible slip (a dropped term, a
wrong sign, an axis mix-up, an off-by-one in a window) fails it.
"""
import numpy as np
import pytest

from splashtank_amd import capi, scenes
from oracle.oracle import Oracle, expf

pytestmark = []


# ---------------------------------------------------------------- helpers
def lattice(d, lo, hi, vel=(0.0, 0.0, 0.0), velfn=None):
    """Particles at lo + (k + 1/2) d (no jitter) inside [lo, hi)."""
    axes = [np.arange(lo[a] + d / 2, hi[a] - 1e-12, d) for a in range(3)]
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    n = X.size
    pos = np.zeros((n, 4), np.float32)
    pos[:, 0], pos[:, 1], pos[:, 2] = X.ravel(), Y.ravel(), Z.ravel()
    v = np.zeros((n, 4), np.float32)
    if velfn is not None:
        vv = velfn(pos[:, :3].astype(np.float64))
        v[:, :3] = vv
    else:
        v[:, :3] = vel
    return pos, v


def make(d, size, pos, vel, fp64=False, open_=0, ratio=2.0):
    o = Oracle(capi.make_opts(d, size, ratio=ratio, open_=open_, max_n=max(len(pos), 1)), fp64=fp64)
    o.upload(pos, vel)
    return o


def dense(sim, field):
    """Canonical per-tile dump -> dense array [k, j, i, (component)] over the tile grid."""
    d = sim.dims
    tiles = sim.dump("tiles")
    a = sim.dump(field)
    per = a.shape[1] if a.ndim == 2 else 1
    fill = 0 if a.dtype.kind in "iu" else np.nan
    out = np.full((d.tz * 8, d.ty * 8, d.tx * 8, per), fill, dtype=a.dtype)
    blocks = a.reshape(len(tiles), 8, 8, 8, per)
    for q, t in enumerate(tiles):
        ti, tj, tk = t % d.tx, (t // d.tx) % d.ty, t // (d.tx * d.ty)
        out[tk * 8:tk * 8 + 8, tj * 8:tj * 8 + 8, ti * 8:ti * 8 + 8] = blocks[q]
    return out if per > 1 else out[..., 0]


def run_to(o, prm, stage):
    o.begin_step(prm)
    for k in range(stage + 1):
        o.stage(k)


# ---------------------------------------------------------------- exp
def test_expf_error_bound():
    xs = np.concatenate([np.linspace(-87.0, 88.0, 20001), np.linspace(-1, 1, 4001), [-0.0, 0.0, 1e-8, -1e-8]])
    xs = xs.astype(np.float32)
    got = np.array([expf(float(x)) for x in xs], np.float32)
    ref = np.exp(xs.astype(np.float64))
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    err = np.abs(got.astype(np.float64) - ref) / ulp
    assert err.max() <= 2.0, err.max()
    assert expf(0.0) == 1.0 and expf(-0.0) == 1.0
    assert expf(-200.0) == 0.0 and np.isinf(expf(100.0))


# ---------------------------------------------------------------- blocks
def test_active_tiles_brute_force():
    rng = np.random.default_rng(7)
    d = 1 / 48
    size = (1.0, 0.75, 0.5)
    n = 400
    pos = np.zeros((n, 4), np.float32)
    pos[:, 0] = rng.uniform(-0.5, 0.5, n)
    pos[:, 1] = rng.uniform(0, 0.75, n)
    pos[:, 2] = rng.uniform(-0.25, 0.25, n)
    pos[:10, :3] = [[-0.5, 0, -0.25], [0.4999, 0.7499, 0.2499], [0, 0, 0], [-0.49, 0.01, 0.24], [0.3, 0.7, -0.2],
                    [0.5, 0.75, 0.25], [-0.5, 0.75, 0.25], [0.5, 0, -0.25], [0.0, 0.375, 0.0], [0.1, 0.2, 0.0]]
    pos[10, 1] = -1e6  # dead: marks nothing
    o = make(d, size, pos, np.zeros_like(pos))
    run_to(o, capi.make_params(dt=1 / 240), capi.ST_STAGE_BLOCKS)
    dm = o.dims
    # brute force, in f64 with cells clamped; particles within 1e-4 cell of a cell
    # boundary are kept away by construction of the random draw (probability ~0)
    x0 = np.array([dm.x0, dm.y0, dm.z0], np.float64)
    cell = np.floor((pos[:, :3].astype(np.float64) - x0) / dm.h).astype(int)
    nmax = np.array([dm.nx, dm.ny, dm.nz]) - 1
    cell = np.clip(cell, 0, nmax)
    tmax = np.array([dm.tx, dm.ty, dm.tz]) * 8 - 1
    want = set()
    for i in range(n):
        if pos[i, 1] < -1e5:
            continue
        lo = np.maximum(cell[i] - 2, 0) // 8
        hi = np.minimum(cell[i] + 2, tmax) // 8
        for tk in range(lo[2], hi[2] + 1):
            for tj in range(lo[1], hi[1] + 1):
                for ti in range(lo[0], hi[0] + 1):
                    want.add(ti + dm.tx * (tj + dm.ty * tk))
    got = o.dump("tiles")
    assert sorted(want) == list(got)
    ctl = o.ctl()
    assert ctl[4] == len(want) and ctl[3] == len(want) and ctl[5] == 0
    assert ctl[2] == dm.scap - 1 - len(want)


def test_tiles_freed_and_reused():
    d = 1 / 32
    pos, vel = lattice(d, (-0.5, 0.0, -0.5), (-0.3, 0.2, -0.3))
    o = make(d, (1, 1, 1), pos, vel)
    prm = capi.make_params(dt=1 / 240)
    run_to(o, prm, capi.ST_STAGE_BLOCKS)
    first = set(o.dump("tiles"))
    # move every particle to the far corner: old tiles freed, new ones allocated from the stack
    pos2 = pos.copy()
    pos2[:, 0] += 0.75
    pos2[:, 2] += 0.75
    o.upload(pos2, vel)
    run_to(o, prm, capi.ST_STAGE_BLOCKS)
    second = set(o.dump("tiles"))
    assert first and second and not (first & second)
    ctl = o.ctl()
    assert ctl[4] == len(second) and ctl[2] == o.dims.scap - 1 - len(second)
    tmap = o.dump("tmap")
    slots = tmap[tmap != 0xFFFFFFFF]
    assert len(set(slots)) == len(slots) and slots.min() >= 1 and slots.max() < o.dims.scap


# ---------------------------------------------------------------- p2g
def test_p2g_particle_at_face_centre():
    d = 1 / 32  # h = 1/16: coordinates and 1/h exact in binary
    h = 2 * d
    i, j, k = 5, 6, 7
    x = -0.5 + i * h
    y = 0.0 + (j + 0.5) * h
    z = -0.5 + (k + 0.5) * h
    pos = np.array([[x, y, z, 0]], np.float32)
    vel = np.array([[1.0, 2.0, -0.5, 0]], np.float32)
    o = make(d, (1, 1, 1), pos, vel)
    run_to(o, capi.make_params(dt=1 / 240), capi.ST_STAGE_P2G)
    aw = dense(o, "accw")
    av = dense(o, "accv")
    cnt = dense(o, "cnt")
    # x faces: the particle sits on face (i, j, k) exactly
    nz = np.argwhere(aw[..., 0] != 0)
    assert nz.tolist() == [[k, j, i]]
    assert aw[k, j, i, 0] == 65536 and av[k, j, i, 0] == 65536
    # y faces sit at (i'+1/2, j', k'+1/2): in face units the particle is at
    # (i-1/2, j+1/2, k) -> 1/4 on x in {i-1, i}, y in {j, j+1}, z = k
    ny = sorted(map(tuple, np.argwhere(aw[..., 1] != 0).tolist()))
    assert ny == sorted([(k, jj, ii) for jj in (j, j + 1) for ii in (i - 1, i)])
    assert all(aw[kk, jj, ii, 1] == 16384 and av[kk, jj, ii, 1] == 32768 for kk, jj, ii in ny)
    # z faces sit at (i'+1/2, j'+1/2, k'): the particle is at (i-1/2, j, k+1/2)
    # -> 1/4 on x in {i-1, i}, y = j, z in {k, k+1}
    nzz = sorted(map(tuple, np.argwhere(aw[..., 2] != 0).tolist()))
    assert nzz == sorted([(kk, j, ii) for kk in (k, k + 1) for ii in (i - 1, i)])
    assert all(aw[kk, jj, ii, 2] == 16384 and av[kk, jj, ii, 2] == -8192 for kk, jj, ii in nzz)
    assert cnt.sum() == 1 and cnt[k, j, i] == 1
    assert np.frombuffer(o.ctl()[6:7].tobytes(), np.float32)[0] == np.float32(np.sqrt(1 + 4 + 0.25))


def test_p2g_weight_partition_and_counts():
    rng = np.random.default_rng(3)
    d = 1 / 64
    n = 3000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-0.3, 0.3, (n, 3))
    pos[:, 1] += 0.5
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 1, (n, 3))
    o = make(d, (1, 1, 1), pos, vel)
    run_to(o, capi.make_params(dt=1 / 240), capi.ST_STAGE_P2G)
    aw = o.dump("accw").astype(np.int64)
    for a in range(3):
        tot = aw[:, a].sum()
        # each particle's 8 rounded weights sum to 65536 within +-4
        assert abs(tot - 65536 * n) <= 4 * n
        # and a dropped corner would lose ~1/8 of the total
    assert o.dump("cnt").sum() == n


def test_uniform_velocity_normalises_to_the_velocity():
    d = 1 / 32
    v = (0.75, -0.375, 1.25)
    pos, vel = lattice(d, (-0.25, 0.25, -0.25), (0.25, 0.75, 0.25), vel=v)
    o = make(d, (1, 1, 1), pos, vel)
    g, dt = 9.81, 1 / 240
    run_to(o, capi.make_params(dt=dt, g=g), capi.ST_STAGE_NORMFACES)
    aw0 = None
    # weights are gone after normFaces; recompute which faces had >= one particle of weight
    o2 = make(d, (1, 1, 1), pos, vel)
    run_to(o2, capi.make_params(dt=dt, g=g), capi.ST_STAGE_P2G)
    aw0 = o2.dump("accw")
    fv, fo = o.dump("facev"), o.dump("faceo")
    for a in range(3):
        m = aw0[:, a] >= 65536
        assert m.sum() > 100
        np.testing.assert_allclose(fo[m, a], v[a], atol=2e-3)
        np.testing.assert_allclose(fv[m, a], v[a] - (g * dt if a == 1 else 0.0), atol=2e-3)
        # faces no particle reached are zero (y: minus g dt)
        z = aw0[:, a] == 0
        np.testing.assert_array_equal(fo[z, a], 0.0)


def test_divergence_of_linear_field():
    d = 1 / 64
    h = 2 * d
    A = np.array([0.5, -1.25, 2.0])
    pos, vel = lattice(d, (-0.25, 0.25, -0.25), (0.25, 0.75, 0.25), velfn=lambda x: x * A)
    o = make(d, (1, 1, 1), pos, vel, fp64=True)
    dt = 1 / 240
    run_to(o, capi.make_params(dt=dt, comp=6.0), capi.ST_STAGE_DIVERGENCE)
    rhs = dense(o, "rhs")
    ct = dense(o, "ctype")
    dm = o.dims
    # interior: cells at least 2 cells inside the fluid block
    lo = np.round((np.array([-0.25, 0.25, -0.25]) - [dm.x0, dm.y0, dm.z0]) / h).astype(int) + 2
    hi = np.round((np.array([0.25, 0.75, 0.25]) - [dm.x0, dm.y0, dm.z0]) / h).astype(int) - 2
    sub = rhs[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
    assert ((ct[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] & 0xFF) == 1).all()
    want = -(A.sum()) * h * h / dt  # tgt = comp * (8/8 - 1) = 0
    np.testing.assert_allclose(sub, want, rtol=2e-4)


# ---------------------------------------------------------------- classify
def test_classify_truth_table():
    d = 1 / 32
    h = 2 * d
    # a dense block, a sparse cell touching it, a sparse cell far away
    dense_p, _ = lattice(d, (-0.25, 0.25, -0.25), (0.0, 0.5, 0.0))
    lone = np.array([[0.27, 0.70, 0.27, 0], [0.28, 0.71, 0.28, 0]], np.float32)  # one cell, far from the block
    touch = np.array([[0.0 + 0.25 * h, 0.25 + 0.25 * h, -0.25 + 0.25 * h, 0]], np.float32)
    pos = np.concatenate([dense_p, lone, touch])
    o = make(d, (1, 1, 1), pos, np.zeros_like(pos))
    run_to(o, capi.make_params(dt=1 / 240, tear=3.0), capi.ST_STAGE_CLASSIFY)
    ct = dense(o, "ctype") & 0xFF
    cnt = dense(o, "cnt")
    dm = o.dims

    def cell(p):
        return tuple(np.floor((np.array(p[:3], np.float64) - [dm.x0, dm.y0, dm.z0]) / h).astype(int)[::-1])

    assert cnt[cell(lone[0])] == 2 and ct[cell(lone[0])] == 0   # sparse, no dense neighbour: torn off
    assert cnt[cell(touch[0])] == 1 and ct[cell(touch[0])] == 1  # sparse next to full cells: water
    assert ct[cell(dense_p[0])] == 1
    # cells past the far wall are never fluid, empty cells are air
    assert (ct[:, :, dm.nx:] == 0).all() and ((ct == 1) <= (cnt > 0)).all()


def test_collider_cells_match_analytic_sphere():
    d = 1 / 32
    h = 2 * d
    r, c = 0.2, np.array([0.05, 0.45, -0.03])
    dims, lo, data = scenes.sphere_sdf(r, h / 2, 3 * h)
    cf = scenes.collider_record(0, dims, lo, h / 2, c)
    pos, vel = lattice(d, (-0.3, 0.2, -0.3), (0.4, 0.7, 0.3))
    o = make(d, (1, 1, 1), pos, vel)
    o.set_colliders(cf, data, 1)
    run_to(o, capi.make_params(dt=1 / 240), capi.ST_STAGE_CLASSIFY)
    ct = dense(o, "ctype")
    tiles = o.dump("tiles")
    dm = o.dims
    k, j, i = np.meshgrid(np.arange(dm.tz * 8), np.arange(dm.ty * 8), np.arange(dm.tx * 8), indexing="ij")
    ctr = np.stack([dm.x0 + (i + 0.5) * h, dm.y0 + (j + 0.5) * h, dm.z0 + (k + 0.5) * h], -1)
    phi = np.linalg.norm(ctr - c, axis=-1) - r
    have = np.zeros(ct.shape, bool)
    for t in tiles:
        ti, tj, tk = t % dm.tx, (t // dm.tx) % dm.ty, t // (dm.tx * dm.ty)
        have[tk * 8:tk * 8 + 8, tj * 8:tj * 8 + 8, ti * 8:ti * 8 + 8] = True
    q0 = 0.3 * h
    sure_in = have & (phi < q0 - 0.05 * h)
    sure_out = have & (phi > q0 + 0.05 * h)
    assert sure_in.sum() > 50
    assert ((ct[sure_in] >> 8) & 0xFF == 1).all()
    assert ((ct[sure_out] >> 8) & 0xFF == 0).all()
    inside = sure_in & (ctr[..., 0] < dm.x1 - 1e-9)
    assert ((ct[inside] & 0xFF) == 2).all()


# ---------------------------------------------------------------- pressure and projection
def _poisson_scene(fp64=True, iters=3000, comp=0.0):
    d = 1 / 32
    rng = np.random.default_rng(5)
    pos, vel = lattice(d, (-0.5, 0.0, -0.5), (-0.05, 0.55, 0.5))
    vel[:, :3] = rng.normal(0, 0.3, (len(pos), 3))
    o = make(d, (1, 1, 1), pos, vel, fp64=fp64)
    prm = capi.make_params(dt=1 / 240, iters=iters, omega=1.7, comp=comp)
    return o, prm


def _types_with_walls(o):
    dm = o.dims
    ct = dense(o, "ctype") & 0xFF
    full = np.full((dm.nz + 2, dm.ny + 2, dm.nx + 2), 2, np.int64)  # closed walls are solid
    full[1:-1, 1:-1, 1:-1] = ct[:dm.nz, :dm.ny, :dm.nx]
    return full


def test_sor_converges_to_direct_solve():
    sp = pytest.importorskip("scipy.sparse")
    spla = pytest.importorskip("scipy.sparse.linalg")
    o, prm = _poisson_scene(iters=4000)
    run_to(o, prm, capi.ST_STAGE_PRESSURE)
    T = _types_with_walls(o)
    rhs = dense(o, "rhs")
    pres = dense(o, "pres")
    dm = o.dims
    fl = np.argwhere(T == 1)
    idx = -np.ones(T.shape, np.int64)
    idx[tuple(fl.T)] = np.arange(len(fl))
    rows, cols, vals = [], [], []
    b = np.zeros(len(fl))
    # textbook 7-point Poisson: air neighbours Dirichlet p = 0, solid neighbours Neumann
    for r, (k, j, i) in enumerate(fl):
        diag = 0
        for dk, dj, di in ((0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)):
            tn = T[k + dk, j + dj, i + di]
            if tn == 2:
                continue
            diag += 1
            if tn == 1:
                rows.append(r); cols.append(idx[k + dk, j + dj, i + di]); vals.append(-1.0)
        rows.append(r); cols.append(r); vals.append(float(diag))
        b[r] = rhs[k - 1, j - 1, i - 1]
    Am = sp.csr_matrix((vals, (rows, cols)), shape=(len(fl), len(fl)))
    x = spla.spsolve(Am, b)
    got = pres[fl[:, 0] - 1, fl[:, 1] - 1, fl[:, 2] - 1]
    assert len(fl) > 1000
    np.testing.assert_allclose(got, x, rtol=1e-7, atol=1e-9 * np.abs(x).max())


def test_projection_is_divergence_free():
    o, prm = _poisson_scene(iters=4000)
    run_to(o, prm, capi.ST_STAGE_DIVERGENCE)
    dm = o.dims
    h = float(dm.h)
    u0 = dense(o, "facev")
    o.stage(capi.ST_STAGE_PRESSURE)
    o.stage(capi.ST_STAGE_PROJECT)
    u = dense(o, "facev")
    ct = dense(o, "ctype") & 0xFF

    def div(U):
        return ((U[:-1, :-1, 1:, 0] - U[:-1, :-1, :-1, 0]) + (U[:-1, 1:, :-1, 1] - U[:-1, :-1, :-1, 1]) +
                (U[1:, :-1, :-1, 2] - U[:-1, :-1, :-1, 2])) / h

    fl = ct[:-1, :-1, :-1] == 1
    d0 = np.abs(div(u0)[fl]).max()
    d1 = np.abs(div(u)[fl]).max()
    assert d0 > 1.0
    assert d1 < 1e-6 * d0, (d0, d1)


def test_sor_residual_decreases():
    o, prm = _poisson_scene(fp64=True, iters=1)
    run_to(o, prm, capi.ST_STAGE_DIVERGENCE)
    T = _types_with_walls(o)
    rhs = dense(o, "rhs")
    res = []
    for _ in range(30):
        o.stage(capi.ST_STAGE_PRESSURE)
        p = dense(o, "pres")
        P = np.zeros(T.shape)
        P[1:-1, 1:-1, 1:-1] = np.nan_to_num(p[:o.dims.nz, :o.dims.ny, :o.dims.nx])
        P[T != 1] = 0.0
        r = 0.0
        fl = np.argwhere(T == 1)
        k, j, i = fl.T
        nns = np.zeros(len(fl))
        s = np.zeros(len(fl))
        for dk, dj, di in ((0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)):
            tn = T[k + dk, j + dj, i + di]
            nns += tn != 2
            s += np.where(tn == 1, P[k + dk, j + dj, i + di], 0.0)
        r = np.linalg.norm(nns * P[k, j, i] - s - rhs[k - 1, j - 1, i - 1])
        res.append(r)
    assert all(b < a for a, b in zip(res, res[1:]))
    assert res[-1] < 0.5 * res[0]


# ---------------------------------------------------------------- g2p
def test_pic_uniform_flow_advects_exactly():
    d = 1 / 32
    dt = 1 / 64
    pos, vel = lattice(d, (-0.5, 0.0, -0.5), (0.5, 1.0, 0.5), vel=(1.0, 0.0, 0.0))
    o = make(d, (1, 1, 1), pos, vel, open_=capi.ST_OPEN_X0 | capi.ST_OPEN_X1)
    o.step(capi.make_params(dt=dt, g=0.0, flip=0.0, comp=0.0, iters=20))
    p1 = o.dump("pos")
    v1 = o.dump("vel")
    alive = p1[:, 1] > -1e5
    left = pos[:, 0] + dt > 0.5
    np.testing.assert_array_equal(alive, ~left)
    np.testing.assert_allclose(p1[alive, 0], pos[alive, 0] + dt, atol=2e-7)
    np.testing.assert_array_equal(p1[alive, 1:3], pos[alive, 1:3])
    np.testing.assert_allclose(v1[alive, :3], np.broadcast_to([1.0, 0.0, 0.0], v1[alive, :3].shape), atol=1e-6)
    assert (p1[~alive, 1] == -1e6).all()


def test_box_clamp_and_wall_normal_removed():
    d = 1 / 32
    pos = np.array([[0.0, 0.02, 0.0, 0]], np.float32)
    vel = np.array([[0.3, -5.0, 0.0, 0]], np.float32)
    o = make(d, (1, 1, 1), pos, vel)
    o.step(capi.make_params(dt=1 / 60, g=0.0, flip=1.0, iters=5))
    p = o.dump("pos")[0]
    v = o.dump("vel")[0]
    margin = np.float32(0.3 * d)
    assert p[1] == np.float32(0.0) + margin   # clamped to the floor inset by the margin
    assert v[1] >= 0.0                          # the inward normal velocity removed


def test_fp32_and_fp64_oracles_agree():
    sc = scenes.dam_break(d=1 / 32)
    a = make(sc.d, sc.size, sc.pos, sc.vel)
    b = make(sc.d, sc.size, sc.pos, sc.vel, fp64=True)
    prm = capi.make_params(**sc.params)
    for _ in range(5):
        a.step(prm)
        b.step(prm)
    pa, pb = a.dump("pos"), b.dump("pos")
    assert np.abs(pa[:, :3] - pb[:, :3]).max() < 1e-4
