"""Pins of the oracle's narrow-band functions (src/flip/wgsl-narrowband.js and the
helpers of wgsl-core.js:155-238) by what the mathematics fixes, on the CPU.

Each kernel is run alone on a state the test sets (Oracle.poke / Oracle.kernel)
and compared with a closed form that a plausible slip would fail:
- phiPut/phiOf: the 16-bit quantisation round trip and its clamps;
- phiInitAll: the analytic box and sphere distances at cell centres, the
  surface clamp to -1.5, 8 outside the domain, velG zeroed, and the
  keep-the-deeper rule of a second fill;
- the depth passes (nbJacA/nbJacB): a planar surface redistances one cell per
  pass, to the exact depth -(d + 1/2) once converged;
- nbPhi: a linear depth field carried by a uniform kept velocity is shifted
  exactly, clamped at the wall;
- nbVel: particle-less faces between two fluid cells, and only those, take the
  uniform kept velocity;
- saveVel: an identity of the projected faces;
- nbKill / nbDeficit / nbSpawn: the killed set, the deficient cells and their
  needs by brute force, and every spawned particle inside its cell at the
  positions of the published integer hash (lowbias32).
"""
import numpy as np

from splashtank_amd import capi
from oracle.oracle import Oracle
from tests.test_oracle_pins import dense, lattice


def sparse(sim, a):
    """Dense [k, j, i, (comp)] over the tile grid -> canonical per-tile order (inverse of dense)."""
    d = sim.dims
    tiles = sim.dump("tiles")
    per = a.shape[3] if a.ndim == 4 else 1
    a = a.reshape(d.tz * 8, d.ty * 8, d.tx * 8, per)
    out = np.zeros((len(tiles), 8, 8, 8, per), a.dtype)
    for q, t in enumerate(tiles):
        ti, tj, tk = t % d.tx, (t // d.tx) % d.ty, t // (d.tx * d.ty)
        out[q] = a[tk * 8:tk * 8 + 8, tj * 8:tj * 8 + 8, ti * 8:ti * 8 + 8]
    return out.reshape(-1, per) if per > 1 else out.reshape(-1)


def quant(phi):
    """phiPut's 16-bit code of a depth in cells: u32(clamp(round((8 - phi) 4096), 0, 65535))."""
    return np.clip(np.rint((8.0 - np.asarray(phi, np.float64)) * 4096.0), 0, 65535).astype(np.uint32)


def phi_of(ctype):
    return 8.0 - (np.asarray(ctype, np.uint32) >> 16).astype(np.float64) / 4096.0


def with_phi(ctype, phi):
    return (np.asarray(ctype, np.uint32) & 0xFFFF) | (quant(phi) << 16)


def centres(sim):
    """cell centres (x, y, z) and the in-domain mask over the dense tile grid"""
    d = sim.dims
    k, j, i = np.meshgrid(np.arange(d.tz * 8), np.arange(d.ty * 8), np.arange(d.tx * 8), indexing="ij")
    x = d.x0 + (i + 0.5) * d.h
    y = d.y0 + (j + 0.5) * d.h
    z = d.z0 + (k + 0.5) * d.h
    dom = (i < d.nx) & (j < d.ny) & (k < d.nz)
    return np.stack([x, y, z], -1), dom, (i, j, k)


def start(d, size, pos, vel, upto, **prm):
    """An oracle with the narrow band on, stepped through stage `upto` of a begun step."""
    o = Oracle(capi.make_opts(d, size, max_n=max(len(pos), 1)))
    o.upload(pos, vel)
    p = capi.make_params(**dict(dict(nb_on=1, dt=1 / 60, g=9.81, iters=10), **prm))
    o.begin_step(p)
    for s in range(upto + 1):
        o.stage(s)
    return o


def box_sdf(x, lo, hi):
    q = np.maximum(lo - x, x - hi)
    return np.linalg.norm(np.maximum(q, 0.0), axis=-1) + np.minimum(q.max(-1), 0.0)


def storage(o):
    """mask of the cells of tiles with storage, over the dense tile grid"""
    dd = o.dims
    has = np.zeros((dd.tz * 8, dd.ty * 8, dd.tx * 8), bool)
    for t in o.dump("tiles"):
        ti, tj, tk = t % dd.tx, (t // dd.tx) % dd.ty, t // (dd.tx * dd.ty)
        has[tk * 8:tk * 8 + 8, tj * 8:tj * 8 + 8, ti * 8:ti * 8 + 8] = True
    return has


def test_phi_init_matches_analytic_box_and_sphere():
    d = 1 / 32  # h = 1/16, 16^3 cells
    h = 2 * d
    box_lo, box_hi = np.array([-0.3, 0.0, -0.5]), np.array([0.1, 0.45, 0.5])
    sc, sr = np.array([0.25, 0.55, 0.0]), 0.15
    fills = [1.0, *box_lo, *box_hi, 0.0, 2.0, *sc, sr, 0, 0, 0]
    pos, vel = lattice(d, (-0.3, 0.0, -0.5), (0.1, 0.45, 0.5))
    o = start(d, (1, 1, 1), pos, vel, 1, nb_init=1, nfill=2, fills=fills)
    x, dom, _ = centres(o)
    ana = np.minimum(box_sdf(x, box_lo, box_hi), np.linalg.norm(x - sc, axis=-1) - sr) / h
    ph = np.clip(ana, -7.9, 7.9)
    ph = np.where(ph < 0, np.minimum(ph, -1.5), ph)
    ph = np.where(dom, ph, 8.0)
    want = quant(ph)
    has = storage(o)
    got = dense(o, "ctype") >> 16
    diff = np.abs(got[has].astype(np.int64) - want[has].astype(np.int64))
    # f32 against f64 distances: at most one code step apart, and most equal
    assert diff.max() <= 1, diff.max()
    assert np.mean(diff == 0) > 0.9
    # water cells are at least 1.5 deep, cells past the far walls hold 8
    assert np.all(8.0 - got[has & dom & (ana < 0)] / 4096.0 <= -1.5)
    assert np.all(got[has & ~dom] == 0)
    # phiInitAll of a restart zeroes the kept velocity
    o.poke("velg", np.full(o.dump("velg").shape, 0.7, np.float32))
    o.kernel("phiInitAll")
    assert np.all(o.dump("velg") == 0.0)
    assert np.array_equal(dense(o, "ctype") >> 16, got)
    # a second fill (nbInit 2) keeps the deeper of old and new
    o2 = start(d, (1, 1, 1), pos, vel, 1, nb_init=2, nfill=1, fills=[2.0, *sc, sr, 0, 0, 0])
    ct = dense(o2, "ctype")
    old = np.where(x[..., 1] < 0.2, -5.0, 6.0)
    o2.poke("ctype", sparse(o2, with_phi(ct, old)))
    o2.kernel("phiInitAll")
    sph = np.clip((np.linalg.norm(x - sc, axis=-1) - sr) / h, -7.9, 7.9)
    sph = np.where(sph < 0, np.minimum(sph, -1.5), sph)
    exp2 = np.minimum(np.where(dom, sph, 8.0), 8.0 - quant(old) / 4096.0)  # the old depth is kept past the walls too
    got2 = dense(o2, "ctype") >> 16
    has2 = storage(o2)
    d2 = np.abs(got2[has2].astype(np.int64) - quant(exp2)[has2].astype(np.int64))
    assert d2.max() <= 1


def _planar(J=7, d=1 / 32):
    pos, vel = lattice(d, (-0.5, 0.0, -0.5), (0.5, J * 2 * d, 0.5))
    o = start(d, (1, 1, 1), pos, vel, 3)
    ct = dense(o, "ctype")
    _, dom, (i, j, k) = centres(o)
    o.poke("ctype", sparse(o, with_phi(ct, np.where(j < J, -7.9, 7.9))))
    return o, dom, j


def test_depth_passes_one_cell_per_pass_then_exact_distance():
    J = 7
    o, dom, j = _planar(J)
    depth = J - 1 - j  # cells below the surface layer
    water = dom & (j < J)
    o.kernel("narrow band")  # nbPhi, nbJacA, nbJacB, nbJacA: three passes
    got = dense(o, "ctype") >> 16
    want = quant(np.where(depth <= 2, -(depth + 0.5), -7.9))
    assert np.array_equal(got[water], want[water])
    assert np.array_equal(got[dom & (j >= J)], quant(np.full(np.sum(dom & (j >= J)), 7.9)))
    assert np.all(got[~dom] == 0)
    # every further stage resolves three more layers; converged, the depth is exact
    o.kernel("narrow band")
    got = dense(o, "ctype") >> 16
    want = quant(np.where(depth <= 5, -(depth + 0.5), -7.9))
    assert np.array_equal(got[water], want[water])
    o.kernel("narrow band")
    got = dense(o, "ctype") >> 16
    assert np.array_equal(got[water], quant(-(depth + 0.5))[water])
    # and a further stage changes nothing (a fixed point)
    o.kernel("narrow band")
    assert np.array_equal(dense(o, "ctype") >> 16, got)


def test_nb_phi_shifts_a_linear_depth_by_the_kept_velocity():
    d = 1 / 32  # h = 1/16
    pos, vel = lattice(d, (-0.5, 0.0, -0.5), (0.5, 1.0, 0.5))
    o = start(d, (1, 1, 1), pos, vel, 3, dt=1 / 32)  # u dt = h/2 for u = 1
    ct = dense(o, "ctype")
    _, dom, (i, j, k) = centres(o)
    phi = -7.0 + 0.25 * i  # -7 .. -3.25 inside the domain
    o.poke("ctype", sparse(o, with_phi(ct, np.where(dom, phi, 8.0))))
    vg = np.zeros(ct.shape + (3,), np.float32)
    vg[..., 0] = 1.0
    o.poke("velg", sparse(o, vg))
    o.kernel("nbPhi")
    rhs = dense(o, "rhs").astype(np.float64)
    want = np.where(i >= 1, -7.0 + 0.25 * (i - 0.5), -7.0)  # phiSample(x - dt u): half a cell upwind, clamped at the wall
    assert np.max(np.abs(rhs[dom] - want[dom])) <= 4e-6
    assert np.all(rhs[~dom] == 8.0)
    # the same field carried the other way reads half a cell downwind
    o.poke("velg", sparse(o, -vg))
    o.kernel("nbPhi")
    rhs = dense(o, "rhs").astype(np.float64)
    nx = o.dims.nx
    want = np.where(i <= nx - 2, -7.0 + 0.25 * (i + 0.5), np.nan)
    sel = dom & (i <= nx - 2)
    assert np.max(np.abs(rhs[sel] - want[sel])) <= 4e-6


def test_nb_vel_fills_only_particle_less_fluid_faces():
    d = 1 / 32
    pos, vel = lattice(d, (-0.3, 0.0, -0.3), (0.3, 0.4, 0.3))
    o = start(d, (1, 1, 1), pos, vel, 5, dt=1 / 64)  # through classify (types set)
    ct = dense(o, "ctype")
    _, dom, (i, j, k) = centres(o)
    fluid = dom & ((ct & 0xFF) == 1)
    has = storage(o)
    u = np.array([0.375, -0.25, 0.125], np.float32)
    vg = np.broadcast_to(u, ct.shape + (3,)).astype(np.float32)
    o.poke("velg", sparse(o, vg))
    accw = dense(o, "accw").copy()
    hole = ((i + 2 * j + k) % 3 == 0)
    accw[hole] = 0
    accv = np.full(accw.shape, 12345, np.int32)
    o.poke("accw", sparse(o, accw))
    o.poke("accv", sparse(o, accv))
    o.kernel("nbVel")
    gv, gw = dense(o, "accv"), dense(o, "accw")
    for ax in range(3):
        lo = np.roll(fluid, 1, axis=2 - ax)
        lo_in = [i, j, k][ax] >= 1
        filled = fluid & lo & lo_in & (accw[..., ax] <= 0)
        assert filled.sum() > 100
        assert np.all(gv[..., ax][filled] == int(round(float(u[ax]) * 65536)))
        assert np.all(gw[..., ax][filled] == 65536)
        rest = has & ~filled
        assert np.all(gv[..., ax][rest] == 12345)
        assert np.array_equal(gw[..., ax][rest], accw[..., ax][rest])


def test_save_vel_is_the_projected_face_velocity():
    d = 1 / 32
    pos, vel = lattice(d, (-0.5, 0.0, -0.5), (0.0, 0.5, 0.5), vel=(0.3, 0.0, -0.2))
    o = start(d, (1, 1, 1), pos, vel, 9)  # through project (+ saveVel)
    fv = o.dump("facev").view(np.uint32)
    assert np.array_equal(o.dump("velg").view(np.uint32), fv)
    o.poke("velg", np.zeros(o.dump("velg").shape, np.float32))
    o.kernel("saveVel")
    assert np.array_equal(o.dump("velg").view(np.uint32), fv)


def lowbias32(x):
    """The published 32-bit integer hash 'lowbias32' (C. Wellons, hash-prospector):
    x ^= x >> 16; x *= 0x7feb352d; x ^= x >> 15; x *= 0x846ca68b; x ^= x >> 16."""
    x = np.asarray(x, np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x.astype(np.uint32)


def test_lowbias32_known_values():
    # fixed points of the construction: 0 maps to 0; the outputs of a few inputs,
    # worked by hand from the definition's five steps
    assert int(lowbias32(0)) == 0
    x = 1
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    assert int(lowbias32(1)) == x


def test_reseed_kill_deficit_and_spawn_by_brute_force():
    d = 1 / 32  # h = 1/16, rest = 8 particles per cell
    h = 2 * d
    J = 8
    rng = np.random.default_rng(11)
    # a sparse pool: about 3 particles per cell below y = J h, some fast
    n = 3 * 16 * J * 16
    pos = np.zeros((n, 4), np.float32)
    pos[:, 0] = rng.uniform(-0.5 + 1e-3, 0.5 - 1e-3, n)
    pos[:, 1] = rng.uniform(1e-3, J * h - 1e-3, n)
    pos[:, 2] = rng.uniform(-0.5 + 1e-3, 0.5 - 1e-3, n)
    vel = np.zeros((n, 4), np.float32)
    fast = rng.random(n) < 0.3
    vel[fast, 0] = 2.0  # 2 m/s > 9 h = 0.5625
    vel[~fast, 2] = 0.1
    o = start(d, (1, 1, 1), pos, vel, 11, nb_band=3.0)  # through g2p
    # the state reseed acts on: depth linear in y, kept velocity zero, particles as set
    ct = dense(o, "ctype")
    x, dom, (i, j, k) = centres(o)
    phi = np.where(j < J, -(J - 0.5 - j), 3.0)
    o.poke("ctype", sparse(o, with_phi(ct, phi)))
    o.poke("velg", np.zeros(o.dump("velg").shape, np.float32))
    p_before, v_before = o.dump("pos").copy(), o.dump("vel").copy()
    cnt = dense(o, "cnt")
    ctype = dense(o, "ctype")
    o.kernel("reseed")
    p_after = o.dump("pos")
    # nbKill: calm (|v| < 9h) particles in cells deeper than the band
    dd = o.dims
    cell = np.clip(np.floor((p_before[:, :3].astype(np.float64) - [dd.x0, dd.y0, dd.z0]) / dd.h).astype(int), 0,
                   [dd.nx - 1, dd.ny - 1, dd.nz - 1])
    cphi = phi_of(ctype[cell[:, 2], cell[:, 1], cell[:, 0]])
    calm = np.linalg.norm(v_before[:, :3].astype(np.float64), axis=1) < 9 * h
    alive0 = p_before[:, 1] >= -1e5
    killed = alive0 & calm & (cphi < -3.0)
    assert killed.sum() > 50
    dead = ~alive0 | killed
    # nbDeficit, in slot-list (tile, then local cell) order; need = min(round(0.85 rest) - n, 64)
    entries = []
    for t in o.dump("tiles"):
        ti, tj, tk = t % dd.tx, (t // dd.tx) % dd.ty, t // (dd.tx * dd.ty)
        for loc in range(512):
            ci, cj, ck = ti * 8 + (loc & 7), tj * 8 + ((loc >> 3) & 7), tk * 8 + (loc >> 6)
            if ci >= dd.nx or cj >= dd.ny or ck >= dd.nz:
                continue
            c = ctype[ck, cj, ci]
            ph = phi_of(c)
            if (c & 0xFF) != 1 or ph > -1.0 or ph < -3.0 or (c >> 8) & 0xFF:
                continue
            nn = max(int(cnt[ck, cj, ci]), 0)
            if nn >= 0.55 * 8:
                continue
            entries.append((ci, cj, ck, min(int(round(0.85 * 8)) - nn, 64)))
    ctl = o.ctl()
    assert ctl[8] == dead.sum() and ctl[9] == len(entries) and ctl[10] == sum(e[3] for e in entries)
    # nbSpawn: dead slots in index order, positions x0 + (cell + hash1(hh + 1..3)) h
    seed = 1  # the first step of a sim runs with seed 1 (sim.js:209)
    slots = np.flatnonzero(dead)
    want = []
    for e, (ci, cj, ck, need) in enumerate(entries):
        for m in range(need):
            hh = ((e * 9781 + m * 6271 + 17) ^ (seed * 2654435761)) & 0xFFFFFFFF
            r = lowbias32([hh + 1, hh + 2, hh + 3]).astype(np.float32) * np.float32(2.3283064e-10)
            cp = np.array([ci, cj, ck], np.float32)
            xw = np.array([dd.x0, dd.y0, dd.z0], np.float32) + (cp + r) * np.float32(dd.h)
            want.append(xw)
    kk = min(len(want), len(slots))
    assert kk > 100
    got = p_after[slots[:kk], :3]
    assert np.array_equal(got.view(np.uint32), np.array(want[:kk], np.float32).view(np.uint32))
    # each spawned particle lies in its own deficient cell and moves with the kept velocity (zero)
    assert np.all(o.dump("vel")[slots[:kk], :3] == 0.0)
    # particles neither killed nor respawned are untouched
    keep = alive0 & ~killed
    assert np.array_equal(p_after[keep].view(np.uint32), p_before[keep].view(np.uint32))
