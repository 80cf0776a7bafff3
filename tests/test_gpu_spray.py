"""GPU whitewater (the spray stage, src/flip/wgsl-whitewater.js:1-163) against
the oracle, through the C ABI: every whitewater slot (position, life, velocity,
kind) and every particle bit for bit after each checked substep, with the air
field on, from a shear pool and from a ball splashing next to a moving box
collider."""
import math
import numpy as np
import pytest

from splashtank_amd import capi, scenes
from splashtank_amd.capi import FlipSimC
from oracle.oracle import Oracle
from tests.helpers import assert_int_equal, opts_for, set_colliders_at, upload_scene
from tests.test_oracle_pins import lattice

pytestmark = pytest.mark.gpu

SPR = dict(spr_on=1, spr_rate=3000.0, spr_min=0.5, spr_life=1.2, spr_mist=0.3, spr_drag=1.0, spr_mdrag=4.0,
           air_on=1, air_couple=3.0, air_decay=0.5, air_blur=0.35, wind=(0.4, 0.0, 0.1))


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _compare(g, o, tag):
    pg, vg = g.read_spray()
    po, vo = o.read_spray()
    assert pg.shape == po.shape, f"{tag}: {pg.shape} vs {po.shape} slots handed out"
    assert_int_equal(_bits(pg), _bits(po), f"{tag} spray positions/life")
    assert_int_equal(_bits(vg), _bits(vo), f"{tag} spray velocities/kind")
    assert_int_equal(_bits(g.dump("pos")), _bits(o.dump("pos")), f"{tag} particle positions")
    return int((pg[:, 3] > 0).sum())


def test_spray_shear_pool_matches_oracle():
    d = 1 / 64
    shear = lambda p: np.stack([0.0 * p[:, 0], 3.0 * np.sin(2 * np.pi * p[:, 0]), 0.0 * p[:, 0]], axis=1)  # noqa: E731
    pos, vel = lattice(d, (-0.5, 0.0, -0.5), (0.5, 0.2, 0.5), velfn=shear)
    opts = capi.make_opts(d, (1, 1, 1), max_n=len(pos))
    g, o = FlipSimC(opts), Oracle(opts)
    for s in (g, o):
        s.upload(pos, vel)
        s.set_spray_cap(1 << 15)
    prm = capi.make_params(dt=1 / 60, g=9.81, iters=20, **SPR)
    print(f"[spray] shear pool 32^3, {len(pos)} particles", flush=True)
    alive = 0
    for step in range(14):
        g.step(prm)
        o.step(prm)
        if step % 3 == 2 or step == 13:
            alive = _compare(g, o, f"step {step}")
            print(f"[spray] step {step}: {alive} living whitewater points", flush=True)
    assert alive > 100
    g.close()


def test_spray_splash_with_collider_matches_oracle():
    sc = scenes.sparse_splash(d=1.0 / 64, iters=15, dt=1.0 / 120)  # 32^3 cells, ball, moving box collider
    opts = opts_for(sc)
    g, o = FlipSimC(opts), Oracle(opts)
    for s in (g, o):
        upload_scene(s, sc)
        s.set_spray_cap(1 << 15)
    prm = capi.make_params(**dict(sc.params, **SPR))
    print(f"[spray] splash 32^3 with a moving box, {sc.n} particles", flush=True)
    t = 0.0
    for step in range(40):
        t += sc.params["dt"]
        set_colliders_at(g, sc, t)
        set_colliders_at(o, sc, t)
        g.step(prm)
        o.step(prm)
        if step % 8 == 7:
            alive = _compare(g, o, f"step {step}")
            print(f"[spray] step {step}: {alive} living whitewater points", flush=True)
    g.close()


def test_spray_pool_resize_and_clear():
    sc = scenes.dam_break(d=1.0 / 32)
    g = FlipSimC(opts_for(sc))
    upload_scene(g, sc)
    assert g.read_spray()[0].shape[0] == 0      # no pool: nothing to read
    g.set_spray_cap(100)
    p, v = g.read_spray()
    assert p.shape == (1, 4) and np.all(p == 0)  # nothing handed out yet: one empty slot
    g.step(capi.make_params(**dict(sc.params, **SPR)))
    g.set_spray_cap(50)                           # a new, empty pool
    p, _ = g.read_spray()
    assert np.all(p[:, 3] == 0)
    g.lib.st_clear(g.h)
    p, _ = g.read_spray()
    assert p.shape == (1, 4) and np.all(p == 0)  # clear() starts the ring again
    g.close()


def test_drips_from_a_box_lifted_out_of_the_water_match_oracle():
    from splashtank_amd import scenes as sc_mod
    d = 1 / 64
    pos, vel = lattice(d, (-0.5, 0.0, -0.5), (0.5, 0.25, 0.5))
    opts = capi.make_opts(d, (1, 1, 1), max_n=len(pos))
    g, o = FlipSimC(opts), Oracle(opts)
    h = 2 * d
    dims, lo, data = sc_mod.box_sdf((0.1,) * 3, h, 4 * h)
    samples = sc_mod.drip_samples(dict(id=1, dims=dims, origin=lo, cell=h, data=data), h, 0)
    path = lambda t: np.array([0.0, 0.15 + 1.2 * t, 0.0])  # noqa: E731  (rises out of the pool)
    for s in (g, o):
        s.upload(pos, vel)
        s.set_spray_cap(1 << 15)
        s.set_drips(samples)
    prm = capi.make_params(dt=1 / 120, g=9.81, iters=15, **dict(SPR, spr_rate=0.0, drip_on=1, drip_rate=40.0, drip_dry=1.5))
    print(f"[drips] box lifted from a pool, {len(samples)} drip samples, {len(pos)} particles", flush=True)
    t = 0.0
    for step in range(36):
        t += 1 / 120
        p1, p0 = path(t), path(t - 1 / 120)
        cf = sc_mod.collider_record(0, dims, lo, h, p1, vel=(p1 - p0) * 120)
        for s in (g, o):
            s.set_colliders(cf, data, 1)
            s.step(prm)
        if step % 6 == 5:
            alive = _compare(g, o, f"step {step}")
            print(f"[drips] step {step}: {alive} living whitewater points", flush=True)
    assert alive > 20
    g.close()


def test_drips_from_a_tilted_spinning_box_match_oracle():
    """The drips of a collider turned off the axes and spinning as it rises out
    of the pool: the drip samples' world positions and surface velocities go
    through the collider's rotation and its spin term (d_dripGeom, colSurfVel)."""
    from splashtank_amd import scenes as sc_mod
    d = 1 / 64
    pos, vel = lattice(d, (-0.5, 0.0, -0.5), (0.5, 0.25, 0.5))
    opts = capi.make_opts(d, (1, 1, 1), max_n=len(pos))
    g, o = FlipSimC(opts), Oracle(opts)
    h = 2 * d
    dims, lo, data = sc_mod.box_sdf((0.14, 0.05, 0.09), h, 4 * h)
    samples = sc_mod.drip_samples(dict(id=1, dims=dims, origin=lo, cell=h, data=data), h, 0)
    ax = np.array([0.3, 1.0, 0.2]) / np.linalg.norm([0.3, 1.0, 0.2])
    spin = 2.5
    R0 = sc_mod.axis_rotation((0, 0, 1), math.radians(25))
    path = lambda t: np.array([0.05, 0.14 + 1.1 * t, -0.03])  # noqa: E731
    for s in (g, o):
        s.upload(pos, vel)
        s.set_spray_cap(1 << 15)
        s.set_drips(samples)
    prm = capi.make_params(dt=1 / 120, g=9.81, iters=15, **dict(SPR, spr_rate=0.0, drip_on=1, drip_rate=40.0, drip_dry=1.5))
    print(f"[drips] tilted spinning box from a pool, {len(samples)} drip samples", flush=True)
    t = 0.0
    alive = 0
    for step in range(36):
        t += 1 / 120
        p1, p0 = path(t), path(t - 1 / 120)
        R = sc_mod.axis_rotation(ax, spin * t) @ R0
        cf = sc_mod.collider_record(0, dims, lo, h, p1, R=R, vel=(p1 - p0) * 120, w=tuple(spin * ax))
        for s in (g, o):
            s.set_colliders(cf, data, 1)
            s.step(prm)
        if step % 6 == 5:
            alive = _compare(g, o, f"step {step}")
            print(f"[drips] step {step}: {alive} living whitewater points", flush=True)
    assert alive > 20
    g.close()
