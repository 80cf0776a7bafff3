"""Pins of the oracle's whitewater stage (src/flip/wgsl-whitewater.js:1-163).

The reference holds no whitewater fixtures, so the restatement is pinned by
closed forms of single points: a spray point in the air follows the drag law
exactly, a point leaving the box dies, mist landing in water dies, a settled
bubble loses life at 0.3 per second; the air field far from any push settles
to the wind at the decay rate; emission waits for the starting water to settle
and stays off without o.spr; and the shared f32 sin/cos is within 1.5e-7 of
the f64 one.
"""
import math

import numpy as np
import pytest

from splashtank_amd import capi
from oracle.oracle import Oracle, sincosf
from tests.test_oracle_pins import lattice


def test_sincos_error_bound():
    xs = np.concatenate([np.linspace(0.0, 2 * np.pi, 40001, dtype=np.float32),
                         np.float32([0.0, np.pi / 2, np.pi, 1.5 * np.pi, 6.2831853, 0.7853982, 2.3561945])])
    worst = 0.0
    for x in xs:
        s, c = sincosf(float(x))
        worst = max(worst, abs(s - math.sin(float(x))), abs(c - math.cos(float(x))))
    assert worst <= 1.5e-7, worst


def _pool_sim(fp64=True, cap=64, vel=(0.0, 0.0, 0.0), top=0.2, velfn=None):
    d = 1 / 32
    pos, v = lattice(d, (-0.5, 0.0, -0.5), (0.5, top, 0.5), vel=vel, velfn=velfn)
    o = Oracle(capi.make_opts(d, (1, 1, 1), max_n=len(pos)), fp64=fp64)
    o.upload(pos, v)
    o.set_spray_cap(cap)
    return o


def _prm(**kw):
    base = dict(dt=1 / 120, g=9.81, iters=20, wind=(0.3, 0.0, -0.2))
    base.update(kw)
    return capi.make_params(**base)


@pytest.mark.parametrize("fp64", [True, False])
def test_spray_point_in_air_follows_the_drag_law(fp64):
    o = _pool_sim(fp64)
    p0 = np.array([[0.1, 0.7, -0.05, 2.0]])   # far above the water: no storage, dry
    v0 = np.array([[1.0, 2.0, 0.5, 0.0]])     # spray, kind held for 0 s
    o.spray_put(p0, v0)
    prm = _prm(spr_on=1, spr_rate=0.0, spr_min=1.0, spr_life=2.0, spr_drag=1.5, spr_mdrag=4.0)
    o.step(prm)
    p1, v1 = o.read_spray()
    dt, g, drag = float(np.float32(prm.dt)), float(np.float32(9.81)), 1.5
    wind = np.float32([0.3, 0.0, -0.2]).astype(np.float64)  # the parameter block holds f32
    vg = v0[0, :3] - np.array([0.0, g * dt, 0.0])
    want_v = vg + (wind - vg) * (1.0 - math.exp(-drag * dt))
    tol = 1e-12 if fp64 else 2e-6
    np.testing.assert_allclose(v1[0, :3], want_v, rtol=tol, atol=tol)
    np.testing.assert_allclose(p1[0, :3], p0[0, :3] + want_v * dt, rtol=tol, atol=tol)
    assert p1[0, 3] == pytest.approx(2.0 - dt, abs=tol)
    assert v1[0, 3] == pytest.approx(0.4 * dt / float(np.float32(0.2)), abs=tol)  # kind 0, held for one step (WWHOLD is an f32 constant)


def test_points_leaving_the_box_or_mist_in_water_die_and_bubbles_fade():
    o = _pool_sim(True)
    p0 = np.array([[0.0, 0.999, 0.0, 1.0],     # about to leave through the top
                   [0.0, 0.05, 0.0, 1.0],      # mist deep in the pool
                   [0.02, 0.06, 0.03, 1.0]])   # a settled bubble deep in the pool
    v0 = np.array([[0.0, 3.0, 0.0, 0.0], [0.0, 0.0, 0.0, 1.0], [0.0, 0.0, 0.0, 3.4]])
    o.spray_put(p0, v0)
    prm = _prm(spr_on=1, spr_rate=0.0, spr_min=1.0, spr_life=2.0, spr_drag=1.0, spr_mdrag=4.0)
    o.step(prm)
    p1, v1 = o.read_spray()
    dt = float(np.float32(prm.dt))
    assert p1[0, 3] == 0.0
    assert p1[1, 3] <= 0.0
    assert p1[2, 3] == pytest.approx(1.0 - 0.3 * dt, abs=1e-12)
    assert 3.0 <= v1[2, 3] <= 3.4 + 1e-12          # still a bubble
    assert v1[2, 1] > 0.0                            # rising toward the water's velocity + rise


def test_air_far_from_pushes_settles_to_the_wind():
    # the water's top cells are half full (4 particles of 8), so their particles drag the air along x
    o = _pool_sim(True, vel=(1.0, 0.0, 0.0), top=0.16)
    prm = _prm(spr_on=1, spr_rate=0.0, spr_min=1.0, spr_life=2.0, spr_drag=1.0, spr_mdrag=4.0,
               air_on=1, air_couple=3.0, air_decay=2.0, air_blur=0.35)
    o.step(prm)
    air, cell = o.read_air()
    assert cell == pytest.approx(2 / 16)              # two solver cells (h = 1/16) per air cell
    n = round(air.shape[0] ** (1 / 3))
    a = air.reshape(n, n, n, 4)                        # [k, j, i]
    dt = float(np.float32(prm.dt))
    want = np.float32([0.3, 0.0, -0.2]).astype(np.float64) * (1.0 - math.exp(-2.0 * dt))
    top = a[:, -1, :, :3].reshape(-1, 3)               # the top layer: nothing pushes there
    np.testing.assert_allclose(top, np.broadcast_to(want, top.shape), rtol=1e-12, atol=1e-15)
    assert np.abs(a[:, 1, :, 0] - want[0]).max() > 0.1  # the surface layer pushes the air cells it is in


def test_emission_waits_for_settling_and_needs_spr():
    # a divergence-free shear (columns of water rising and falling at up to 3 m/s):
    # fast, exposed water at the top of the rising columns
    shear = lambda p: np.stack([0.0 * p[:, 0], 3.0 * np.sin(2 * np.pi * p[:, 0]), 0.0 * p[:, 0]], axis=1)  # noqa: E731
    o = _pool_sim(True, cap=4096, velfn=shear)
    on = _prm(dt=1 / 60, spr_on=1, spr_rate=4000.0, spr_min=0.5, spr_life=1.0, spr_mist=0.3, spr_drag=1.0, spr_mdrag=4.0)
    o.step(on)                                         # tm = 0: the starting water is settling
    p, v = o.read_spray()
    assert np.all(p[:, 3] == 0.0)
    off = _prm(dt=1 / 60)
    for _ in range(8):
        o.step(off)                                     # no o.spr: nothing thrown off
    p, v = o.read_spray()
    assert np.all(p[:, 3] == 0.0)
    o.step(on)                                          # tm > 0.1 now
    p, v = o.read_spray()
    alive = p[:, 3] > 0.0
    assert alive.sum() > 10
    kinds = v[alive, 3]
    assert set(np.unique(kinds)) <= {0.0, 1.0, 3.0}
    assert np.all(p[alive, 3] <= 1.6 * 1.4 + 1e-9)      # life <= spray life x (0.6 + 0.8) x mist factor


def _drip_box(o, pos, vel=(0.0, 0.0, 0.0)):
    """a 0.08 half-size box collider at `pos` in oracle or GPU sim `o`, and its drip samples"""
    from splashtank_amd import scenes
    h = 2 / 32
    dims, lo, data = scenes.box_sdf((0.08,) * 3, h, 4 * h)
    cf = scenes.collider_record(0, dims, lo, h, pos, vel=vel)
    o.set_colliders(cf, data, 1)
    return scenes.drip_samples(dict(id=1, dims=dims, origin=lo, cell=h, data=data), h, 0)


def test_drips_dry_out_and_let_go_of_drops():
    o = _pool_sim(True, cap=8192)
    samples = _drip_box(o, (0.0, 0.6, 0.0))        # well above the water: every sample exposed
    samples[:, 7] = 1.0                             # and wet, as if just lifted out
    o.set_drips(samples)
    prm = _prm(spr_on=1, spr_rate=0.0, spr_min=1.0, spr_life=1.5, spr_drag=1.0, spr_mdrag=4.0,
               drip_on=1, drip_rate=1.0e7, drip_dry=2.0)
    o.step(prm)
    p, v = o.read_spray()
    n = len(samples)
    assert n > 20 and p.shape[0] == n               # every sample let go of one drop (rate x dt >> 1)
    assert np.all(p[:, 3] == pytest.approx(3.0))   # spray life x 2
    assert np.all(v[:, 3] == 0.0)                  # spray
    h = 1 / 16
    speed = np.linalg.norm(v[:, :3], axis=1)
    assert speed.max() <= 2 * h + 1e-12            # a static box: only the push off its surface, h x 2 x hash
    o.step(prm)                                     # wetness fell to 1 - dt / dry: fewer, still every sample
    p2, _ = o.read_spray()
    assert p2.shape[0] == min(2 * n, 8192)
    dry = _pool_sim(True, cap=64)
    samples2 = _drip_box(dry, (0.0, 0.6, 0.0))      # dry samples never drip
    dry.set_drips(samples2)
    dry.step(prm)
    assert np.all(dry.read_spray()[0][:, 3] == 0.0)
