"""HIP graph replay of substeps (st_step_n / st_batch_step_n with use_graph).

A replayed substep must give exactly the bits of the eager one.  Each test
replays captured graphs several times, with eager substeps and synchronizations
of the program between the replays, and checks the state after every chunk
against sims stepped eagerly.
"""
import numpy as np
import pytest

from splashtank_amd import capi, scenes
from splashtank_amd.capi import FlipBatchC, FlipSimC, StError
from tests.helpers import assert_int_equal, opts_for, set_colliders_at, upload_scene

pytestmark = pytest.mark.gpu


def _same_bits(a, b, tag):
    for f in ("pos", "vel", "pres", "ctype", "cnt", "tmap"):
        x, y = a.dump(f), b.dump(f)
        if x.dtype.kind == "f":
            x, y = x.view(np.uint32), y.view(np.uint32)
        assert_int_equal(x, y, f"{tag} {f}")


def test_graph_replay_matches_eager_dam_break():
    sc = scenes.dam_break(d=1.0 / 64, iters=20)
    print(f"[graph] dam break 32^3, {sc.n} particles", flush=True)
    a, b = FlipSimC(opts_for(sc)), FlipSimC(opts_for(sc))
    upload_scene(a, sc)
    upload_scene(b, sc)
    prm = capi.make_params(**sc.params)
    done = 0
    for chunk in range(4):
        n = 3 + chunk  # odd and even chunk lengths: both particle-buffer parities replay
        b.step_n(prm, n, graph=True)
        b.sync()
        for _ in range(n):
            a.step(prm)
        a.sync()
        done += n
        _same_bits(a, b, f"after {done} substeps (graph chunk {chunk})")
        # eager work of the program between replays
        a.step(prm)
        b.step(prm)
        done += 1
        b.sync()
        a.sync()
        _same_bits(a, b, f"after {done} substeps (eager between chunks)")
    a.close()
    b.close()


def test_graph_replay_with_moving_collider():
    sc = scenes.sparse_splash(d=1.0 / 96, iters=15)  # 48^3 cells, box collider moving every substep
    print(f"[graph] splash 48^3 with a moving collider, {sc.n} particles", flush=True)
    a, b = FlipSimC(opts_for(sc)), FlipSimC(opts_for(sc))
    upload_scene(a, sc)
    upload_scene(b, sc)
    prm = capi.make_params(**sc.params)
    t = 0.0
    for step in range(8):
        t += sc.params["dt"]
        set_colliders_at(a, sc, t)
        set_colliders_at(b, sc, t)
        a.step(prm)
        b.step_n(prm, 1, graph=True)
        if step % 3 == 2:
            a.sync()
            b.sync()
            _same_bits(a, b, f"step {step}")
    a.sync()
    b.sync()
    _same_bits(a, b, "end")
    a.close()
    b.close()


def test_graph_refuses_narrow_band():
    sc = scenes.dam_break(d=1.0 / 32)
    s = FlipSimC(opts_for(sc))
    upload_scene(s, sc)
    with pytest.raises(StError):
        s.step_n(capi.make_params(**dict(sc.params, nb_on=1)), 1, graph=True)
    s.close()


def test_batch_graph_replay_matches_eager():
    tanks = scenes.batched_tanks(count=4, seed=13, nmin=16, nmax=40, iters=10)
    prm = capi.make_params(**tanks[0].params)
    solo = [FlipSimC(opts_for(sc)) for sc in tanks]
    bt = [FlipSimC(opts_for(sc)) for sc in tanks]
    for sc, a, b in zip(tanks, solo, bt):
        upload_scene(a, sc)
        upload_scene(b, sc)
    batch = FlipBatchC(bt)
    try:
        done = 0
        for chunk in range(3):
            n = 2 + chunk
            print(f"[batch-graph] chunk {chunk}: {n} replays", flush=True)
            batch.step_n(prm, n, graph=True)
            batch.step(prm)  # eager batched substep between replays
            batch.sync()
            for a in solo:
                for _ in range(n + 1):
                    a.step(prm)
                a.sync()
            done += n + 1
            for sc, a, b in zip(tanks, solo, bt):
                _same_bits(a, b, f"after {done} substeps {sc.name}")
    finally:
        batch.close()
        for s in solo + bt:
            s.close()


def test_graph_replay_follows_whitewater_switches():
    """Replays must follow the spray, air and drip switches of each substep and the
    number of drip samples: a graph captured with one setting is never replayed
    for another.  Eager and replayed sims get the same sequence of settings
    (spray on, off, drips on, fewer drip samples) and must agree bit for bit,
    whitewater slots included."""
    from tests.test_gpu_spray import SPR
    from tests.test_oracle_pins import lattice
    d = 1 / 64
    pos, vel = lattice(d, (-0.5, 0.0, -0.5), (0.5, 0.25, 0.5))
    opts = capi.make_opts(d, (1, 1, 1), max_n=len(pos))
    h = 2 * d
    dims, lo, data = scenes.box_sdf((0.1,) * 3, h, 4 * h)
    samples = scenes.drip_samples(dict(id=1, dims=dims, origin=lo, cell=h, data=data), h, 0)
    a, b = FlipSimC(opts), FlipSimC(opts)
    for s in (a, b):
        s.upload(pos, vel)
        s.set_spray_cap(1 << 14)
        s.set_drips(samples)
    base = dict(dt=1 / 120, g=9.81, iters=12)
    plans = [("spray+air", dict(base, **SPR)),
             ("off", dict(base)),
             ("spray+drips", dict(base, **dict(SPR, spr_rate=500.0, drip_on=1, drip_rate=40.0, drip_dry=1.0))),
             ("fewer drips", None),
             ("spray+air again", dict(base, **SPR))]
    path = lambda t: np.array([0.0, 0.15 + 1.2 * t, 0.0])  # noqa: E731  (a box rising out of the pool)
    t = 0.0
    last = None
    for name, kw in plans:
        if kw is None:  # fewer samples, same allocation: the replay must not scan the old count
            for s in (a, b):
                s.set_drips(samples[: len(samples) // 2])
            kw = last
        last = kw
        prm = capi.make_params(**kw)
        print(f"[graph] whitewater plan '{name}'", flush=True)
        for _ in range(4):
            t += 1 / 120
            p1, p0 = path(t), path(t - 1 / 120)
            cf = scenes.collider_record(0, dims, lo, h, p1, vel=(p1 - p0) * 120)
            for s in (a, b):
                s.set_colliders(cf, data, 1)
            a.step(prm)
            b.step_n(prm, 1, graph=True)
        a.sync()
        b.sync()
        _same_bits(a, b, f"plan {name}")
        pa, va = a.read_spray()
        pb, vb = b.read_spray()
        assert pa.shape == pb.shape, f"plan {name}: {pa.shape} vs {pb.shape} whitewater slots"
        assert_int_equal(pa.view(np.uint32), pb.view(np.uint32), f"plan {name} whitewater positions")
        assert_int_equal(va.view(np.uint32), vb.view(np.uint32), f"plan {name} whitewater velocities")
    a.close()
    b.close()
