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
