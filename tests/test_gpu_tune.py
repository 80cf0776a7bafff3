"""The pressure-solve autotuner (st_tune_sor, st_batch_tune_sor).

Every candidate path gives the same bits, so tuning may change only speed: a
sim tuned midway and then stepped on the tuned path must match, bit for bit, the
same sim stepped without tuning; the chosen candidate is the fastest one timed.
"""
import numpy as np
import pytest

from splashtank_amd import capi, scenes
from splashtank_amd.capi import FlipBatchC, FlipSimC, StError
from tests.helpers import assert_int_equal, opts_for, set_colliders_at, upload_scene

pytestmark = pytest.mark.gpu


def _same_bits(a, b, tag):
    for f in ("pos", "vel", "pres", "ctype", "cnt"):
        x, y = a.dump(f), b.dump(f)
        if x.dtype.kind == "f":
            x, y = x.view(np.uint32), y.view(np.uint32)
        assert_int_equal(x, y, f"{tag} {f}")


def _check_choice(t):
    ms = [c["ms"] for c in t["candidates"]]
    assert len(ms) >= 2 and all(m > 0 for m in ms), t
    best = t["candidates"][int(np.argmin(ms))]
    assert (t["mode"], t["width"]) == (best["mode"], best["width"]), t


def test_tune_keeps_the_bits():
    sc = scenes.sparse_splash(d=1.0 / 128, iters=30)  # 64^3 cells, moving box collider
    print(f"[tune] splash 64^3, {sc.n} particles", flush=True)
    a, b = FlipSimC(opts_for(sc)), FlipSimC(opts_for(sc))
    upload_scene(a, sc)
    upload_scene(b, sc)
    prm = capi.make_params(**sc.params)
    tuned = capi.make_params(**dict(sc.params, sor_mode=capi.ST_SOR_TUNED))
    t = 0.0
    for step in range(6):
        t += sc.params["dt"]
        set_colliders_at(a, sc, t)
        set_colliders_at(b, sc, t)
        if step == 3:
            res = b.tune_sor(tuned, reps=2)
            print(f"[tune] {res}", flush=True)
            _check_choice(res)
        a.step(prm)
        b.step(tuned if step >= 3 else prm)
        a.sync()
        b.sync()
        _same_bits(a, b, f"step {step}")
    want = {0: 1, 1: 0, 2: 2}[res["mode"]]  # st_sor_path: 0 per iteration, 1 resident, 2 streaming
    if res["mode"] == 0:
        assert b.sor_path() in (1, 2)
    else:
        assert b.sor_path() == want
    a.close()
    b.close()


def test_tune_refuses_without_particles():
    s = FlipSimC(capi.make_opts(1.0 / 32, (1.0, 1.0, 1.0), max_n=10))
    with pytest.raises(StError):
        s.tune_sor(capi.make_params())
    s.close()


def test_batch_tune_keeps_the_bits():
    tanks = scenes.batched_tanks(count=4, seed=9, nmin=16, nmax=40, iters=12)
    prm = capi.make_params(**tanks[0].params)
    solo = [FlipSimC(opts_for(sc)) for sc in tanks]
    bt = [FlipSimC(opts_for(sc)) for sc in tanks]
    for sc, a, b in zip(tanks, solo, bt):
        upload_scene(a, sc)
        upload_scene(b, sc)
    batch = FlipBatchC(bt)
    try:
        for step in range(4):
            print(f"[batch-tune] step {step}", flush=True)
            if step == 2:
                res = batch.tune_sor(prm, reps=2)
                print(f"[batch-tune] {res}", flush=True)
                _check_choice(res)
            for a in solo:
                a.step(prm)
            batch.step(prm)
            batch.sync()
            for sc, a, b in zip(tanks, solo, bt):
                a.sync()
                _same_bits(a, b, f"step {step} {sc.name}")
    finally:
        batch.close()
        for s in solo + bt:
            s.close()
