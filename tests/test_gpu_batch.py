"""Batched tanks (st_batch_*): many tanks advanced with one launch per kernel
for all of them must give every tank exactly the bits st_step gives it alone,
and (through that) the oracle's results.  Compared bit for bit: particle
positions and velocities (id order), pressure, cell types, fixed-point face
sums and the control words, after every substep."""
import numpy as np
import pytest

from splashtank_amd import capi, scenes
from splashtank_amd.capi import FlipBatchC, FlipSimC, StError
from oracle.oracle import Oracle
from tests.helpers import assert_f32_close, assert_int_equal, opts_for, set_colliders_at, upload_scene

pytestmark = pytest.mark.gpu

FIELDS = ["pos", "vel", "pres", "ctype", "cnt", "accv", "tmap"]


def _tank_set():
    """Small tanks of uneven sizes (16..48 cells a side, 4 to 8 particles per
    cell), one with a moving box collider, plus an empty tank."""
    tanks = scenes.batched_tanks(count=5, seed=5, nmin=16, nmax=48, iters=12)
    splash = scenes.sparse_splash(d=1.0 / 64, iters=9)  # 32^3 cells, moving box collider
    tanks.append(splash)
    return tanks


def _params(tanks):
    # uneven SOR iteration counts, odd and even, so the batched solve has tanks
    # that stop early and tanks whose result ends in the partner buffer
    its = [12, 7, 20, 1, 16, 9]
    return [capi.make_params(**dict(sc.params, iters=its[k % len(its)])) for k, sc in enumerate(tanks)]


def _same(a, b, tag):
    for f in FIELDS:
        x, y = a.dump(f), b.dump(f)
        if x.dtype.kind == "f":
            assert_int_equal(x.view(np.uint32), y.view(np.uint32), f"{tag} {f}")
        else:
            assert_int_equal(x, y, f"{tag} {f}")
    assert_int_equal(a.ctl()[:7], b.ctl()[:7], f"{tag} ctl")


def test_batch_matches_single_tanks():
    tanks = _tank_set()
    prms = _params(tanks)
    solo = [FlipSimC(opts_for(sc)) for sc in tanks]
    bt = [FlipSimC(opts_for(sc)) for sc in tanks]
    empty_solo = FlipSimC(capi.make_opts(1.0 / 64, (1.0, 1.0, 1.0), max_n=1000))
    empty_bt = FlipSimC(capi.make_opts(1.0 / 64, (1.0, 1.0, 1.0), max_n=1000))
    for sc, a, b in zip(tanks, solo, bt):
        upload_scene(a, sc)
        upload_scene(b, sc)
    batch = FlipBatchC(bt[:3] + [empty_bt] + bt[3:])
    prm_b = prms[:3] + [prms[0]] + prms[3:]
    try:
        t = 0.0
        for step in range(5):
            t += tanks[0].params["dt"]
            print(f"[batch] {len(tanks)} tanks + 1 empty, step {step}", flush=True)
            for sc, a, b, p in zip(tanks, solo, bt, prms):
                set_colliders_at(a, sc, t)
                set_colliders_at(b, sc, t)
                a.step(p)
            batch.step(prm_b)
            batch.sync()
            for sc, a, b in zip(tanks, solo, bt):
                a.sync()
                _same(a, b, f"step {step} {sc.name}")
        assert empty_bt.ctl()[4] == 0 and empty_bt.read_particles().shape[0] == 0
        # the tanks still run alone while batched, on the batch's stream
        for sc, a, b, p in zip(tanks, solo, bt, prms):
            a.step(p)
            b.step(p)
            b.sync()
            _same(a, b, f"solo step after batching {sc.name}")
    finally:
        batch.close()
    # released: each tank steps on its own stream again and can be destroyed
    bt[0].step(prms[0])
    bt[0].sync()
    for s in solo + bt + [empty_solo, empty_bt]:
        s.close()
        assert s.h is None


@pytest.mark.parametrize("grid", [None, 0, 8, 2])
def test_batch_resident_solve_phases(grid):
    """The batched SOR in one cooperative launch (default), in several phases
    and with tanks that stream tile windows when the launch is capped at 8 or
    2 workgroups, and one launch per iteration (0): the same bits as each tank
    stepped alone, over three substeps."""
    tanks = _tank_set()
    prms = _params(tanks)
    solo = [FlipSimC(opts_for(sc)) for sc in tanks]
    bt = [FlipSimC(opts_for(sc)) for sc in tanks]
    for sc, a, b in zip(tanks, solo, bt):
        upload_scene(a, sc)
        upload_scene(b, sc)
    batch = FlipBatchC(bt)
    try:
        if grid is not None:
            batch.sor_grid(grid)
        t = 0.0
        for step in range(3):
            t += tanks[0].params["dt"]
            print(f"[batch] resident solve, grid cap {grid}, step {step}", flush=True)
            for sc, a, b, p in zip(tanks, solo, bt, prms):
                set_colliders_at(a, sc, t)
                set_colliders_at(b, sc, t)
                a.step(p)
                a.sync()
            batch.step(prms)
            batch.sync()
            assert batch.sor_path() == (grid != 0)
            for sc, a, b in zip(tanks, solo, bt):
                _same(a, b, f"grid {grid} step {step} {sc.name}")
    finally:
        batch.close()
    for s in solo + bt:
        s.close()


def test_batch_one_param_block_and_oracle():
    """One parameter block for all tanks; the tanks' particles after two
    substeps against the oracle (f32 bar, bit-identical counted)."""
    tanks = scenes.batched_tanks(count=3, seed=11, nmin=16, nmax=24, iters=10)
    prm = capi.make_params(**tanks[0].params)
    gs = [FlipSimC(opts_for(sc)) for sc in tanks]
    os_ = [Oracle(opts_for(sc)) for sc in tanks]
    for sc, g, o in zip(tanks, gs, os_):
        upload_scene(g, sc)
        upload_scene(o, sc)
    batch = FlipBatchC(gs)
    try:
        for step in range(2):
            print(f"[batch-oracle] step {step}", flush=True)
            batch.step(prm)
            for o in os_:
                o.step(prm)
        batch.sync()
        for sc, g, o in zip(tanks, gs, os_):
            for f in ("pos", "vel", "pres"):
                r = assert_f32_close(g.dump(f), o.dump(f), f"{sc.name} {f}")
                print(sc.name, f, r, flush=True)
            assert_int_equal(g.dump("ctype"), o.dump("ctype"), f"{sc.name} ctype")
            assert_int_equal(g.dump("cnt"), o.dump("cnt"), f"{sc.name} cnt")
    finally:
        batch.close()
        for g in gs:
            g.close()


def test_batch_errors():
    sc = scenes.dam_break(d=1.0 / 32)
    a = FlipSimC(opts_for(sc))
    b = FlipSimC(opts_for(sc))
    upload_scene(a, sc)
    upload_scene(b, sc)
    with pytest.raises(StError):
        FlipBatchC([a, a])
    batch = FlipBatchC([a, b])
    try:
        with pytest.raises(StError):
            FlipBatchC([b])  # already in a batch
        assert a.lib.st_destroy(a.h) == capi.ST_ERR_STATE  # refused while batched
        with pytest.raises(StError):
            batch.step(capi.make_params(**dict(sc.params, nb_on=1)))
        with pytest.raises(StError):
            batch.step([capi.make_params(**sc.params)] * 3)  # neither 1 nor one per tank
        batch.step(capi.make_params(**sc.params))
        batch.sync()
    finally:
        batch.close()
    a.close()
    b.close()
