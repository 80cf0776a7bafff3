"""The full-feature case on the GPU: the ocean side boundary over a deep pool in
narrow-band mode (band 3), a ball of water dropped in, whitewater with the air
field, and two SDF colliders (a tilted box spinning as it moves, a rotated
sphere sinking and rising) -- every branch of the substep at once.

At 48^3 the HIP substep is compared with the oracle: stage by stage for the
first substeps (integer fields bit-exact, f32 fields within 1e-5), then every
few substeps the particles (id order), every whitewater slot and the narrow
band's control words, up to when whitewater is being thrown off.  At 512^3
(about a tenth of the tiles in use) the same scene runs through properties that
do not depend on size: finite state inside the box, the particle bookkeeping of
the kill and reseed, no tile without room, the whitewater pool bounds, and the
FP64 norms of the pressure solve.
"""
import numpy as np
import pytest

from splashtank_amd import capi, scenes
from splashtank_amd.capi import FlipSimC
from oracle.oracle import Oracle
from tests.helpers import assert_f32_close, assert_int_equal, opts_for, set_colliders_at, upload_scene

pytestmark = pytest.mark.gpu

INT_FIELDS = {"tiles", "tmap", "ctype", "cnt", "accv", "accw"}
CHECKS = {
    capi.ST_STAGE_BLOCKS: ["tiles", "tmap", "ctype", "velg"],
    capi.ST_STAGE_P2G: ["cnt", "accv", "accw"],
    capi.ST_STAGE_NARROWBAND: ["rhs", "ctype"],
    capi.ST_STAGE_CLASSIFY: ["ctype", "pres", "accv", "accw"],
    capi.ST_STAGE_NORMFACES: ["facev", "faceo"],
    capi.ST_STAGE_DIVERGENCE: ["rhs"],
    capi.ST_STAGE_PRESSURE: ["pres"],
    capi.ST_STAGE_PROJECT: ["facev", "velg"],
    capi.ST_STAGE_G2P: ["pos", "vel"],
    capi.ST_STAGE_RESEED: ["pos", "vel"],
}


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _step_params(sc, step, t):
    kw = sc.params_at(t)
    if step == 0:
        fills = list(sc.nb_fills)
        kw.update(nb_init=1, nfill=len(fills) // 8, fills=fills)
    return capi.make_params(**kw)


def _check(g, o, fields, tag):
    for f in fields:
        a, b = g.dump(f), o.dump(f)
        if f in INT_FIELDS:
            assert_int_equal(a, b, f"{tag}: {f}")
        else:
            assert_f32_close(a, b, f"{tag}: {f}")


def _check_spray(g, o, tag):
    pg, vg = g.read_spray()
    po, vo = o.read_spray()
    assert pg.shape == po.shape, f"{tag}: {pg.shape} vs {po.shape} whitewater slots handed out"
    assert_int_equal(_bits(pg), _bits(po), f"{tag}: whitewater positions/life")
    assert_int_equal(_bits(vg), _bits(vo), f"{tag}: whitewater velocities/kind")
    return int((pg[:, 3] > 0).sum()) if len(pg) else 0


def test_full_feature_matches_oracle_48():
    sc = scenes.stress(d=1.0 / 96, iters=20, dt=1.0 / 120)  # 48^3 cells
    opts = opts_for(sc)
    g, o = FlipSimC(opts), Oracle(opts)
    for s in (g, o):
        upload_scene(s, sc)
        s.set_spray_cap(sc.spray_cap)
    print(f"[stress] 48^3, {sc.n} particle slots, {len(sc.params['waves'])} waves, 2 colliders, spray cap {sc.spray_cap}",
          flush=True)
    t = 0.0
    made = 0
    for step in range(36):
        t += sc.params["dt"]
        for s in (g, o):
            set_colliders_at(s, sc, t)
        prm = _step_params(sc, step, t)
        if step < 3:
            g.begin_step(prm)
            o.begin_step(prm)
            for st in range(capi.ST_STAGE_COUNT):
                g.stage(st)
                o.stage(st)
                _check(g, o, CHECKS.get(st, []), f"step {step} after {capi.STAGES[st]}")
        else:
            g.step(prm)
            o.step(prm)
        if step < 3 or step % 6 == 5:
            tag = f"step {step}"
            _check(g, o, ["pos", "vel", "ctype", "pres"], tag)
            assert list(g.ctl()[7:11]) == list(o.ctl()[7:11]), (tag, g.ctl()[7:11], o.ctl()[7:11])
            made = _check_spray(g, o, tag)
            alive = int((o.dump("pos")[:, 1] > -1e5).sum())
            print(f"[stress] {tag}: {alive} particles, {made} living whitewater points, reseed ctl {list(o.ctl()[7:11])}",
                  flush=True)
    ty = g.dump("ctype") & 0xFF
    col = (g.dump("ctype") >> 8) & 0xFF
    assert np.sum((ty == 2) & (col == 0)) > 1000, "the ocean's boundary holds solid cells"
    assert np.sum(col == 1) > 50 and np.sum(col == 2) > 50, "both colliders own cells"
    assert made > 50, "whitewater is thrown off by the end"
    g.close()


def test_full_feature_512_properties():
    sc = scenes.stress(d=1.0 / 1024, level=0.1, layer=0.06, drop=(0.1, 0.3, -0.05, 0.06), iters=40, dt=1.0 / 480)
    opts = opts_for(sc)
    g = FlipSimC(opts)
    upload_scene(g, sc)
    g.set_spray_cap(sc.spray_cap)
    d = g.dims
    n = sc.n
    print(f"[stress] 512^3 ({d.nx}x{d.ny}x{d.nz}), {n} particle slots, spray cap {sc.spray_cap}", flush=True)
    t = 0.0
    for step in range(24):
        t += sc.params["dt"]
        set_colliders_at(g, sc, t)
        g.step(_step_params(sc, step, t))
        if step % 8 == 7:
            p = g.dump("pos")
            v = g.dump("vel")
            alive = p[:, 1] > -1e5
            c = g.ctl()
            assert np.isfinite(p[alive]).all() and np.isfinite(v[alive]).all(), "finite particles"
            q = p[alive, :3]
            assert np.all((q[:, 0] >= d.x0) & (q[:, 0] <= d.x1) & (q[:, 1] >= d.y0) & (q[:, 1] <= d.y1) &
                          (q[:, 2] >= d.z0) & (q[:, 2] <= d.z1)), "particles inside the box"
            # the reseed: ctl[8] dead slots after the kill, ctl[10] particles asked for; it spawned
            # min(asked, dead), so the living count is what the kill left plus those
            assert int(alive.sum()) == n - int(c[8]) + min(int(c[10]), int(c[8])), (int(alive.sum()), list(c[7:11]))
            assert c[5] == 0, "every tile found room"
            tiles = int(c[4])
            nrm = g.solve_norms()
            rel = nrm["res_l2"] / max(nrm["rhs_l2"], 1e-300)
            sp, sv = g.read_spray()
            living = sp[:, 3] > 0 if len(sp) else np.zeros(0, bool)
            print(f"[stress] 512^3 step {step}: {int(alive.sum())} particles, {tiles} of {d.ntiles} tiles in use "
                  f"({tiles / d.ntiles:.1%}), relative residual {rel:.3g} after {sc.params['iters']} sweeps, "
                  f"{int(living.sum())} whitewater points; norms {nrm}", flush=True)
            assert np.isfinite([nrm[k] for k in nrm]).all() and rel < 1.0, nrm
            # the projection leaves div - tgt = -(dt / h^2) r in every fluid cell: the same in norm
            k = float(np.float32(sc.params["dt"])) / float(d.h) ** 2
            assert nrm["err_l2"] == pytest.approx(k * nrm["res_l2"], rel=1e-3), nrm
            assert len(sp) <= sc.spray_cap
            if living.any():
                w, sw = sp[living, :3], sv[living]
                assert np.isfinite(w).all() and np.isfinite(sw).all()
                assert np.all((w[:, 1] >= d.y0) & (w[:, 1] <= d.y1))
                assert set(np.unique(np.floor(sw[:, 3] + 0.5)).astype(int)) <= {0, 1, 2, 3}
    assert 0.05 <= tiles / d.ntiles <= 0.2, "about a tenth of the tiles in use"
    g.close()
