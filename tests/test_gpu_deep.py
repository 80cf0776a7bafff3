"""Deep water points of the narrow band look (st_read_deep; deepPick,
wgsl-whitewater.js:279-297) against the oracle's restatement on the same
narrow-band pool: the same cells in the same active-cell order, bit-identical
(centres and speeds), and what the rule fixes (every point a fluid cell centre
whose cell holds no particle; speed word 4096 + |velG| capped at 63)."""
import numpy as np
import pytest

from splashtank_amd import capi, scenes
from oracle.oracle import Oracle
from tests.helpers import opts_for

pytestmark = pytest.mark.gpu


def test_deep_points_match_oracle():
    sc = scenes.nb_pool(d=1.0 / 64, depth=0.5, iters=10)
    g = capi.FlipSimC(opts_for(sc))
    o = Oracle(opts_for(sc))
    try:
        for s in (g, o):
            s.upload(sc.pos, sc.vel)
        fills = [float(x) for x in np.asarray(sc.nb_fills, np.float64).ravel()]
        assert len(g.read_deep()) == 0  # nothing stepped yet
        for k in range(3):
            kw = dict(sc.params)
            if k == 0:  # the narrow band starts from the fill shapes
                kw.update(nb_init=1, nfill=len(fills) // 8, fills=fills)
            prm = capi.make_params(**kw)
            g.step(prm)
            o.step(prm)
            a = g.read_deep()
            b = o.read_deep()
            print(f"[deep] substep {k}: {len(a)} points (oracle {len(b)})", flush=True)
            assert len(a) > 100 and len(a) == len(b)
            assert np.array_equal(a, b.astype(np.float32))
            w = a[:, 3] - 4096.0
            assert (w >= 0).all() and (w <= 63).all()
            h = float(o.dims.h)
            cell = (a[:, :3] - np.array([o.dims.x0, o.dims.y0, o.dims.z0])) / h - 0.5
            assert np.allclose(cell, np.round(cell), atol=1e-3)  # cell centres
        p4 = g.dump("pos")
        occupied = {tuple(c) for c in np.floor((p4[:, :3] - np.array([o.dims.x0, o.dims.y0, o.dims.z0])) / h).astype(int)}
        deep = {tuple(c) for c in np.round(cell).astype(int)}
        assert not (deep & occupied)  # no point where a particle sits
    finally:
        g.close()
        o.close()
