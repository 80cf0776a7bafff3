"""Order independence of the substep: the same particles uploaded in another
order (so with other ids and another memory order) give, particle for particle,
the same bits after several substeps with a moving collider.  This is what
makes the reference's optional particle sort (opt.sort / sortEvery,
src/flip/sim.js:263-276, wgsl-blocks.js:62-99) a matter of memory order only
for the substep: p2g sums in fixed point, every other stage is per particle or
per cell.  Here particles are binned by tile every substep instead, so opt.sort
is accepted and has nothing left to do."""
import numpy as np
import pytest

from splashtank_amd import capi, scenes
from tests.helpers import opts_for, set_colliders_at

pytestmark = pytest.mark.gpu


def _run(sc, pos, vel, steps):
    g = capi.FlipSimC(opts_for(sc))
    g.upload(pos, vel)
    prm = capi.make_params(**sc.params)
    t = 0.0
    for _ in range(steps):
        t += sc.params["dt"]
        set_colliders_at(g, sc, t)
        g.step(prm)
    p = g.read_particles()
    g.close()
    return p


def test_particle_order_changes_nothing():
    sc = scenes.sparse_splash(d=1.0 / 64, iters=20, dt=1.0 / 240)
    perm = np.random.default_rng(11).permutation(sc.n)
    print(f"[order] sparse splash 32^3, {sc.n} particles, shuffled", flush=True)
    a = _run(sc, sc.pos, sc.vel, 6)
    b = _run(sc, sc.pos[perm], sc.vel[perm], 6)
    # particle perm[k] of the first run is particle k of the second
    assert np.array_equal(a[perm].view(np.uint32), b.view(np.uint32))
