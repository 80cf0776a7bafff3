"""Asynchronous readback through the C ABI (st_read_begin / st_read_wait /
st_read_free), the path behind the Node facade's Promises: the result is the
state when the read was begun, whatever is enqueued after it, and it equals the
blocking read of that state."""
import numpy as np
import pytest

from splashtank_amd import capi, scenes


def _sim(sc):
    g = capi.FlipSimC(capi.make_opts(sc.d, sc.size, ratio=sc.ratio, open_=sc.open_, max_n=sc.n))
    g.upload(sc.pos, sc.vel)
    return g


@pytest.mark.gpu
def test_async_read_is_the_state_at_begin():
    sc = scenes.dam_break(d=1.0 / 64, iters=30)
    prm = capi.make_params(**sc.params)
    ref = _sim(sc)
    for _ in range(6):
        ref.step(prm)
    want = ref.read_particles()
    ctl_want = ref.ctl()
    ref.close()
    g = _sim(sc)
    for _ in range(6):
        g.step(prm)
    print("[async] reads begun after 6 substeps, 4 more enqueued behind them", flush=True)
    pend = g.read_particles_async()
    pctl = capi.PendingRead(g.lib, g.h, capi.ST_READ_CTL, 16)
    for _ in range(4):
        g.step(prm)
    got = pend.wait()
    ctl = pctl.wait()
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.array_equal(ctl, np.asarray(ctl_want).view(np.uint32)[:16])
    # the later substeps did run: a fresh read differs
    assert not np.array_equal(g.read_particles().view(np.uint32), want.view(np.uint32))
    # a ticket left unwaited is released cleanly, and outlives its sim
    left = g.read_particles_async(100)
    g.close()
    left.close()
