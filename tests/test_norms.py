"""FP64 norms of the pressure solve and projection (st_solve_norms).

CPU: the numpy restatement (oracle/norms.py) is pinned by what the projection
fixes -- after a substep, div - tgt = -(dt / h^2) r cell by cell (exact in the
f64 oracle, to f32 rounding in the f32 one) -- and by a solve run to
convergence, whose residual falls to rounding level.
GPU: st_solve_norms (f64 reductions on the device) against the restatement on
the GPU sim's own dumps and on the f32 oracle's (the same bits), and the f64
oracle's norms stated beside them.
"""
import numpy as np
import pytest

from splashtank_amd import capi, scenes
from oracle import norms as onorms
from oracle.oracle import Oracle
from tests.helpers import opts_for, upload_scene


def _dam(d=1 / 32, iters=40):
    """dam break whose particles start displaced by up to half a spacing and with
    random velocities (uneven counts, a solve with work to do)"""
    sc = scenes.dam_break(d=d, iters=iters, dt=1 / 240)
    rng = np.random.default_rng(11)
    sc.pos[:, :3] += rng.uniform(-0.5 * d, 0.5 * d, (sc.n, 3)).astype(np.float32)
    sc.pos[:, 1] = np.maximum(sc.pos[:, 1], np.float32(0.25 * d))
    sc.vel[:, :3] = rng.normal(0.0, 0.5, (sc.n, 3)).astype(np.float32)
    return sc


def _run_oracle(sc, steps, fp64):
    o = Oracle(opts_for(sc), fp64=fp64)
    upload_scene(o, sc)
    prm = capi.make_params(**sc.params)
    for _ in range(steps):
        o.step(prm)
    return o, prm


@pytest.mark.parametrize("fp64", [True, False])
def test_projection_identity_pins_the_norms(fp64):
    sc = _dam(iters=5)
    o, prm = _run_oracle(sc, 3, fp64)
    t = onorms.cell_terms(o, prm.comp, sc.open_)
    h = float(np.float32(o.dims.h))
    dt = float(np.float32(prm.dt))
    assert t["r"].size > 500
    assert np.abs(t["tgt"]).max() > 0.1  # the density target takes part
    # five iterations leave a large residual, so e is far from zero ...
    scale = np.abs(t["e"]).max()
    assert scale > 0.1
    # ... and the projection turns exactly that residual into e
    gap = np.abs(t["e"] + dt / (h * h) * t["r"]).max()
    tol = 1e-12 * scale if fp64 else 1e-5 * max(1.0, np.abs(t["div"]).max())
    assert gap <= tol, (gap, scale)
    # a sign slip in r (or in the projection term) would leave a gap of 2|e|
    assert np.abs(t["e"] - dt / (h * h) * t["r"]).max() > 1e4 * tol


def test_converged_solve_has_rounding_level_residual():
    sc = _dam(iters=3000)
    o, prm = _run_oracle(sc, 1, True)
    n = onorms.norms(o, prm.comp, sc.open_)
    assert n["fluid"] > 1000
    assert n["res_l2"] < 1e-9 * n["rhs_l2"], n
    # at 40 iterations the same scene leaves a residual far above that
    o2, prm2 = _run_oracle(_dam(iters=40), 1, True)
    n2 = onorms.norms(o2, prm2.comp, sc.open_)
    assert n2["res_l2"] > 1e-6 * n2["rhs_l2"], n2


@pytest.mark.gpu
def test_gpu_norms_match_restatement():
    from splashtank_amd.capi import FlipSimC
    sc = _dam(d=1 / 64)
    print(f"[norms] dam break d=1/64, {sc.n} particles, 5 substeps", flush=True)
    g = FlipSimC(opts_for(sc))
    upload_scene(g, sc)
    prm = capi.make_params(**sc.params)
    with pytest.raises(capi.StError):
        g.solve_norms()  # nothing solved yet
    for _ in range(5):
        g.step(prm)
    got = g.solve_norms()
    ref = onorms.norms(g, prm.comp, sc.open_)
    o, _ = _run_oracle(sc, 5, False)
    ref_o = onorms.norms(o, prm.comp, sc.open_)
    o64, _ = _run_oracle(sc, 5, True)
    ref64 = onorms.norms(o64, prm.comp, sc.open_)
    print(f"[norms] gpu {got}\n[norms] f32 oracle {ref_o}\n[norms] f64 oracle {ref64}", flush=True)
    assert got["fluid"] == ref["fluid"] == ref_o["fluid"]
    for k in ("res_l2", "res_max", "rhs_l2", "div_l2", "div_max", "err_l2", "err_max"):
        # same f32 state, f64 sums in another order
        assert got[k] == pytest.approx(ref[k], rel=1e-10, abs=1e-300), k
        assert got[k] == pytest.approx(ref_o[k], rel=1e-10, abs=1e-300), k
    # the f32 path leaves the divergence error the same solve leaves in f64 (40 iterations
    # dominate it, not the precision)
    assert got["err_l2"] == pytest.approx(ref64["err_l2"], rel=0.2)
    g.close()
