"""Comparison helpers shared by the parity tests."""
from __future__ import annotations

import numpy as np

from splashtank_amd import capi

# Tolerance for f32 fields, from the stated bar: max relative error <= 1e-5 with an
# absolute floor for values near zero.  The kernels are written to give the same
# bits as the oracle, so the tests also count values that are not bit-identical.
F32_RTOL = 1e-5
F32_ATOL = 1e-6


def f32_report(a, b):
    a = np.asarray(a, np.float32).reshape(-1)
    b = np.asarray(b, np.float32).reshape(-1)
    assert a.shape == b.shape, (a.shape, b.shape)
    same = np.sum(a.view(np.uint32) == b.view(np.uint32))
    both_nan = np.isnan(a) & np.isnan(b)
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    d[both_nan] = 0.0
    scale = np.maximum(np.abs(b.astype(np.float64)), F32_ATOL / F32_RTOL)
    rel = d / scale
    rel[both_nan] = 0.0
    one_nan = np.isnan(a) ^ np.isnan(b)
    rel[one_nan] = np.inf
    return {"n": int(a.size), "bit_identical": int(same), "max_abs": float(d.max(initial=0.0)),
            "max_rel": float(rel.max(initial=0.0))}


def assert_f32_close(a, b, what=""):
    r = f32_report(a, b)
    assert r["max_rel"] <= F32_RTOL, f"{what}: {r}"
    return r


def assert_int_equal(a, b, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    bad = np.flatnonzero(a.reshape(-1) != b.reshape(-1))
    assert bad.size == 0, f"{what}: {bad.size} of {a.size} differ, first at {bad[:5]}: {a.reshape(-1)[bad[:5]]} vs {b.reshape(-1)[bad[:5]]}"


def upload_scene(sim, sc, t=0.0):
    sim.upload(sc.pos, sc.vel)
    if sc.colliders:
        cf, sdf, sig = sc.collider_state(t, sc.params["dt"])
        sim.set_colliders(cf, sdf, sig)


def set_colliders_at(sim, sc, t):
    if sc.colliders:
        cf, sdf, sig = sc.collider_state(t, sc.params["dt"])
        sim.set_colliders(cf, sdf, sig)


def opts_for(sc, **kw):
    return capi.make_opts(sc.d, sc.size, ratio=sc.ratio, open_=sc.open_, max_n=kw.pop("max_n", sc.n),
                          slots=kw.pop("slots", sc.slots), **kw)
