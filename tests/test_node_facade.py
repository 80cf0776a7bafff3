"""The JS FlipSim facade (Node, N-API addon) over the same C ABI.

CPU: the addon loads in Node and fails loudly when no GPU is present.
GPU: Node drives FlipSim.add / step / stepTimed / readbackInto on a scene and
the Python C-ABI driver runs the same particles and settings; the two results
must be bit-identical (same library, same kernels).
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from splashtank_amd import capi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "splashtank_amd", "flip", "splashtank_hip.node")
NODE = shutil.which("node")

needs_node = pytest.mark.skipif(NODE is None or not os.path.isfile(ADDON), reason="node or the N-API addon is missing")


@needs_node
def test_addon_loads_and_reports_missing_gpu():
    code = ("const m = require(process.argv[1]); const d = m.initHip(0); console.log(d.version);"
            "try { new m.FlipSim(d, {d: 1/32, maxN: 8, size: [1,1,1]}); console.log('GPU'); }"
            "catch (e) { console.log('ERR ' + e.message); }")
    r = subprocess.run([NODE, "-e", code, os.path.join(ROOT, "splashtank_amd", "flip")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[0].startswith("splashtank-mi355x")
    if lines[1] != "GPU":
        assert lines[1].startswith("ERR ") and "device" in lines[1]


@needs_node
@pytest.mark.gpu
def test_facade_matches_c_abi(tmp_path):
    sc = scenes.dam_break(d=1.0 / 48, frac=(0.4, 0.6, 1.0))
    steps = 5
    parts = np.concatenate([sc.pos, sc.vel], axis=1).astype(np.float32)
    (tmp_path / "p.bin").write_bytes(parts.tobytes())
    cfg = dict(particles=str(tmp_path / "p.bin"), out=str(tmp_path / "out.bin"), d=sc.d, size=list(sc.size), open=sc.open_,
               ratio=sc.ratio, steps=steps, params=dict(sc.params))
    (tmp_path / "cfg.json").write_text(json.dumps(cfg))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "facade_run.js"), str(tmp_path / "cfg.json")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["n"] == sc.n and info["stages"][0] == "blocks" and "pressure" in info["stages"]
    js = np.frombuffer((tmp_path / "out.bin").read_bytes(), np.float32).reshape(-1, 4)
    # the same run through the C ABI: the facade's stepTimed is one more full step
    g = capi.FlipSimC(capi.make_opts(sc.d, sc.size, ratio=sc.ratio, open_=sc.open_, max_n=sc.n))
    g.upload(sc.pos, sc.vel)
    prm = capi.make_params(**sc.params)
    for _ in range(steps + 1):
        g.step(prm)
    ref = g.read_particles()
    np.testing.assert_array_equal(js.view(np.uint32), ref.view(np.uint32))
    assert info["stats"]["sc"] == int(g.ctl()[4])
    # the FP64 norms: same state, same fixed reduction order, so the same doubles
    nrm = g.solve_norms()
    assert info["norms"]["resL2"] == nrm["res_l2"] and info["norms"]["errMax"] == nrm["err_max"]
    assert info["norms"]["fluid"] == nrm["fluid"] > 0
    # the live view: the facade's pump() frame against the C ABI's packed words
    from oracle import pack
    L, words = g.pack_view()
    u = pack.unpack(words, L)
    view = info["view"]
    assert view["n"] == L["nW"] == sc.n
    assert view["pos"] == int(u["pos"].astype(np.int64).sum()) and view["spd"] == int(u["spd"].astype(np.int64).sum())
    assert view["fm"] == int(u["fm"].astype(np.int64).sum())


@needs_node
@pytest.mark.gpu
def test_facade_batch_matches_c_abi(tmp_path):
    """FlipBatch in Node over three tanks of different sizes and SOR iteration
    counts against each tank stepped alone through the C ABI, bit for bit."""
    tanks = scenes.batched_tanks(count=3, seed=7, nmin=16, nmax=32, iters=10)
    iters = [10, 7, 16]
    steps = 4
    cfgs = []
    for k, sc in enumerate(tanks):
        parts = np.concatenate([sc.pos, sc.vel], axis=1).astype(np.float32)
        (tmp_path / f"p{k}.bin").write_bytes(parts.tobytes())
        cfgs.append(dict(particles=str(tmp_path / f"p{k}.bin"), out=str(tmp_path / f"out{k}.bin"), d=sc.d,
                         size=list(sc.size), open=sc.open_, ratio=sc.ratio, params=dict(sc.params, iters=iters[k])))
    (tmp_path / "cfg.json").write_text(json.dumps(dict(tanks=cfgs, steps=steps)))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "batch_run.js"), str(tmp_path / "cfg.json")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    for k, sc in enumerate(tanks):
        js = np.frombuffer((tmp_path / f"out{k}.bin").read_bytes(), np.float32).reshape(-1, 4)
        g = capi.FlipSimC(capi.make_opts(sc.d, sc.size, ratio=sc.ratio, open_=sc.open_, max_n=sc.n))
        g.upload(sc.pos, sc.vel)
        prm = capi.make_params(**dict(sc.params, iters=iters[k]))
        for _ in range(steps):
            g.step(prm)
        assert info["n"][k] == sc.n
        np.testing.assert_array_equal(js.view(np.uint32), g.read_particles().view(np.uint32))
        assert info["stats"][k] == int(g.ctl()[4])
        g.close()


@needs_node
@pytest.mark.gpu
def test_facade_extensions(tmp_path):
    """stepN (with and without graph replay), the autotuner, snapshots and the
    adjoint through Node: stepN gives the bits of step() one at a time, the
    tuner and a snapshot round trip leave the state as it was, and the adjoint's
    gradient equals the C ABI's (capi.AdjointC) to float rounding: the backward
    sums face gradients with float atomics, whose order varies from run to run
    (two runs differ by at most a few ulps, 2.4e-7 relative seen)."""
    sc = scenes.dam_break(d=1.0 / 32, iters=15)  # 16^3 cells, closed box
    steps = 4
    parts = np.concatenate([sc.pos, sc.vel], axis=1).astype(np.float32)
    (tmp_path / "p.bin").write_bytes(parts.tobytes())
    cfg = dict(particles=str(tmp_path / "p.bin"), out=str(tmp_path / "g.bin"), d=sc.d, size=list(sc.size), open=sc.open_,
               ratio=sc.ratio, steps=steps, params=dict(sc.params))
    (tmp_path / "cfg.json").write_text(json.dumps(cfg))
    print(f"[facade] extensions, {sc.n} particles, {steps} substeps", flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "ext_run.js"), str(tmp_path / "cfg.json")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["stepN"] and info["stepNGraph"] and info["afterTune"] and info["restore"], info
    assert info["tune"]["candidates"] >= 3
    js = np.frombuffer((tmp_path / "g.bin").read_bytes(), np.float32).reshape(2, -1, 4)
    g = capi.FlipSimC(capi.make_opts(sc.d, sc.size, ratio=sc.ratio, open_=sc.open_, max_n=sc.n))
    try:
        g.upload(sc.pos, sc.vel)
        a = capi.AdjointC(g, steps)
        prm = capi.make_params(**sc.params)
        for _ in range(steps):
            a.record(prm)
        gp = np.zeros((sc.n, 4), np.float32)
        gp[:, 0], gp[:, 1], gp[:, 2] = 1, 2, 3
        a.seed(gp, np.zeros_like(gp))
        for _ in range(steps):
            a.backward()
        gpos, gvel = a.grad()
        a.close()
    finally:
        g.close()
    for got, want in ((js[0][:, :3], gpos[:, :3]), (js[1][:, :3], gvel[:, :3])):
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6 * float(np.abs(want).max()))
    assert np.abs(gvel[:, :3]).max() > 0


@needs_node
@pytest.mark.gpu
def test_async_readback_keeps_the_event_loop_running(tmp_path):
    """readbackInto returns a Promise at once; the event loop turns while the
    copy is pending; the result is the state at the call although substeps were
    enqueued right after it; readStats and readSpray resolve what the blocking
    reads give."""
    sc = scenes.dam_break(d=1.0 / 128, iters=40)  # the 64^3 case: 502,656 particles
    parts = np.concatenate([sc.pos, sc.vel], axis=1).astype(np.float32)
    (tmp_path / "p.bin").write_bytes(parts.tobytes())
    cfg = dict(particles=str(tmp_path / "p.bin"), d=sc.d, size=list(sc.size), open=sc.open_, ratio=sc.ratio,
               steps=20, after=10, sprayCap=20000, sprSteps=6, params=dict(sc.params))
    (tmp_path / "cfg.json").write_text(json.dumps(cfg))
    print(f"[facade] async readback, {sc.n} particles", flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "async_run.js"), str(tmp_path / "cfg.json")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout.strip().splitlines()[-1])
    print(o)
    assert o["pendingAfterCall"] and o["loopTurnsWhilePending"] >= 1
    assert o["resolvedN"] == sc.n and o["stateAtCall"]
    assert o["stats"] and o["spray"]


@needs_node
@pytest.mark.gpu
def test_facade_pool_prime_and_growth(tmp_path):
    """A facade sim with a 6-slot pool and opt.sort on: prime() grows the pool
    until the first step's tiles fit with the reference's headroom, maybeGrow
    keeps it ahead while stepping, and the readback equals a sim with every tile
    pooled and opt.sort off, bit for bit (the pool's size and the particle
    order in memory change nothing of the result)."""
    sc = scenes.dam_break(d=1.0 / 48)
    parts = np.concatenate([sc.pos, sc.vel], axis=1).astype(np.float32)
    (tmp_path / "p.bin").write_bytes(parts.tobytes())
    cfg = dict(particles=str(tmp_path / "p.bin"), d=sc.d, size=list(sc.size), open=sc.open_, ratio=sc.ratio,
               steps=12, params=dict(sc.params))
    (tmp_path / "cfg.json").write_text(json.dumps(cfg))
    print(f"[facade] pool growth, {sc.n} particles", flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "pool_run.js"), str(tmp_path / "cfg.json")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout.strip().splitlines()[-1])
    print(o)
    assert o["slotsAtStart"] == 6 and o["slotsAfterPrime"] > 6
    assert o["primeStats"]["ovf"] == 0 and o["primeStats"]["sc"] <= 0.8 * o["slotsAfterPrime"]
    assert o["ovfTotal"] == 0 and o["slotsAtEnd"] < o["fullPool"]
    assert o["same"]
