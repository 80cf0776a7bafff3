"""Scene driving on the JS facade (splashtank_amd/flip/runner.js): SceneRunner's
collider velocities, emitters and keyframed poses against a recording stand-in
sim (CPU), and a short bake through FlipSim on the GPU."""
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "splashtank_amd", "flip", "splashtank_hip.node")
NODE = shutil.which("node")
needs_node = pytest.mark.skipif(NODE is None, reason="node is missing")


def _run(script):
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", script)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


@needs_node
def test_runner_collider_velocities_emitters_keyframes():
    o = _run("runner_unit.js")
    # no previous pose on the first step: no velocity; then the pose difference
    assert o["firstVel"] == [0, 0]
    assert abs(o["vel"][0] - 0.5) < 1e-9 and o["vel"][1] == 0 and o["vel"][2] == 0
    # 2 rad/s about y from R Rprev^T: sin(w dt) / dt
    assert abs(o["omega"][1] - o["omegaExpect"]) < 1e-12 and abs(o["omega"][0]) < 1e-12 and abs(o["omega"][2]) < 1e-12
    assert o["jumpVel"] == 0  # 120 m/s between steps is a jump, not a motion
    # a smooth emitter releases one lattice disc per spacing travelled, at the emitter speed
    assert o["emitted"] == o["emittedExpect"] > 0
    assert abs(o["emitVel"][1] + 2) < 1e-12
    assert o["keyHeld"] == [0, 3] and o["keyLinear"] == 2
    assert o["keySpline"][0] == 1 and o["keySpline"][1] == 3 and abs(o["keySpline"][2] - 2.0625) < 1e-12
    assert abs(o["objSpeed"] - 1.0) < 1e-9


@needs_node
@pytest.mark.gpu
@pytest.mark.skipif(not os.path.isfile(ADDON), reason="the N-API addon is not built")
def test_bake_drives_flipsim():
    o = _run("bake_run.js")
    assert o["ok"]
    fr = o["frames"]
    assert len(fr) == 5 and fr[0]["sub"] == 0
    assert all(f["finite"] and f["inside"] for f in fr)
    # the emitter and the dropped ball add water; substeps follow the speeds
    assert fr[-1]["n"] > fr[0]["n"]
    assert all(2 <= f["sub"] <= 12 for f in fr[1:])
    assert any(f["objectSpeed"] > 0.4 for f in fr[1:])
    # the frames (packed as the page's cache holds them) written, read back unchanged, inside the box
    assert o["cacheRoundTrip"] and o["cacheBytes"] > 0
    assert o["worstQuant"] == 0


@needs_node
def test_splc_cache_layout():
    """The SPLC cache (cache.js) on hand-made frames: version 2 bytes at fixed
    offsets, the round trip, and a hand-built version 1 file."""
    o = _run("cache_unit.js")
    assert o["n"] == 2  # the two dead slots after the last water point are dropped
    assert o["pos"] == [32768, 32768, 49151, 0, 65535, 65535]
    assert o["spd"] == [40, 191] and o["fm"] == [51, 0]  # 2.5 m/s in 16ths; 63.5 capped at 191; foam level 3 * 17
    assert o["objs"] == [1, 2, 3, 0, 0.5, 0]
    assert o["magic"] == "SPLC" and o["version"] == 2 and o["jsonLen"][0] == o["jsonLen"][1]
    assert o["frames"] == 2 and o["frame0"] == {"n": 2, "flags": 1, "yaw": 0.5, "pos0": 32768, "spd1": 191, "fm0": 51}
    assert o["length"][0] == o["length"][1] and o["round"]
    assert o["v1"] == {"fps": 30, "n": 2, "yaw": -0.75, "pos": [1, 2, 3, 4, 5, 6], "spd": [33, 44], "ids": [9]}


@needs_node
def test_cache_writers_match_reference_golden():
    """packFrame (water, whitewater codes and deep points of a readback),
    packDeepFrame, packSprayFrame, packSprayCode and writeCache give, byte for
    byte, what the reference's packFrameBuf, packDeep, packSprayFrame,
    packSprayJS and writeCacheParts wrote from the same inputs
    (tests/golden/make_cache_golden.js), including the spray (flag 2) and
    deep-water (flag 4) sections; readCache reads them back."""
    o = _run("cache_golden.js")
    assert all(o[k] for k in ("frame", "sprayFull", "sprayCapped", "codes", "deep", "file", "readBack")), o
    assert o["fileBytes"][0] == o["fileBytes"][1]


@needs_node
def test_scene_driving_matches_reference_golden():
    """keyPose, sampleState (keys, recorded path, spin), objectSpeed,
    SceneRunner.step (collider velocities and spins from pose changes, jumps
    past colMax, smooth and random emitters with spread and turbulence),
    bakeDiag and nbSeed (boxes on and off the walls, a sphere, the ocean, a
    full pool) give exactly what the reference's own functions gave on the same
    inputs and random draws (tests/golden/make_runner_golden.js)."""
    o = _run("runner_golden.js")
    parts = ("keyPose", "sampleState", "objectSpeed", "runnerSteps", "runnerAdds", "bakeDiag", "nbSeed", "nbSeedDigest")
    assert all(o.get(k) is True for k in parts), o
    assert o["runnerCounts"][0] == o["runnerCounts"][1] > 100
    assert not o.get("nbSeedSkipped"), "the ocean case needs the N-API addon (host code only)"
