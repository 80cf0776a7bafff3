"""Drip samples (dripSamples, src/flip/bake.js:19-40) against golden vectors
written by the reference's own function (tests/golden/make_drip_golden.js) on
the reference's own buildSDF grids: the Python helper (scenes.drip_samples)
and the JS facade's (dripSamples in splashtank_amd/flip/sdf.js) must give the
same f32 samples bit for bit."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from splashtank_amd import scenes

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _grids():
    meta = json.load(open(os.path.join(GOLD, "sdf_golden.json")))
    blob = np.fromfile(os.path.join(GOLD, "sdf_golden.bin"), dtype="<f4")
    return {c["name"]: dict(id=q + 1, dims=c["dims"], origin=c["origin"], cell=c["sdfCell"],
                            data=blob[c["offset"]:c["offset"] + c["count"]]) for q, c in enumerate(meta["cases"])}


def _cases():
    meta = json.load(open(os.path.join(GOLD, "drip_golden.json")))
    blob = np.fromfile(os.path.join(GOLD, "drip_golden.bin"), dtype="<f4")
    return [(c, blob[c["offset"]:c["offset"] + c["count"] * 8].reshape(-1, 8)) for c in meta["cases"]]   # offsets count floats


@pytest.mark.parametrize("case", _cases(), ids=lambda c: f'{c[0]["grid"]}-h{c[0]["h"]:.4f}')
def test_python_drip_samples_match_reference(case):
    c, want = case
    got = scenes.drip_samples(_grids()[c["grid"]], c["h"], c["k"])
    assert got.shape == want.shape
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_js_drip_samples_match_reference():
    node = shutil.which("node")
    if node is None:
        pytest.skip("node is missing")
    r = subprocess.run([node, os.path.join(os.path.dirname(GOLD), "js", "drip_check.js")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert len(res) == 6 and all(v == 0 for v in res.values()), res
