"""The collider distance-grid builder (st_build_sdf) against the reference's
buildSDF (src/shared.js:98-181).

Golden vectors: tests/golden/sdf_golden.{json,bin}, written by
tests/golden/make_sdf_golden.js from the reference's own buildSDF on three of
its own meshes (cube, sphere, cylinder) at the page's collider settings; the
library must give the same f32 grid bit for bit.  Beside them, what the
mathematics fixes for the cube: exact distances to the box near the surface,
the sign from the box, and chamfer values within a few percent further out.
Host code only: no GPU is needed.
"""
import json
import os

import numpy as np
import pytest

from splashtank_amd import capi

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cases():
    meta = json.load(open(os.path.join(GOLD, "sdf_golden.json")))
    blob = np.fromfile(os.path.join(GOLD, "sdf_golden.bin"), dtype="<f4")
    return [(c, blob[c["offset"]:c["offset"] + c["count"]]) for c in meta["cases"]]


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0]["name"])
def test_matches_reference_golden(case):
    c, want = case
    got = capi.build_sdf(np.array(c["v"]), np.array(c["tris"]).reshape(-1, 3), c["cell"], c["margin"])
    assert got["dims"] == c["dims"]
    np.testing.assert_array_equal(np.array(got["origin"]), np.array(c["origin"]))
    assert got["cell"] == c["sdfCell"]
    bad = np.flatnonzero(got["data"].view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, f"{bad.size} of {want.size} differ, first {bad[:5]}: {got['data'][bad[:5]]} vs {want[bad[:5]]}"


def test_cube_against_the_analytic_box():
    c = [x for x, _ in _cases() if x["name"] == "cube"][0]
    g = capi.build_sdf(np.array(c["v"]), np.array(c["tris"]).reshape(-1, 3), 0.025, 0.16)
    nx, ny, nz = g["dims"]
    d = g["data"].reshape(nz, ny, nx)
    ax = [g["origin"][k] + np.arange(g["dims"][k]) * g["cell"] for k in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    q = np.stack([np.abs(X) - 0.1, np.abs(Y) - 0.1, np.abs(Z) - 0.1])
    exact = np.linalg.norm(np.maximum(q, 0), axis=0) + np.minimum(q.max(axis=0), 0)   # box SDF, half size 0.1
    band = np.abs(exact) <= 2.0 * g["cell"]
    assert band.sum() > 1000
    np.testing.assert_allclose(d[band], exact[band], atol=2e-7)           # exact near the surface
    off = np.abs(exact) > 2.0 * g["cell"]
    assert np.all(np.sign(d[off]) == np.sign(exact[off]))                 # inside negative, outside positive
    rel = np.abs(np.abs(d[off]) - np.abs(exact[off])) / np.abs(exact[off])
    assert rel.max() < 0.10                                               # chamfer: within a few percent


def test_rejects_bad_meshes():
    with pytest.raises(capi.StError):
        capi.build_sdf(np.zeros((3, 3)), np.array([[0, 1, 5]]))


def test_node_facade_builds_the_same_grids():
    import shutil
    import subprocess
    node = shutil.which("node")
    addon = os.path.join(os.path.dirname(GOLD), "..", "splashtank_amd", "flip", "splashtank_hip.node")
    if node is None or not os.path.isfile(addon):
        pytest.skip("node or the N-API addon is missing")
    r = subprocess.run([node, os.path.join(os.path.dirname(GOLD), "js", "sdf_check.js")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert set(res) == {"cube", "sphere", "cylinder"}
    assert all(v["bad"] == 0 for v in res.values()), res
    assert [res[k]["id"] for k in ("cube", "sphere", "cylinder")] == [1, 2, 3]   # ids as the reference's SDF_ID
