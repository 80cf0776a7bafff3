"""The keyframe map of the differentiable rollout (splashtank_amd.adjoint.
pose_weights, KeyedCollider), pinned by what keyPose (shared.js:608-621) fixes:
the pose holds outside the keys, passes through every key, is linear between
keys without `smooth`, and with it is the Catmull-Rom curve, whose tangent at
key i is half the difference of its neighbours (the end keys repeated) and which
reproduces a path linear in time; the weights sum to one."""
import numpy as np
import pytest

from splashtank_amd.adjoint import KeyedCollider, pose_weights

TIMES = [0.0, 0.1, 0.2, 0.3]


@pytest.mark.parametrize("smooth", [False, True])
def test_holds_outside_and_passes_through_keys(smooth):
    assert np.array_equal(pose_weights(TIMES, -1.0, smooth), [1, 0, 0, 0])
    assert np.array_equal(pose_weights(TIMES, 0.5, smooth), [0, 0, 0, 1])
    for i, t in enumerate(TIMES):
        w = pose_weights(TIMES, t, smooth)
        assert np.allclose(w, np.eye(4)[i], atol=1e-12), (i, w)
    for t in np.linspace(-0.05, 0.35, 41):
        assert abs(pose_weights(TIMES, t, smooth).sum() - 1.0) < 1e-12


def test_linear_between_keys():
    w = pose_weights(TIMES, 0.125, False)
    assert np.allclose(w, [0, 0.75, 0.25, 0], atol=1e-12)


def test_catmull_rom_tangents_and_linear_precision():
    rng = np.random.default_rng(0)
    keys = rng.standard_normal((4, 3))
    e = 1e-5

    def pose(t):
        return pose_weights(TIMES, t, True) @ keys

    for i in range(4):
        lo, hi = max(i - 1, 0), min(i + 1, 3)
        want = (keys[hi] - keys[lo]) / 2 / 0.1  # per unit time, key spacing 0.1
        # second-order one-sided differences inside the segments on either side of key i
        t = TIMES[i]
        if i < 3:
            d = (-3 * pose(t) + 4 * pose(t + e) - pose(t + 2 * e)) / (2 * e)
            assert np.allclose(d, want, rtol=1e-5, atol=1e-5), (i, d, want)
        if i > 0:
            d = (3 * pose(t) - 4 * pose(t - e) + pose(t - 2 * e)) / (2 * e)
            assert np.allclose(d, want, rtol=1e-5, atol=1e-5), (i, d, want)
    # keys on a line at even times: the interior segment is that line
    lin = np.array([[0.0, 1.0, 2.0], [1.0, 1.5, 2.0], [2.0, 2.0, 2.0], [3.0, 2.5, 2.0]])
    for t in np.linspace(0.1, 0.2, 11):
        p = pose_weights(TIMES, t, True) @ lin
        assert np.allclose(p, lin[0] + (lin[1] - lin[0]) * t / 0.1, atol=1e-12)


def test_keyed_collider_velocity_is_the_pose_difference():
    kc = KeyedCollider([4, 4, 4], [0, 0, 0], 0.1, np.zeros(64, np.float32), TIMES, smooth=True)
    keys = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.2, 0.1, 0.0], [0.2, 0.2, 0.1]])
    dt = 0.01
    c, v, wc, wv = kc.at(keys, 0.15, dt)
    c0, _, _, _ = kc.at(keys, 0.15 - dt, dt)
    assert np.allclose(v, (c - c0) / dt, atol=1e-12)
    assert np.allclose(wc @ keys, c) and np.allclose(wv @ keys, v)
    fast = KeyedCollider([4, 4, 4], [0, 0, 0], 0.1, np.zeros(64, np.float32), TIMES, lim=1.0)
    _, v, _, wv = fast.at(keys, 0.15, dt)  # |v| above the limit: a jump, no velocity
    assert not v.any() and not wv.any()
