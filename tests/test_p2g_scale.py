"""p2g rounds each corner's momentum as round(w * vc * scale) (wgsl-solver.js:
33, the product taken left to right in f32).  The HIP kernel forms vcs =
vc * scale once per particle and axis and rounds w * vcs instead, when the
scale is a power of two in 2^0 .. 2^20 and |vcs| < 2147483520 (it also skips the
saturating clamp there, since w is in [0, 1]), and there forms the scaled weight
wS = (fx * fy) * (fz * scale) once for the weight sum, rint(wS), and the
momentum sum, rint(wS * vc).  These tests check on the CPU, in
IEEE f32 (numpy), that the two products round to the same integer for weights
and velocities spread over the whole exponent range, subnormals and zeros
included, and that the rewrite is wrong for a scale that is not a power of two
(so the kernel's condition is needed)."""
import numpy as np
import pytest

F = np.float32


def _inputs(rng, scale, n=1 << 21):
    w = rng.random(n, dtype=np.float32)
    # weights as p2g makes them: products of three fractions in [0, 1], plus the ends
    w = (w * rng.random(n, dtype=np.float32) * rng.random(n, dtype=np.float32)).astype(F)
    w[:64] = [0.0, 1.0, np.finfo(F).tiny, np.float32(1e-45)] * 16
    # velocities: sign, exponent spread over the range where |vc * scale| < 2^31 - 128
    e = rng.uniform(-149.0, np.log2(2147483520.0 / scale) - 1e-3, n)
    vc = (np.sign(rng.standard_normal(n)) * np.exp2(e)).astype(F)
    vc[64:128] = [0.0, -0.0, np.float32(1e-45), -np.float32(1e-45)] * 16
    return w, vc


def _written(w, vc, scale):
    return np.rint((w * vc) * F(scale))


def _kernel(w, vc, scale):
    return np.rint(w * (vc * F(scale)))


def _fractions(rng, n):
    """three fractions per corner as win_axis gives them (f or 1 - f), with
    exact zeros, ones and tiny values among them"""
    f = rng.random((3, n), dtype=np.float32)
    f[:, :48] = np.array([0.0, 1.0, 1e-20, np.finfo(F).tiny, 2.0 ** -60, 0.5], F).repeat(8)[None, :]
    f[1:, 48:96] = F(2.0 ** -50)  # products below the normal range
    return f


@pytest.mark.parametrize("scale", [1.0, 2.0, 65536.0, 1048576.0])
def test_scaled_weight_serves_both_sums(scale):
    """wS = (fx * fy) * (fz * scale): rint(wS) against rint(((fx * fy) * fz) * scale)
    and rint(wS * vc) against rint(((fx * fy) * fz * vc) * scale)"""
    rng = np.random.default_rng(int(scale) + 19)
    n = 1 << 21
    fx, fy, fz = _fractions(rng, n)
    _, vc = _inputs(rng, scale, n)
    w = (fx * fy) * fz
    wS = (fx * fy) * (fz * F(scale))
    assert np.array_equal(np.rint(wS).astype(np.int64), np.rint(w * F(scale)).astype(np.int64))
    assert np.array_equal(np.rint(wS * vc).astype(np.int64), _written(w, vc, scale).astype(np.int64))


@pytest.mark.parametrize("scale", [1.0, 2.0, 65536.0, 1048576.0])
def test_power_of_two_scale_rounds_the_same(scale):
    rng = np.random.default_rng(int(scale) + 7)
    w, vc = _inputs(rng, scale)
    vcs = vc * F(scale)
    assert np.all(np.abs(vcs) < F(2147483520.0))
    a, b = _written(w, vc, scale), _kernel(w, vc, scale)
    # equal as integers (a signed zero converts to 0 either way)
    assert np.array_equal(a.astype(np.int64), b.astype(np.int64))
    # the clamp at 2147483520 cannot act: |w * vcs| <= |vcs|
    assert np.all(np.abs(b) <= F(2147483520.0))


def test_other_scales_need_the_written_product():
    rng = np.random.default_rng(3)
    w, vc = _inputs(rng, 65535.0)
    a, b = _written(w, vc, 65535.0), _kernel(w, vc, 65535.0)
    assert np.count_nonzero(a != b) > 0
