"""The live view's packing (st_pack_view; packDisp/packSpray/sprayDisp,
src/flip/wgsl-whitewater.js:184-253, packLayout sim.js:376-380).

CPU: the numpy restatement (oracle/pack.py) is pinned by unpacking -- positions
come back within half a 16-bit step, speed bytes are the 16ths of m/s capped at
191 (255 for a dead point), white-water bytes 17 x the level, whitewater codes
carry the kind, the speed in quarters and the direction (octahedral, within
0.4 degrees) or the life left.  GPU: the packed words equal the restatement's
on the GPU sim's own particles and whitewater, bit for bit, with and without a
water sample stride.
"""
import numpy as np
import pytest

from oracle import pack

f32 = np.float32


def _decode_dir(code):
    u = (code & 1023).astype(np.float64) / 1023 * 2 - 1
    w = ((code >> 10) & 1023).astype(np.float64) / 1023 * 2 - 1
    z = 1 - np.abs(u) - np.abs(w)
    neg = z < 0
    u2 = np.where(neg, (1 - np.abs(w)) * np.sign(u + 1e-30), u)
    w2 = np.where(neg, (1 - np.abs(u)) * np.sign(w + 1e-30), w)
    n = np.stack([u2, w2, z], axis=1)
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def test_pack_unpack_round_trip():
    rng = np.random.default_rng(3)
    lo, hi = np.array([-0.5, 0.0, -0.5]), np.array([0.5, 1.0, 0.5])
    k = (65535.0 / (hi - lo)).astype(f32)
    n = 1001
    water = np.zeros((n, 4), f32)
    water[:, :3] = rng.uniform(lo, hi, (n, 3)).astype(f32)
    speed = rng.uniform(0, 14, n).astype(f32)
    level = rng.integers(0, 16, n)
    water[:, 3] = speed + f32(64) * level
    water[7, 1] = f32(-1.0e6)                                   # a dead particle
    m = 37
    sv = rng.normal(0, 2, (m, 3)).astype(f32)
    kinds = rng.integers(0, 4, m).astype(np.uint32)
    life = rng.uniform(0, 1, m).astype(f32)
    codes = pack.pack_spray(sv, kinds, life)
    spray = np.zeros((m, 4), f32)
    spray[:, :3] = rng.uniform(lo, hi, (m, 3)).astype(f32)
    spray[:, 3] = codes.view(f32)
    L = pack.layout(n, 64)
    words = pack.pack(L, water, spray, lo.astype(f32), k, 1.0)
    u = pack.unpack(words, L)
    alive = np.ones(n, bool)
    alive[7] = False
    back = lo + u["pos"][alive].astype(np.float64) / k
    assert np.abs(back - water[alive, :3]).max() <= 0.5 / k.min() + 2e-7   # half a step, plus f32 rounding of (p - lo) k
    np.testing.assert_array_equal(u["spd"][alive], np.minimum(np.floor(speed[alive].astype(np.float64) * 16), 191))
    assert u["spd"][7] == 255 and u["fm"][7] == 0
    np.testing.assert_array_equal(u["fm"][alive], 17 * level[alive])
    c = u["sp_code"]
    assert len(c) == m and np.all(c >> 31 == 1)
    kind_bits = (c >> 29) & 3
    np.testing.assert_array_equal(kind_bits, np.array([0, 2, 1, 3])[kinds])       # 00 spray, 10 mist, 01 foam, 11 bubble
    spd = np.sqrt((sv.astype(np.float64) ** 2).sum(1))
    np.testing.assert_array_equal((c >> 21) & 255, np.minimum(np.floor(spd * 4), 255))
    flying = kinds < 2
    ang = np.degrees(np.arccos(np.clip((_decode_dir(c[flying]) * (sv[flying] / spd[flying, None])).sum(1), -1, 1)))
    assert ang.max() < 0.4
    np.testing.assert_array_equal(c[~flying] & 1023, np.floor(life[~flying].astype(np.float64) * 1023))


@pytest.mark.gpu
@pytest.mark.parametrize("disp_max", [0, 7000])
def test_gpu_pack_matches_restatement(disp_max):
    from splashtank_amd import capi
    from tests.test_oracle_pins import lattice
    from tests.test_gpu_spray import SPR
    d = 1 / 64
    shear = lambda p: np.stack([0.0 * p[:, 0], 3.0 * np.sin(2 * np.pi * p[:, 0]), 0.0 * p[:, 0]], axis=1)  # noqa: E731
    pos, vel = lattice(d, (-0.5, 0.0, -0.5), (0.5, 0.2, 0.5), velfn=shear)
    g = capi.FlipSimC(capi.make_opts(d, (1, 1, 1), max_n=len(pos)))
    g.upload(pos, vel)
    g.set_spray_cap(1 << 14)
    prm = capi.make_params(dt=1 / 60, g=9.81, iters=20, **dict(SPR, vs=1.25))
    for _ in range(12):
        g.step(prm)
    L, words = g.pack_view(disp_max)
    print(f"[view] disp_max {disp_max}: {L}", flush=True)
    n = len(pos)
    stride = max(1, -(-n // (disp_max or 300000)))
    assert L["stride"] == (stride if not L["from_pos"] else 1) and L["nS"] == 1 << 14
    water = g.read_particles()[::L["stride"]][:L["nW"]]
    spos, svel = g.read_spray(1 << 14)
    handed = len(spos) if len(spos) > 1 else 0
    spray = pack.spray_sample(spos, svel, handed, 1 << 14, min(1 << 14, 200000), prm.ww_foam_life or 3.0)
    dm = g.dims
    lo = np.array([dm.x0, dm.y0, dm.z0], f32)
    k = np.array([65535.0 / (float(dm.x1) - float(dm.x0)), 65535.0 / (float(dm.y1) - float(dm.y0)),
                  65535.0 / (float(dm.z1) - float(dm.z0))]).astype(f32)
    want = pack.pack(pack.layout(L["nW"], L["nS"]), water, spray, lo, k, 1.25)
    assert words[0] == len(spray) > 100
    bad = np.flatnonzero(words != want)
    assert bad.size == 0, f"{bad.size} of {words.size} words differ, first at {bad[:5]}"
    g.close()
