"""TEST INFRASTRUCTURE: the live view's packing (packLayout, sim.js:376-380;
packDisp and packSpray, src/flip/wgsl-whitewater.js:184-253; sprayDisp 201-211)
restated with numpy, and its unpacking (unpackLive, sim.js:526-537).

Only tests/ use it, as the checker of st_pack_view.  Arithmetic is f32 where
the WGSL's is: positions quantised as round-half-even((p - lo) * k), k =
f32(65535 / extent); the speed byte clamp((w - 64 floor(w / 64)) 16 vs, 0, 191);
the white-water byte 17 clamp(floor(w / 64), 0, 15).
"""
from __future__ import annotations

import numpy as np

f32 = np.float32
DEADY = f32(-1.0e5)


def layout(nW, nS, nD=0):
    gW, gS, gD = (nW + 3) // 4, (nS + 3) // 4, (nD + 3) // 4
    L = {"nW": nW, "nS": nS, "nD": nD, "oWp": 4}
    L["oWs"] = L["oWp"] + gW * 6
    L["oWf"] = L["oWs"] + gW
    L["oSp"] = L["oWf"] + gW
    L["oSk"] = L["oSp"] + gS * 6
    L["oDp"] = L["oSk"] + gS * 4
    L["oDs"] = L["oDp"] + gD * 6
    L["words"] = L["oDs"] + gD
    return L


def _u32sat(x):
    x = np.asarray(x, np.float32)
    out = np.where(np.isnan(x) | (x <= 0), 0.0, np.minimum(x, f32(4294967040.0)))
    return np.trunc(out).astype(np.uint64).astype(np.uint32)


def q16(p, lo, k):
    """(n, 3) f32 positions -> (n, 3) u32 in 0..65535"""
    v = (p.astype(f32) - lo.astype(f32)) * k.astype(f32)
    return _u32sat(np.clip(np.rint(v), f32(0), f32(65535)))


def pack_spray(v, kind, life_frac):
    """packSpray: velocities (n, 3) f32, kinds (n,) u32, life fractions (n,) f32 -> (n,) u32 codes"""
    v = v.astype(f32)
    s = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]).astype(f32)
    code = np.uint32(0x80000000) | (np.minimum(_u32sat(s * f32(4)), 255).astype(np.uint32) << np.uint32(21))
    code = code | np.select([kind == 1, kind == 2, kind == 3], [np.uint32(0x40000000), np.uint32(0x20000000), np.uint32(0x60000000)],
                            np.uint32(0)).astype(np.uint32)
    foamish = kind >= 2
    life_code = _u32sat(np.clip(life_frac.astype(f32), f32(0), f32(1)) * f32(1023))
    n = np.where((s > f32(1e-6))[:, None], v / np.where(s > 0, s, f32(1))[:, None], np.array([0, 1, 0], f32)).astype(f32)
    l1 = ((np.abs(n[:, 0]) + np.abs(n[:, 1])) + np.abs(n[:, 2])).astype(f32)
    px, py = (n[:, 0] / l1).astype(f32), (n[:, 1] / l1).astype(f32)
    neg = n[:, 2] < 0
    qx = ((f32(1) - np.abs(py)) * np.where(px >= 0, f32(1), f32(-1))).astype(f32)
    qy = ((f32(1) - np.abs(px)) * np.where(py >= 0, f32(1), f32(-1))).astype(f32)
    px, py = np.where(neg, qx, px), np.where(neg, qy, py)
    u = _u32sat(np.clip(np.rint((px * f32(0.5) + f32(0.5)) * f32(1023)), f32(0), f32(1023)))
    w = _u32sat(np.clip(np.rint((py * f32(0.5) + f32(0.5)) * f32(1023)), f32(0), f32(1023)))
    return np.where(foamish, code | life_code, code | (w << np.uint32(10)) | u).astype(np.uint32)


def spray_sample(spos, svel, handed, smax, sd_max, foam_life):
    """sprayDisp: the living slots of every stride-th slot handed out, in slot order -> (n, 4) f32 with the code in w"""
    used = min(handed, smax)
    stride = max(1, (used + sd_max - 1) // sd_max)
    j = np.arange(sd_max, dtype=np.int64) * stride
    j = j[j < used]
    j = j[spos[j, 3] > 0]
    out = np.zeros((len(j), 4), f32)
    out[:, :3] = spos[j, :3]
    kinds = _u32sat(svel[j, 3].astype(f32) + f32(0.5))
    codes = pack_spray(svel[j, :3], kinds, (spos[j, 3].astype(f32) / max(f32(foam_life), f32(0.01))).astype(f32))
    out[:, 3] = codes.view(np.float32)
    return out


def pack(L, water, spray, lo, k, vs):
    """packDisp: water (nW, 4) f32 (y < -1e5 dead), spray (<= nS, 4) f32 -> words (u32)"""
    out = np.zeros(L["words"], np.uint32)
    nW, nS = L["nW"], L["nS"]
    ns = min(len(spray), nS)
    out[0] = ns
    gW = (nW + 3) // 4
    w = np.zeros((gW * 4, 4), f32)
    w[:, 1] = f32(-1.0e6)
    w[:nW] = water[:nW]
    alive = w[:, 1] > DEADY
    q = np.where(alive[:, None], q16(w[:, :3], lo, k), 0).astype(np.uint32)
    lv = np.floor(w[:, 3] / f32(64)).astype(f32)
    s = np.where(alive, _u32sat(np.clip((w[:, 3] - lv * f32(64)) * f32(16) * f32(vs), f32(0), f32(191))), 255).astype(np.uint32)
    f = np.where(alive, _u32sat(np.clip(lv, f32(0), f32(15))) * 17, 0).astype(np.uint32)
    out[L["oWp"]:L["oWp"] + gW * 6] = _put6(q)
    sh = np.uint32(8) * (np.arange(gW * 4, dtype=np.uint32) % 4)
    out[L["oWs"]:L["oWs"] + gW] = (s << sh).reshape(gW, 4).sum(axis=1, dtype=np.uint64).astype(np.uint32)
    out[L["oWf"]:L["oWf"] + gW] = (f << sh).reshape(gW, 4).sum(axis=1, dtype=np.uint64).astype(np.uint32)
    gS = (nS + 3) // 4
    sp = np.zeros((gS * 4, 4), f32)
    sp[:, 1] = f32(-1.0e6)
    sp[:ns] = spray[:ns]
    sal = sp[:, 1] > DEADY
    qs = np.where(sal[:, None], q16(sp[:, :3], lo, k), 0).astype(np.uint32)
    codes = np.where(sal, sp[:, 3].view(np.uint32), 0).astype(np.uint32)
    out[L["oSp"]:L["oSp"] + gS * 6] = _put6(qs)
    out[L["oSk"]:L["oSk"] + gS * 4] = codes
    return out


def _put6(q):
    q = q.reshape(-1, 4, 3).astype(np.uint32)
    sh = np.uint32(16)
    w = np.stack([q[:, 0, 0] | (q[:, 0, 1] << sh), q[:, 0, 2] | (q[:, 1, 0] << sh), q[:, 1, 1] | (q[:, 1, 2] << sh),
                  q[:, 2, 0] | (q[:, 2, 1] << sh), q[:, 2, 2] | (q[:, 3, 0] << sh), q[:, 3, 1] | (q[:, 3, 2] << sh)], axis=1)
    return w.reshape(-1)


def unpack(words, L):
    """unpackLive (sim.js:526-537) -> dict(pos (nW, 3) u16, spd, fm, sp_pos, sp_code)"""
    b = words.view(np.uint8)
    nW, nS = L["nW"], min(int(words[0]), L["nS"])
    pos = b[L["oWp"] * 4:L["oWp"] * 4 + nW * 6].view(np.uint16).reshape(nW, 3)
    spd = b[L["oWs"] * 4:L["oWs"] * 4 + nW]
    fm = b[L["oWf"] * 4:L["oWf"] * 4 + nW]
    sp_pos = b[L["oSp"] * 4:L["oSp"] * 4 + nS * 6].view(np.uint16).reshape(nS, 3)
    sp_code = words[L["oSk"]:L["oSk"] + nS]
    return {"pos": pos, "spd": spd, "fm": fm, "sp_pos": sp_pos, "sp_code": sp_code}
