"""The hand-written adjoint of the substep (st_adj_*, splashtank_amd/adjoint.py)
against the FP64 reference gradient (torch autograd of tests/adjoint_ref.py,
itself pinned against central differences and the oracle on the CPU), and
against central differences of that FP64 forward with the discrete choices
frozen.  Bar: relative error <= 1e-3, in norm and along 16 random directions."""
import numpy as np
import pytest
import torch

from splashtank_amd import adjoint, capi, scenes
from splashtank_amd.capi import StError
from oracle.oracle import Oracle
from tests.adjoint_ref import RefSim, keyed_sphere, ref_cols
from tests.helpers import opts_for

pytestmark = pytest.mark.gpu
RTOL = 1e-3
# Per direction: the GPU gradient is that of the f32 forward, within RTOL of the
# FP64 gradient in norm; a single directional derivative carries that error
# divided by its share of the norm, so directions down to a fifth of the RMS are
# held to twice the norm bar.
RTOL_DIR = 2e-3


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _dirs_check(ga, gb, rng, k=16):
    """Directional derivatives along k random directions: the worst difference,
    relative to the RMS of the reference's directional derivatives (a random
    direction can be nearly orthogonal to the gradient, which would make a
    per-direction ratio meaningless)."""
    xs, ys = [], []
    for _ in range(k):
        d = rng.standard_normal(ga.shape)
        xs.append(float((ga * d).sum()))
        ys.append(float((gb * d).sum()))
    xs, ys = np.array(xs), np.array(ys)
    return float(np.abs(xs - ys).max() / max(np.sqrt(np.mean(ys * ys)), 1e-30))


def _gpu_clamps(sc, ref, dims, pos0, vel0, prm, K, keyed=None):
    """The reference follows the box-clamp branches of the f32 run: a particle
    resting on a wall has a normal velocity within rounding of zero, and which
    side of the bound it lands on is decided by f32 rounding (where the
    derivative itself jumps).  keyed: (KeyedCollider, keys) moved as the
    rollout moves it."""
    g = capi.FlipSimC(opts_for(sc))
    try:
        p4 = np.zeros((len(pos0), 4), np.float32)
        v4 = np.zeros_like(p4)
        p4[:, :3] = pos0
        v4[:, :3] = vel0
        g.upload(p4, v4)
        out = []
        dt = float(prm.dt)
        for s in range(K):
            if keyed is not None:
                kc, keys = keyed
                c, v, _, _ = kc.at(keys, s * dt, dt)
                g.set_colliders(scenes.collider_record(0, kc.dims, kc.origin, kc.cell, c, vel=v), kc.data, 1)
            g.step(prm)
            out.append(ref.wall_branches(dims, g.dump("pos")))
        return out
    finally:
        g.close()


def test_adjoint_matches_fp64_reference_linear_loss():
    sc = scenes.dam_break(d=1.0 / 32, iters=20)  # 16^3 cells
    rng = np.random.default_rng(1)
    n = sc.n
    vel0 = (sc.vel[:, :3] + 0.05 * rng.standard_normal((n, 3))).astype(np.float32)
    pos0 = sc.pos[:, :3].astype(np.float32)
    A = rng.standard_normal((n, 3))
    B = rng.standard_normal((n, 3))
    K = 5
    prm = capi.make_params(**sc.params)
    ro = adjoint.Rollout(opts_for(sc), prm, K, every=2)
    try:
        print("[adjoint] forward + backward, 16^3, 5 substeps, snapshots every 2", flush=True)
        _, _, snaps = ro.forward_raw(pos0, vel0, keep_snaps=True)
        gp, gv, _ = ro.backward_raw(snaps, A.astype(np.float32), B.astype(np.float32))
    finally:
        ro.close()
    dims = Oracle(opts_for(sc), fp64=True).dims
    ref = RefSim(dims, sc.params)
    clamps = _gpu_clamps(sc, ref, dims, pos0, vel0, prm, K)
    p0 = torch.tensor(pos0, dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor(vel0, dtype=torch.float64, requires_grad=True)
    pK, vK, _ = ref.rollout(p0, v0, K, clamps=clamps)
    ((torch.tensor(A) * pK).sum() + (torch.tensor(B) * vK).sum()).backward()
    rp, rv = p0.grad.numpy(), v0.grad.numpy()
    ev, ep = _rel(gv, rv), _rel(gp, rp)
    dv, dp = _dirs_check(gv, rv, rng), _dirs_check(gp, rp, rng)
    print(f"[adjoint] rel err vel {ev:.2e} pos {ep:.2e}; 16 directions vel {dv:.2e} pos {dp:.2e}", flush=True)
    assert ev <= RTOL and ep <= RTOL and dv <= RTOL and dp <= RTOL


def test_adjoint_matches_fp64_reference_64_cubed():
    """The same bar at the size the adjoint is benchmarked on: a 64^3 dam break
    (502,656 particles), 4 substeps with snapshots every 2 (the FP64 reference
    keeps its autograd graph on the host, about 12 GB).  Longer rollouts add
    particles whose f32 and f64 positions land on different sides of a cell
    boundary, so the two runs take different discrete branches for them and
    their gradients part: over 8 substeps the position gradient's relative
    error is 3.6e-3, 88% of it from 100 of the 502,656 particles (surface
    particles; 1.3e-3 without them), while the velocity gradient stays at 2.9e-4."""
    sc = scenes.dam_break(d=1.0 / 128, iters=20)
    rng = np.random.default_rng(11)
    n = sc.n
    vel0 = (sc.vel[:, :3] + 0.05 * rng.standard_normal((n, 3))).astype(np.float32)
    pos0 = sc.pos[:, :3].astype(np.float32)
    A = rng.standard_normal((n, 3))
    B = rng.standard_normal((n, 3))
    K = 4
    prm = capi.make_params(**sc.params)
    ro = adjoint.Rollout(opts_for(sc), prm, K, every=2)
    try:
        print(f"[adjoint] forward + backward, 64^3, {n} particles, {K} substeps, snapshots every 2", flush=True)
        _, _, snaps = ro.forward_raw(pos0, vel0, keep_snaps=True)
        gp, gv, _ = ro.backward_raw(snaps, A.astype(np.float32), B.astype(np.float32))
    finally:
        ro.close()
    dims = Oracle(opts_for(sc), fp64=True).dims
    ref = RefSim(dims, sc.params)
    clamps = _gpu_clamps(sc, ref, dims, pos0, vel0, prm, K)
    p0 = torch.tensor(pos0, dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor(vel0, dtype=torch.float64, requires_grad=True)
    print("[adjoint] FP64 reference rollout and its gradient", flush=True)
    pK, vK, _ = ref.rollout(p0, v0, K, clamps=clamps)
    ((torch.tensor(A) * pK).sum() + (torch.tensor(B) * vK).sum()).backward()
    rp, rv = p0.grad.numpy(), v0.grad.numpy()
    ev, ep = _rel(gv, rv), _rel(gp, rp)
    dv, dp = _dirs_check(gv, rv, rng), _dirs_check(gp, rp, rng)
    print(f"[adjoint] 64^3: rel err vel {ev:.2e} pos {ep:.2e}; 16 directions vel {dv:.2e} pos {dp:.2e}", flush=True)
    assert ev <= RTOL and ep <= RTOL and dv <= RTOL and dp <= RTOL


def test_adjoint_checkpoint_spacing_does_not_change_the_gradient():
    sc = scenes.dam_break(d=1.0 / 32, iters=11)
    rng = np.random.default_rng(2)
    n = sc.n
    pos0 = sc.pos[:, :3].astype(np.float32)
    vel0 = (0.05 * rng.standard_normal((n, 3))).astype(np.float32)
    A = rng.standard_normal((n, 3)).astype(np.float32)
    prm = capi.make_params(**sc.params)
    out = []
    for every in (1, 3, 4):
        ro = adjoint.Rollout(opts_for(sc), prm, 4, every=every)
        try:
            _, _, snaps = ro.forward_raw(pos0, vel0, keep_snaps=True)
            out.append(ro.backward_raw(snaps, A, np.zeros_like(A))[:2])
        finally:
            ro.close()
    for gp, gv in out[1:]:
        # the g2p face gradients are f32 atomics: order-dependent roundings only
        assert _rel(gv, out[0][1]) < 1e-5 and _rel(gp, out[0][0]) < 1e-5


def _assert_directions(ans, fds, what):
    """Directional derivatives of the GPU gradient against central differences:
    every direction within RTOL of the RMS of the differences, and within RTOL_DIR
    of its own difference wherever that is not near zero (below a fifth of the RMS
    a relative error measures the cancellation, not the gradient)."""
    rms = float(np.sqrt(np.mean(fds * fds)))
    err = np.abs(ans - fds)
    own = err / np.abs(fds)
    big = np.abs(fds) >= 0.2 * rms
    print(f"[adjoint] {what} against central differences: worst {err.max() / rms:.2e} of the RMS, worst per "
          f"direction {own[big].max():.2e} over {int(big.sum())} directions (all: {own.round(6).tolist()})", flush=True)
    assert err.max() <= RTOL * rms
    assert big.sum() >= len(fds) // 2
    assert own[big].max() <= RTOL_DIR


def test_adjoint_density_loss_block_velocities():
    """The inverse problem's gradient: per-tile initial velocities, a dam break,
    the L2 distance of the particle density after K substeps to Gaussian blobs;
    GPU adjoint through torch autograd (adjoint.Rollout) against the FP64
    reference gradient and central differences of the FP64 forward."""
    sc = scenes.dam_break(d=1.0 / 64, iters=20)  # 32^3 cells, 4^3 tiles
    o = Oracle(opts_for(sc), fp64=True)
    dims = o.dims
    nx, ny, nz = dims.nx, dims.ny, dims.nz
    origin = (dims.x0, dims.y0, dims.z0)
    h = float(dims.h)
    rng = np.random.default_rng(3)
    cent = [(rng.uniform(-0.3, 0.3), rng.uniform(0.1, 0.4), rng.uniform(-0.3, 0.3)) for _ in range(3)]
    target = torch.tensor(adjoint.gaussian_target((nx, ny, nz), origin, h, cent, 4 * h, float(dims.rest)))
    cells = np.floor((sc.pos[:, :3] - np.array(origin)) / h).astype(np.int64)
    tx, ty = dims.tx, dims.ty
    blk = torch.tensor((cells[:, 0] >> 3) + tx * ((cells[:, 1] >> 3) + ty * (cells[:, 2] >> 3)))
    nb = dims.tx * dims.ty * dims.tz
    u0 = 0.2 * rng.standard_normal((nb, 3))
    K = 6
    prm = capi.make_params(**sc.params)
    pos0 = torch.tensor(sc.pos[:, :3], dtype=torch.float64)

    def loss(pK):
        return ((adjoint.splat_density(pK, origin, h, (nx, ny, nz)) - target) ** 2).sum()

    ro = adjoint.Rollout(opts_for(sc), prm, K, every=3)
    try:
        print("[adjoint] density loss, 32^3, 6 substeps", flush=True)
        u = torch.tensor(u0, requires_grad=True)
        pK, _ = ro(pos0.to(torch.float32), u[blk].to(torch.float32))
        L = loss(pK.to(torch.float64))
        L.backward()
        g_gpu = u.grad.numpy().copy()
    finally:
        ro.close()
    ref = RefSim(dims, sc.params)
    clamps = _gpu_clamps(sc, ref, dims, sc.pos[:, :3].astype(np.float32),
                         torch.tensor(u0)[blk].to(torch.float32).numpy(), prm, K)
    ur = torch.tensor(u0, requires_grad=True)
    pr, _, rec = ref.rollout(pos0, ur[blk], K, clamps=clamps)
    Lr = loss(pr)
    Lr.backward()
    g_ref = ur.grad.numpy()
    e = _rel(g_gpu, g_ref)
    print(f"[adjoint] loss gpu {float(L):.6e} ref {float(Lr):.6e}; grad rel err {e:.2e}", flush=True)
    assert abs(float(L) - float(Lr)) <= 1e-4 * abs(float(Lr))
    assert e <= RTOL
    # central differences of the FP64 forward (discrete choices frozen) along 16 directions
    eps = 1e-6
    fds, ans = [], []
    for _ in range(16):
        d = torch.tensor(rng.standard_normal((nb, 3)))
        with torch.no_grad():
            a, _, _ = ref.rollout(pos0, (torch.tensor(u0) + eps * d)[blk], K, frozen=rec)
            b, _, _ = ref.rollout(pos0, (torch.tensor(u0) - eps * d)[blk], K, frozen=rec)
            fds.append((float(loss(a)) - float(loss(b))) / (2 * eps))
        ans.append(float((torch.tensor(g_gpu) * d).sum()))
    fds, ans = np.array(fds), np.array(ans)
    worst = float(np.abs(ans - fds).max() / np.sqrt(np.mean(fds * fds)))
    _assert_directions(ans, fds, "16 directions")


def test_adjoint_sphere_keyframes_and_block_velocities():
    """The inverse problem with both parameter sets: per-tile initial
    velocities and the 4 key positions of a sphere collider that moves into the
    water (solid faces at its velocity, particles pushed out of it), density
    loss after K substeps; GPU adjoint (st_adj_col_grad through the keyframe
    map) against the FP64 reference gradient and central differences of the
    FP64 forward along 16 joint directions."""
    sc = scenes.dam_break(d=1.0 / 64, iters=20)  # 32^3 cells, 4^3 tiles
    o = Oracle(opts_for(sc), fp64=True)
    dims = o.dims
    nx, ny, nz = dims.nx, dims.ny, dims.nz
    origin = (dims.x0, dims.y0, dims.z0)
    h = float(dims.h)
    dt = float(sc.params["dt"])
    rng = np.random.default_rng(4)
    cent = [(rng.uniform(-0.3, 0.3), rng.uniform(0.1, 0.4), rng.uniform(-0.3, 0.3)) for _ in range(3)]
    target = torch.tensor(adjoint.gaussian_target((nx, ny, nz), origin, h, cent, 4 * h, float(dims.rest)))
    cells = np.floor((sc.pos[:, :3] - np.array(origin)) / h).astype(np.int64)
    tx, ty = dims.tx, dims.ty
    blk = torch.tensor((cells[:, 0] >> 3) + tx * ((cells[:, 1] >> 3) + ty * (cells[:, 2] >> 3)))
    nb = dims.tx * dims.ty * dims.tz
    u0 = 0.2 * rng.standard_normal((nb, 3))
    kc, keys = keyed_sphere(dims, dt)
    K = 6
    prm = capi.make_params(**sc.params)
    pos0 = torch.tensor(sc.pos[:, :3], dtype=torch.float64)

    def loss(pK):
        return ((adjoint.splat_density(pK, origin, h, (nx, ny, nz)) - target) ** 2).sum()

    ro = adjoint.Rollout(opts_for(sc), prm, K, every=3, cols=[kc])
    try:
        print("[adjoint] sphere keyframes + tile velocities, 32^3, 6 substeps", flush=True)
        u = torch.tensor(u0, requires_grad=True)
        kt = torch.tensor(keys, requires_grad=True)
        pK, _ = ro(pos0.to(torch.float32), u[blk].to(torch.float32), kt)
        L = loss(pK.to(torch.float64))
        L.backward()
        gu, gk = u.grad.numpy().copy(), kt.grad.numpy().copy()
    finally:
        ro.close()
    ref = RefSim(dims, sc.params)
    vel0 = torch.tensor(u0)[blk].to(torch.float32).numpy()
    clamps = _gpu_clamps(sc, ref, dims, sc.pos[:, :3].astype(np.float32), vel0, prm, K, keyed=(kc, keys))
    ur = torch.tensor(u0, requires_grad=True)
    kr = torch.tensor(keys, requires_grad=True)
    pr, _, rec = ref.rollout(pos0, ur[blk], K, clamps=clamps, cols=ref_cols([kc], [kr], K, dt))
    hits = sum(int(br["hits"][0].sum()) for _, br in rec)
    Lr = loss(pr)
    Lr.backward()
    ru, rk = ur.grad.numpy(), kr.grad.numpy()
    eu, ek = _rel(gu, ru), _rel(gk, rk)
    print(f"[adjoint] {hits} pushes; loss gpu {float(L):.6e} ref {float(Lr):.6e}; "
          f"grad rel err tiles {eu:.2e} keys {ek:.2e}", flush=True)
    print(f"[adjoint] key gradient gpu {gk.ravel().round(4).tolist()}", flush=True)
    print(f"[adjoint] key gradient ref {rk.ravel().round(4).tolist()}", flush=True)
    assert hits > 100
    assert abs(float(L) - float(Lr)) <= 1e-4 * abs(float(Lr))
    assert eu <= RTOL and ek <= RTOL
    eps = 1e-6
    fds, ans = [], []
    for _ in range(16):
        du = rng.standard_normal((nb, 3))
        dk = rng.standard_normal(keys.shape) * 0.05
        with torch.no_grad():
            a, _, _ = ref.rollout(pos0, torch.tensor(u0 + eps * du)[blk], K, frozen=rec,
                                  cols=ref_cols([kc], [torch.tensor(keys + eps * dk)], K, dt))
            b, _, _ = ref.rollout(pos0, torch.tensor(u0 - eps * du)[blk], K, frozen=rec,
                                  cols=ref_cols([kc], [torch.tensor(keys - eps * dk)], K, dt))
            fds.append((float(loss(a)) - float(loss(b))) / (2 * eps))
        ans.append(float((gu * du).sum() + (gk * dk).sum()))
    _assert_directions(np.array(ans), np.array(fds), "16 joint directions")
    # the keys alone, at the same scale as the tile velocities (unit directions in
    # key space), so the key gradient meets central differences on its own.  The
    # collider's distance is trilinear in its grid, so the loss is smooth only
    # inside one grid cell: the differences take steps of 1e-7 (at 1e-6 the
    # pushed particles cross cell boundaries of the grid and the FP64 reference
    # gradient itself misses its differences by 2.5e-3 of the RMS; at 1e-7 by 2e-5)
    eps_k = 1e-7
    fds, ans = [], []
    for _ in range(16):
        dk = rng.standard_normal(keys.shape)
        with torch.no_grad():
            a, _, _ = ref.rollout(pos0, torch.tensor(u0)[blk], K, frozen=rec,
                                  cols=ref_cols([kc], [torch.tensor(keys + eps_k * dk)], K, dt))
            b, _, _ = ref.rollout(pos0, torch.tensor(u0)[blk], K, frozen=rec,
                                  cols=ref_cols([kc], [torch.tensor(keys - eps_k * dk)], K, dt))
            fds.append((float(loss(a)) - float(loss(b))) / (2 * eps_k))
        ans.append(float((gk * dk).sum()))
    _assert_directions(np.array(ans), np.array(fds), "16 key-only directions")


def test_adjoint_refuses_what_it_does_not_cover():
    sc = scenes.sparse_splash(d=1.0 / 64, iters=5)  # has a box collider
    g = capi.FlipSimC(opts_for(sc))
    try:
        g.upload(sc.pos, sc.vel)
        cf, sdf, sig = sc.collider_state(0.0, sc.params["dt"])
        g.set_colliders(np.tile(cf, 5), sdf, sig)  # five colliders: one more than the adjoint follows
        a = capi.AdjointC(g, 2)
        with pytest.raises(StError):
            a.record(capi.make_params(**sc.params))
        g.set_colliders(cf, sdf, sig)
        a.record(capi.make_params(**sc.params))
        g.set_colliders(cf, sdf, sig + 1)  # new grids: the recorded substep cannot be reversed
        with pytest.raises(StError):
            a.backward()
        a.close()
    finally:
        g.close()
    sc = scenes.dam_break(d=1.0 / 32)
    g = capi.FlipSimC(opts_for(sc))
    try:
        g.upload(sc.pos, sc.vel)
        a = capi.AdjointC(g, 1)
        with pytest.raises(StError):
            a.record(capi.make_params(**dict(sc.params, drag=0.5)))
        with pytest.raises(StError):
            a.backward()  # nothing recorded
        a.record(capi.make_params(**sc.params))
        with pytest.raises(StError):
            a.record(capi.make_params(**sc.params))  # one tape only
        a.close()
    finally:
        g.close()


def test_adjoint_one_launch_solve_matches_launches_per_sweep():
    """The transposed solve as one cooperative launch (a grid barrier between
    its sweeps) against one launch per sweep: the same sweeps in the same
    order, so the gradients after 3 reversed substeps of the 32^3 dam break
    agree up to the order of the f32 atomics of the reversed g2p and p2g
    (relative 1e-5 in norm; a sweep reading a stale neighbour would be far off);
    the default path is one launch per sweep."""
    sc = scenes.dam_break(d=1.0 / 64, iters=20)
    prm = capi.make_params(**sc.params)
    out = []
    for one in (True, False):
        g = capi.FlipSimC(opts_for(sc))
        a = None
        try:
            g.upload(sc.pos, sc.vel)
            a = capi.AdjointC(g, 3)
            if one:
                a.set_sor_path(True)
            for _ in range(3):
                a.record(prm)
            rng = np.random.default_rng(5)
            n = g.count()
            a.seed(rng.standard_normal((n, 3)), rng.standard_normal((n, 3)))
            paths = []
            for _ in range(3):
                a.backward()
                paths.append(a.sor_path())
            gp, gv = a.grad()
            out.append((gp.copy(), gv.copy(), paths))
        finally:
            if a is not None:
                a.close()
            g.close()
    print(f"[adjoint] one-launch solve paths {out[0][2]}, per-sweep paths {out[1][2]}", flush=True)
    assert all(out[0][2]) and not any(out[1][2])
    ep, ev = _rel(out[0][0], out[1][0]), _rel(out[0][1], out[1][1])
    print(f"[adjoint] one launch vs per sweep: rel diff pos {ep:.2e} vel {ev:.2e}", flush=True)
    assert ep <= 1e-5 and ev <= 1e-5
