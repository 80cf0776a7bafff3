"""Benchmark of the FLIP substep on one MI355X.

Workload (N=1): the 256^3 sparse splash -- a unit box at d = 1/512 (h = 1/256),
a pool 0.25 deep, a ball of water r = 0.12 falling at 2 m/s from y = 0.7, a
0.1 half-extent box collider moving along x at 0.5 m/s, 80 red-black SOR
iterations, dt = 1/480, f32.  Particles and grid stay in HBM; the timed region
is K full substeps (every stage, colliders updated each substep).

Prints one JSON line: particle-substeps/s, the roofline of the dominant kernel
(measured with HIP events on the sim's stream), the batched-tanks throughput
next to the same tanks stepped one at a time, the cost of the adjoint
(gradient) pass against the forward pass, and the CPU oracle timed on a
bounded sample of the same scene on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_scene(args):
    from splashtank_amd import scenes
    if args.workload == "splash256":
        return scenes.sparse_splash(d=1.0 / 512, iters=80, dt=1.0 / 480)
    if args.workload == "dam64":
        return scenes.dam_break(d=1.0 / 128, iters=40, dt=1.0 / 240)
    raise SystemExit(f"unknown workload {args.workload}")


def run_gpu(sc, args):
    from splashtank_amd import capi
    opts = capi.make_opts(sc.d, sc.size, ratio=sc.ratio, open_=sc.open_, max_n=sc.n, device=0)
    sim = capi.FlipSimC(opts)
    sim.upload(sc.pos, sc.vel)
    prm = capi.make_params(**sc.params)
    dt = sc.params["dt"]
    t = 0.0

    def one():
        nonlocal t
        t += dt
        if sc.colliders:
            cf, sdf, sig = sc.collider_state(t, dt)
            sim.set_colliders(cf, sdf, sig)
        sim.step(prm)

    log(f"[bench] warmup {args.warmup} substeps, n={sc.n}")
    for _ in range(args.warmup):
        one()
    sim.sync()
    log(f"[bench] timing {args.steps} substeps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one()
    sim.sync()
    wall = time.perf_counter() - t0
    pos = sim.read_particles()
    alive = int(np.sum(pos[:, 1] >= -1e5))
    ctl = sim.ctl()
    stats = dict(blocks=int(ctl[4]), fluid=int(ctl[0] + ctl[1]), vmax=float(ctl[6:7].view(np.float32)[0]))
    # FP64 norms (on the GPU) of the last timed substep's solve and projection
    nrm = sim.solve_norms()
    stats["solve_f64"] = dict(rel_residual=nrm["res_l2"] / max(nrm["rhs_l2"], 1e-300), res_max=nrm["res_max"],
                              div_l2=nrm["div_l2"], div_minus_target_l2=nrm["err_l2"],
                              div_minus_target_max=nrm["err_max"], fluid_cells=int(nrm["fluid"]))
    sim.close()
    # per-stage times of the same substeps: the scene again from its start, the
    # same warm-up, then HIP events around each stage of the timed substeps on the
    # sim's stream (the stream every kernel of the stage is launched on)
    stage_ms = None
    nst = min(args.stage_steps, args.steps) if args.stage_steps >= 0 else args.steps
    if nst:
        sim = capi.FlipSimC(opts)
        sim.upload(sc.pos, sc.vel)
        t = 0.0
        for _ in range(args.warmup):
            one()
        runs = []
        for _ in range(nst):
            t += dt
            if sc.colliders:
                cf, sdf, sig = sc.collider_state(t, dt)
                sim.set_colliders(cf, sdf, sig)
            runs.append(sim.time_stages(prm))
        stage_ms = {k: float(np.mean([r[k] for r in runs])) for k in runs[0]}
        stats["sor_path"] = ["per-iteration", "resident", "streaming"][sim.sor_path()]
        sim.close()
        log(f"[bench] stage times over {nst} substeps")
    return wall, alive, stats, stage_ms


def run_batched(args):
    """Batched tanks: 64 independent tanks (n^3 cells, n log-uniform in 32..128,
    4-8 particles per cell, 40 SOR iterations) advanced (a) one tank after
    another with st_step, each on its own stream, and (b) together with
    st_batch_step, one launch per kernel for all tanks (every tank's SOR
    resident in one cooperative launch).  Same scenes, same warm-up, same
    number of timed substeps; fresh tanks for each way.  (c) From one snapshot
    after the autotuner: the batched SOR as one launch per iteration at the
    tuned width, with stream launches and as HIP graph replays (graphs always
    launch per iteration)."""
    from splashtank_amd import capi, scenes
    tanks = scenes.batched_tanks(count=args.batch_tanks, seed=3)
    ntot = sum(sc.n for sc in tanks)
    prm = capi.make_params(**tanks[0].params)

    def make():
        sims = []
        for sc in tanks:
            s = capi.FlipSimC(capi.make_opts(sc.d, sc.size, ratio=sc.ratio, open_=sc.open_, max_n=sc.n, device=0))
            s.upload(sc.pos, sc.vel)
            sims.append(s)
        return sims

    def sync(sims):
        for s in sims:
            s.sync()

    K, W = args.batch_steps, args.batch_warmup
    log(f"[bench] batched tanks: {len(tanks)} tanks, {ntot} particles; one tank at a time")
    sims = make()
    for _ in range(W):
        for s in sims:
            s.step(prm)
    sync(sims)
    t0 = time.perf_counter()
    for _ in range(K):
        for s in sims:
            s.step(prm)
    sync(sims)
    seq = time.perf_counter() - t0
    for s in sims:
        s.close()
    log(f"[bench] batched tanks: one launch per kernel for all tanks")
    sims = make()
    batch = capi.FlipBatchC(sims)
    for _ in range(W):
        batch.step(prm)
    batch.sync()
    t0 = time.perf_counter()
    for _ in range(K):
        batch.step(prm)
    batch.sync()
    bat = time.perf_counter() - t0
    # algorithmic bytes of one substep summed over the tanks (the model of the
    # single-tank line, each tank with its own blocks, fluid cells and tiles)
    byts = act = 0
    for sc, s in zip(tanks, sims):
        ctl = s.ctl()
        act += int(ctl[4])
        n = int(round(1.0 / (sc.d * sc.ratio)))
        T = ((n + 1 + 7) // 8) ** 3
        byts += 112 * sc.n + 224 * int(ctl[4]) * 512 + (20 + 16 * int(sc.params["iters"])) * int(ctl[0] + ctl[1]) + 16 * T
    batch.close()
    for s in sims:
        s.close()
    # (c) the autotuner (fresh tanks, same warm-up), then from one snapshot K
    # substeps with one SOR launch per iteration at the tuned width, stream
    # launches, then K more as HIP graph replays
    log(f"[bench] batched tanks: autotuner, then launches per iteration (stream, graph replay)")
    sims = make()
    batch = capi.FlipBatchC(sims)
    for _ in range(W):
        batch.step(prm)
    tune = batch.tune_sor(prm, reps=2)
    batch.sync()
    pipe = [c for c in tune["candidates"] if c["mode"] == 1]
    pipe_width = min(pipe, key=lambda c: c["ms"])["width"]
    snaps = [capi.SnapshotC(s) for s in sims]

    def restore():
        for q in snaps:
            q.restore()

    batch.sor_grid(0)
    restore()
    t0 = time.perf_counter()
    batch.step_n(prm, K, graph=False)
    batch.sync()
    tun = time.perf_counter() - t0
    # the graphs' warm-up: the first substep after a restore and the rest are
    # captured on their first use, and the first replays of a new graph run
    # slower (measured 7.3 against 5.7 ms over 20 substeps after 2 warm-up ones),
    # so the graph leg warms up over as many substeps as it times
    restore()
    batch.step_n(prm, K, graph=True)
    batch.sync()
    restore()
    t0 = time.perf_counter()
    batch.step_n(prm, K, graph=True)
    batch.sync()
    gra = time.perf_counter() - t0
    for q in snaps:
        q.close()
    batch.close()
    for s in sims:
        s.close()
    return {"workload": f"{len(tanks)}_tanks_32-128^3", "tanks": len(tanks), "particles": ntot,
            "sor_iters": int(tanks[0].params["iters"]), "steps": K, "warmup": W, "active_tiles": act,
            "value": ntot * K / bat, "unit": "particle-substeps/s", "ms_per_step": bat * 1e3 / K,
            "sequential_value": ntot * K / seq, "sequential_ms_per_step": seq * 1e3 / K,
            "speedup_vs_sequential": seq / bat,
            "sor": "resident, one cooperative launch for all tanks",
            "tuner": {"choice": ["resident", "per-iteration"][tune["mode"]], "width": tune["width"],
                      "candidates_ms": {(["resident", "per-iteration"][c["mode"]] + f' {c["width"]}'): c["ms"]
                                        for c in tune["candidates"]}},
            "per_iteration": {"ms_per_step": tun * 1e3 / K, "value": ntot * K / tun, "sor_width": pipe_width,
                              "speedup_vs_sequential": seq / tun},
            "per_iteration_graph": {"ms_per_step": gra * 1e3 / K, "value": ntot * K / gra, "sor_width": pipe_width,
                                    "speedup_vs_sequential": seq / gra},
            "substep_roofline": {"algorithmic_bytes": byts, "achieved": byts / (bat / K) / 1e9, "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": byts / (bat / K) / 1e9 / HBM_PEAK_GBS}}


def run_adjoint(args):
    """The inverse problem's gradient at 64^3: a dam break (502,656 particles,
    40 SOR iterations) with a sphere collider (r 0.12) on 4 keyframes dropping
    into the water column, 60 substeps with a snapshot every 10; the forward pass
    (st_step, snapshots taken) and the backward pass (each segment re-run with
    its tapes recorded, then reversed, the collider gradient read per substep and
    mapped onto the keys) timed on the GPU, per substep."""
    from splashtank_amd import adjoint, capi, scenes
    sc = scenes.dam_break(d=1.0 / 128, iters=40, dt=1.0 / 240)
    K, every = args.adj_steps, 10
    prm = capi.make_params(**sc.params)
    opts = capi.make_opts(sc.d, sc.size, ratio=sc.ratio, open_=sc.open_, max_n=sc.n, device=0)
    h = 2.0 * sc.d
    dt = float(sc.params["dt"])
    gd, lo, data = scenes.sphere_sdf(0.12, h, 2 * h)
    sphere = adjoint.KeyedCollider(gd, lo, h, data, [0.0, K / 3 * dt, 2 * K / 3 * dt, (K - 1) * dt], smooth=True)
    ro = adjoint.Rollout(opts, prm, K, every=every, cols=[sphere])
    ro.set_keys([np.array([[-0.30, 0.75, 0.0], [-0.30, 0.62, 0.02], [-0.28, 0.52, 0.0], [-0.25, 0.46, -0.02]])])
    rng = np.random.default_rng(4)
    g_pos = rng.standard_normal((sc.n, 3)).astype(np.float32)
    try:
        log(f"[bench] adjoint: 64^3 dam break, {K} substeps, snapshot every {every}")
        ro.forward_raw(sc.pos[:, :3], sc.vel[:, :3])  # warm-up
        ro.sim.sync()
        t0 = time.perf_counter()
        ro.forward_raw(sc.pos[:, :3], sc.vel[:, :3])
        ro.sim.sync()
        plain = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, _, snaps = ro.forward_raw(sc.pos[:, :3], sc.vel[:, :3], keep_snaps=True)
        ro.sim.sync()
        fwd = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, _, gkeys = ro.backward_raw(snaps, g_pos, np.zeros_like(g_pos))
        bwd = time.perf_counter() - t0
        for q in snaps:
            q.close()
    finally:
        ro.close()
    return {"workload": "dam_break_64^3 + keyframed sphere", "particles": sc.n, "substeps": K, "checkpoint_every": every,
            "parameters": "particle start state (pos, vel) + 4 sphere keyframes",
            "key_gradient_norm": float(np.linalg.norm(gkeys[0])),
            "forward_ms_per_substep": plain * 1e3 / K, "forward_with_snapshots_ms_per_substep": fwd * 1e3 / K,
            "backward_ms_per_substep": bwd * 1e3 / K,
            "backward_over_forward": bwd / plain,
            "note": "backward = each segment re-run forward with its tapes recorded, then reversed (st_adj_backward)"}


def counter_summary(workload: str):
    """The committed counter summary of this workload (profiles/<workload>_traffic.json,
    written by --summarize-counters from rocprofv3 --pmc passes over this bench), or None."""
    path = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def measured_traffic(workload: str, kernel: str):
    """HBM-side bytes per dispatch of `kernel` from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, each scaled by what it reads for a
    device copy of known size), or None when there is none for this workload."""
    d = counter_summary(workload)
    try:
        k = d["kernels"][kernel]
        return k["read_bytes"] + k["write_bytes"]
    except (TypeError, KeyError):
        return None


# MI355X: 256 CUs; a CU issues at most one wave64 VALU instruction per clock (four
# SIMDs, four clocks each); 2.4 GHz peak engine clock
CUS, CLOCK_HZ = 256, 2.4e9


def summarize_counters(src: str, workload: str, command: str, calib_mb: int):
    """Per-kernel summary of rocprofv3 --pmc passes (one sub-directory per pass,
    each with run_counter_collection.csv) into profiles/<workload>_traffic.json:
    HBM-side read and write bytes per dispatch (FETCH_SIZE and WRITE_SIZE in KiB,
    scaled by what they read for the calibration copy of --calib-copy-mb), and the
    instruction, wait and LDS counters with the VALU issue share they imply."""
    import csv
    import glob
    per = {}
    durs = {}
    calib_rows = {}
    for f in sorted(glob.glob(os.path.join(src, "*", "run_counter_collection.csv"))):
        rows = list(csv.DictReader(open(f)))
        # the calibration copy: the copy dispatches of the largest grid (the small
        # ones are the parameter uploads)
        cg = max([int(r["Grid_Size"]) for r in rows if "copyBuffer" in r["Kernel_Name"]] or [0])
        for r in rows:
            k = r["Kernel_Name"].split("(")[0]
            if "copyBuffer" in k:
                if int(r["Grid_Size"]) == cg:
                    calib_rows.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
                continue
            per.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
            durs.setdefault(k, {})[(f, r["Dispatch_Id"])] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
    copy = ["calibration copy"] if calib_rows else []
    per.update({"calibration copy": calib_rows} if calib_rows else {})
    mean = lambda v: sum(v) / len(v)
    fs = ws = 1.0
    calib = None
    if copy:
        c = per[copy[0]]
        nbytes = calib_mb << 20
        fs = nbytes / (mean(c["FETCH_SIZE"]) * 1024.0) if "FETCH_SIZE" in c else 1.0
        ws = nbytes / (mean(c["WRITE_SIZE"]) * 1024.0) if "WRITE_SIZE" in c else 1.0
        calib = {"copy_bytes": nbytes, "fetch_kib": mean(c.get("FETCH_SIZE", [0])), "write_kib": mean(c.get("WRITE_SIZE", [0]))}
    out = {"workload": workload, "fetch_scale": fs, "write_scale": ws, "calibration": calib, "source": command, "kernels": {}}
    for k, c in per.items():
        if k in copy:
            continue
        e = {"dispatches": max(len(v) for v in c.values())}
        if "FETCH_SIZE" in c:
            e["read_bytes"] = mean(c["FETCH_SIZE"]) * 1024.0 * fs
        if "WRITE_SIZE" in c:
            e["write_bytes"] = mean(c["WRITE_SIZE"]) * 1024.0 * ws
        for name, v in c.items():
            if name.startswith("SQ_") or name.startswith("GRBM_"):
                e[name] = mean(v)
        d = mean(list(durs[k].values()))
        e["dispatch_s"] = d
        if "SQ_INSTS_VALU" in c:
            e["valu_issue_share"] = e["SQ_INSTS_VALU"] / (CUS * CLOCK_HZ * d)
        if "SQ_INSTS_LDS" in c:
            e["lds_instr_per_cu_clock"] = e["SQ_INSTS_LDS"] / (CUS * CLOCK_HZ * d)
        out["kernels"][k] = e
    path = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


def calib_copy(mb: int):
    """A plain device-to-device copy of `mb` MiB (HIP's blit kernel), so a counter
    pass can scale FETCH_SIZE / WRITE_SIZE by what they read for known traffic."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = C.c_size_t(mb << 20)
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), n) or hip.hipMalloc(C.byref(b), n):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, n)
    for _ in range(3):
        hip.hipMemcpy(b, a, n, 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] calibration copy of {mb} MiB x3 done")


def cpu_baseline(sc, args):
    from splashtank_amd import capi
    from oracle.oracle import Oracle, threads
    opts = capi.make_opts(sc.d, sc.size, ratio=sc.ratio, open_=sc.open_, max_n=sc.n)
    o = Oracle(opts)
    o.upload(sc.pos, sc.vel)
    prm = capi.make_params(**sc.params)
    dt = sc.params["dt"]
    t = 0.0
    for k in range(args.cpu_warmup + args.cpu_steps):
        if k == args.cpu_warmup:
            t0 = time.perf_counter()
        t += dt
        if sc.colliders:
            cf, sdf, sig = sc.collider_state(t, dt)
            o.set_colliders(cf, sdf, sig)
        o.step(prm)
        log(f"[bench] cpu oracle substep {k} done")
    wall = time.perf_counter() - t0
    o.close()
    return sc.n * args.cpu_steps / wall, threads(), wall


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), with the logical CPUs this process may use."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count()
    return f"{name} ({n} logical CPUs available)"


REF_BAKER = os.path.join(ROOT, "oracle", "_ref", "splash")
# the reference's native baker on a dam break of the 64^3 case's shape in its own
# 3 x 2.6 x 3 box (fluid 0.4 x 0.6 x 1 of it): d 0.0224 gives 506,520 particles;
# 60 fps with 4 substeps (dt 1/240), 4 density iterations
REF_SCENE = """splash-scene 1
d 0.0224
fps 60
frames {frames}
substeps 4
iters 4
gravity 9.81
maxn 600000
fill box -1.5 -0.3 0 1.56 -1.5 1.5
"""


def reference_cpu(frames: int):
    """The reference's CPU path (native/splash.cpp, PBF: a different algorithm,
    a baseline only), built from its source into oracle/_ref by the oracle's
    Makefile, timed by its own per-frame clock: sum of N x substeps over elapsed."""
    import re
    import subprocess
    import tempfile
    if not os.path.exists(REF_BAKER):
        return None
    with tempfile.TemporaryDirectory() as td:
        scene = os.path.join(td, "scene.txt")
        with open(scene, "w") as f:
            f.write(REF_SCENE.format(frames=frames))
        log(f"[bench] reference baker: {frames} frames")
        r = subprocess.run([REF_BAKER, scene, os.devnull], capture_output=True, text=True, timeout=300)
    thr = re.search(r"threads: (\d+)", r.stdout)
    rows = re.findall(r"frame (\d+)/\d+\s+(\d+) points\s+([0-9.]+)s elapsed", r.stdout)
    if r.returncode != 0 or not rows:
        return {"error": f"exit {r.returncode}", "stdout_tail": r.stdout[-300:]}
    work = sum(int(n) * 4 for _, n, _ in rows)
    el = float(rows[-1][2])
    return {"value": work / el, "unit": "particle-substeps/s", "cores": int(thr.group(1)) if thr else None,
            "kind": "reference", "algorithm": "PBF (native/splash.cpp), not FLIP: baseline only",
            "sample": f"dam break, {rows[0][1]} particles, {len(rows)} frames x 4 substeps, 4 density iterations; "
                      f"its own elapsed clock ({el:.1f} s)",
            "build": "g++ -O3 -fopenmp native/splash.cpp (oracle/Makefile target ref; -march=native left out)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="splash256", choices=["splash256", "dam64"])
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-warmup", type=int, default=1)
    ap.add_argument("--ref-frames", type=int, default=36, help="frames of the reference's CPU baker to time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stage-steps", type=int, default=-1, help="substeps timed stage by stage (default: all timed ones)")
    ap.add_argument("--calib-copy-mb", type=int, default=0)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-tanks line")
    ap.add_argument("--batch-tanks", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--batch-warmup", type=int, default=5)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint line")
    ap.add_argument("--adj-steps", type=int, default=60)
    ap.add_argument("--summarize-counters", metavar="DIR",
                    help="summarize rocprofv3 --pmc passes under DIR (one sub-directory per pass) into "
                         "profiles/<workload>_traffic.json and exit; --calib-copy-mb names the calibration copy")
    ap.add_argument("--counter-command", default="", help="the rocprofv3 commands, recorded in the summary")
    args = ap.parse_args()
    if args.summarize_counters:
        summarize_counters(args.summarize_counters, args.workload, args.counter_command, args.calib_copy_mb or 256)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the substep runs on a single MI355X; "
              "multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    if args.calib_copy_mb:
        calib_copy(args.calib_copy_mb)
    sc = make_scene(args)
    wall, alive, stats, stage_ms = run_gpu(sc, args)
    solve_f64 = stats.pop("solve_f64", None)
    value = alive * args.steps / wall
    ms_per_step = wall * 1e3 / args.steps
    iters = int(sc.params["iters"])
    fc, blocks = stats["fluid"], stats["blocks"]
    A = blocks * 512
    T = ((256 + 1 + 7) // 8) ** 3 if args.workload == "splash256" else ((64 + 1 + 7) // 8) ** 3
    # algorithmic bytes of one substep: 112 N + 224 A + (20 + 16 I) Fc + 16 T
    sub_bytes = 112 * alive + 224 * A + (20 + 16 * iters) * fc + 16 * T
    roof = None
    kernels = None
    if stage_ms:
        # the three large kernels, each with its algorithmic bytes per launch and
        # its time per launch (HIP events around its stage on the sim's stream):
        #  - pressure: 16 B per fluid cell per red-black iteration (rhs, type,
        #    pressure read and write); the cooperative solve runs all iterations
        #    in one launch (k_sor_resident; the stage also holds three small list
        #    kernels of a few us), the per-iteration path one per iteration;
        #  - g2p: 64 B per particle (state read and written) and 28 B per active
        #    cell (new and old face velocities, count);
        #  - p2g: 32 B per particle (state read) and 56 B per active cell (three
        #    face sums of weight and velocity written, count).
        # The roofline is the HBM one; `limiter` says what the counters in the
        # committed summary show bounds the kernel instead, when it is not HBM.
        coop = stats.get("sor_path") in ("resident", "streaming")
        per_launch_iters = iters if coop else 1
        sor_name = "k_sor_resident" if coop else "k_sor_pipe"
        cands = {
            sor_name: (16 * fc * per_launch_iters, stage_ms["pressure"] * 1e3 / (iters // per_launch_iters),
                       f"{sor_name} (all {iters} red-black SOR iterations in one launch, {stats['sor_path']})" if coop
                       else "k_sor_pipe (one red-black SOR iteration per launch)"),
            "k_g2p_tile": (64 * alive + 28 * A, stage_ms["g2p"] * 1e3, "k_g2p_tile (g2p and the next substep's bin marks)"),
            "k_p2g_tile": (32 * alive + 56 * A, stage_ms["p2g"] * 1e3, "k_p2g_tile (p2g, one workgroup per tile)"),
        }
        summ = counter_summary(args.workload)
        kernels = {}
        for kname, (byts, us, desc) in cands.items():
            ach = byts / (us * 1e-6) / 1e9
            e = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                 "traffic": measured_traffic(args.workload, kname), "kernel": desc,
                 "algorithmic_bytes_per_launch": byts, "launch_us": us}
            k = (summ or {}).get("kernels", {}).get(kname)
            if k:
                e["traffic_source"] = f"profiles/{args.workload}_traffic.json (rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE, calibrated by a device copy; committed, not measured in this run)"
                if "valu_issue_share" in k:
                    share = k["valu_issue_share"]
                    e["valu_issue_share"] = share
                    if share >= 0.5:
                        e["limiter"] = (f"VALU issue: the counters put the kernel's VALU instructions at {share:.0%} of the "
                                        f"CUs' issue rate (one wave64 instruction per CU per clock at 2.4 GHz), "
                                        f"HBM traffic at {(e['traffic'] or 0) / (us * 1e-6) / 1e9 / HBM_PEAK_GBS:.0%} of peak")
            kernels[kname] = e
        dom = max(kernels, key=lambda q: kernels[q]["launch_us"])
        roof = dict(kernels[dom])
    p2g_atomics = None
    if stage_ms:
        # k_p2g_tile: per particle 24 packed 64-bit LDS adds (weight and weighted
        # velocity of 8 faces on 3 axes) and one LDS count add; no global atomics
        # for tiles with storage (own cells stored, the 488-cell halo shell
        # written to the tile's record and added in by k_p2g_merge), only for
        # cells of tiles without storage.  The reference makes 50 global atomics
        # per particle (48 face sums, the count, the speed maximum).
        p2g_s = stage_ms["p2g"] * 1e-3
        p2g_atomics = {"reference_global_per_particle": 50, "lds_per_particle": 25,
                       "global_per_particle": 0, "halo_record_words_per_launch": blocks * 2928,
                       "lds_atomics_per_s": 25 * alive / p2g_s,
                       "reference_atomics_per_s_equivalent": 50 * alive / p2g_s, "p2g_ms": stage_ms["p2g"],
                       # no published figure: the LDS moves 128 B per clock per CU, so at most 16 lanes'
                       # 64-bit adds per clock per CU, x 256 CUs x 2.4 GHz (bank conflicts only lower it)
                       "lds_atomic_peak_per_s": 16 * CUS * CLOCK_HZ,
                       "lds_atomic_peak_basis": "128 B/clock/CU of LDS = 16 64-bit adds/clock/CU x 256 CUs x 2.4 GHz (model)",
                       "lds_atomic_frac": 25 * alive / p2g_s / (16 * CUS * CLOCK_HZ)}
    out = {
        "metric": "particle-substeps/s", "value": value, "unit": "particle-substeps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": sc.name + ("_256^3" if args.workload == "splash256" else "_64^3"), "particles": sc.n,
                   "sor_iters": sc.params["iters"], "dt": sc.params["dt"], **stats},
        "stage_ms": stage_ms,
        "solve_f64": solve_f64,
        "roofline": roof,
        "kernel_rooflines": kernels,
        "p2g_atomics": p2g_atomics,
        "substep_roofline": {"algorithmic_bytes": sub_bytes, "achieved": sub_bytes / (ms_per_step * 1e-3) / 1e9,
                             "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": sub_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    if not args.no_batch:
        out["batched"] = run_batched(args)
    if not args.no_adjoint:
        out["adjoint"] = run_adjoint(args)
    if not args.no_cpu:
        out["reference_cpu"] = reference_cpu(args.ref_frames)
        v, cores, cw = cpu_baseline(sc, args)
        out["cpu_baseline"] = {"value": v, "unit": "particle-substeps/s", "cores": cores, "kind": "port",
                               "cpu": cpu_model(),
                               "sample": f"{args.cpu_steps} substeps of the same scene after {args.cpu_warmup} warm-up, "
                                         f"{cw:.1f} s"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
