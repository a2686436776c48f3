"""Benchmark of the gfx950 rigid-body step.

Default workload: 65,536 convex hulls (a third unit cubes, a third random
8-vertex and a third random 32-vertex hulls) falling onto a static ground,
dt 1/60, 10 solver iterations, FP32, 500 steps -- body-steps/s of the whole
step (broadphase, narrowphase, solver, integration) with all state resident
in HBM, the step replayed from one captured hipGraph.  The scene is first
advanced untimed into its contact-rich regime (step 200 of its 500, when
most hulls have landed); the timed window follows the warmup steps after
that, and the average over the whole 500-step run is reported beside it.

Prints one JSON line.  Options:
  --gpus N        only 1 is supported (one GPU)
  --steps K       timed steps (default 200)
  --warmup W      untimed steps before timing, after the settling steps (default 20)
  --settle S      untimed steps before the warmup (default: the workload's)
  --workload      hulls64k (default) | lattice4k | batched4k
  --cpu-seconds S budget of the CPU-oracle baseline sample (default 15, 0 = skip)
  --per-scene-baseline  (batched4k) also time the same scenes stepped one
                  world and one launch set (captured graph) per scene
  --no-extras     only the workload's own figures (profiling runs)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0

# steps each workload is advanced untimed before the warmup: the 65,536
# hulls fall from up to 122 m (most have landed by step 200 of 500), the
# lattice lands within 40 steps, the piles settle within 120
SETTLE = {"hulls64k": 200, "lattice4k": 40, "batched4k": 120}

# algorithmic bytes per unit for each stage: what each stage must read and
# write once per body, pair, contact or sort pass (stage_bytes below)
STAGES = ["setup", "sort", "pairs", "narrowphase", "constraints", "colouring", "solve",
          "integrate"]

# k_tile_solve, per row and sweep: normal sweeps on precomputed rows read
# 368 B and write the 16 B of impulses; friction sweeps read the full 312 B
# row and write 16 B; scene tiles run every sweep in one launch and write
# both impulse vectors once (328 B).  Per tile body and launch: velocities
# read and written (4 float4 = 64 B).
ROW_NORMAL, ROW_FRICTION, ROW_SCENE, TILE_BODY = 384, 328, 328, 64


def solve_bytes(contacts, tile_bodies, sweeps, scene_tiles):
    """Algorithmic bytes of the solver kernel per step."""
    if scene_tiles:
        return ROW_SCENE * contacts + TILE_BODY * tile_bodies
    half = sweeps // 2
    return half * (ROW_NORMAL * contacts + TILE_BODY * tile_bodies) + \
        half * (ROW_FRICTION * contacts + TILE_BODY * tile_bodies)


def stage_bytes(stage, bodies, pairs, contacts, passes, sweeps, tile_bodies, scene_tiles):
    if stage == "setup":      # read pos,quat,linvel,invI_local,mass,shape,scene,large; write aabb,kv,linvel,invI
        return 184 * bodies
    if stage == "sort":       # histogram read + P passes of read+write (key,value)
        return (2 * passes + 1) * 8 * bodies
    if stage == "pairs":      # sorted kv + AABB per body, one int2 per pair, count scan
        return 40 * bodies + 8 * bodies + 8 * pairs
    if stage == "narrowphase":  # pair, two body transforms/shape ids, contact slot + flag
        return (8 + 2 * 40 + 96 + 4) * pairs
    if stage == "constraints":  # flag scan, slot read, contact + constraint write
        return 8 * pairs + (96 + 96 + 176) * contacts
    if stage == "colouring":  # tile key sort + one pass over rows
        return (3 * 2 + 1) * 8 * contacts + 24 * contacts
    if stage == "solve":
        return solve_bytes(contacts, tile_bodies, sweeps, scene_tiles)
    if stage == "integrate":  # read+write pos, quat, linvel, angvel; read inverse mass
        return 108 * bodies
    return 0


def traffic_profile(workload):
    """HBM bytes per k_tile_solve dispatch measured by rocprofv3 --pmc
    (FETCH_SIZE, WRITE_SIZE, scaled by a copy kernel's readings) at the
    state the default command times, from profiles/hbm_traffic_<workload>.json
    (written with the counter runs; the .md beside it gives the commands)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d


def build_scene(name):
    from experiments_amd import scenes
    if name == "hulls64k":
        return scenes.random_hulls(n=65536, box=120.0, pool=(1024, 1024), steps=500, iterations=10)
    if name == "lattice4k":
        return scenes.lattice_cubes(16, 16, 16)
    if name == "batched4k":
        return scenes.batched_piles(n_scenes=4096, lo=8, hi=512, iterations=10)
    raise SystemExit("unknown workload %s" % name)


def cpu_baseline(world, scene, budget_s):
    """The oracle port (single-threaded C restatement) stepping the same scene
    from the world's current state for about budget_s seconds."""
    import oracle.step as OS
    snap = OS.Snapshot(world)
    n = snap.state["pos"].shape[0]
    steps = 0
    t0 = time.perf_counter()
    while True:
        out = OS.oracle_step(snap, None)
        for k in ("pos", "quat", "linvel", "angvel"):
            snap.state[k] = out[k]
        steps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    import oracle as O
    threads = int(O.lib().orc_num_threads())
    return {"value": n * steps / dt, "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": "%d step(s) of the same %d-body scene from the GPU run's state after "
                      "warmup, oracle/rb_oracle.c (C; narrowphase on %d OpenMP threads, the "
                      "other stages serial), %.1f s" % (steps, n, threads, dt)}


def per_scene_baseline(scene, steps, warmup, max_scenes=None):
    """The batched scenes stepped one world (own buffers, own captured graph)
    per scene, one after another: the per-scene launch-set comparison."""
    from experiments_amd import scenes as S
    subs = S.split_scenes(scene)
    if max_scenes:
        subs = subs[:max_scenes]
    n_sc = len(subs)
    worlds, bodies = [], 0
    for sub in subs:
        bodies += sub.num_bodies
        w = S.build_world(sub, max_pairs=max(4096, 16 * sub.num_bodies),
                          max_contacts=max(2048, 8 * sub.num_bodies), max_hull_vertices=1 << 16)
        w.use_graph(True)
        w.step(max(warmup, 1))
        worlds.append(w)
    total_ms = 0.0
    t0 = time.perf_counter()
    for w in worlds:
        w.step(steps)
        total_ms += w.last_step_ms()
    host_s = time.perf_counter() - t0
    for w in worlds:
        w.close()
    return {"value": bodies * steps / host_s, "scenes": n_sc, "bodies": bodies,
            "host_s": round(host_s, 4), "gpu_ms_sum": round(total_ms, 3)}


def sap_on_world(w, steps=20):
    """The sweep-and-prune pair search (mi_rb_sap_*) on the world's current
    AABBs (after the stage-timing step), timed with its HIP events; checked
    against the world's grid pair list of the same AABBs."""
    import ctypes
    import torch
    from experiments_amd import _lib as L
    mn, mx = w.aabbs()
    n = len(mn)
    dmn = torch.from_numpy(np.ascontiguousarray(mn)).cuda()
    dmx = torch.from_numpy(np.ascontiguousarray(mx)).cuda()
    # the grid's overflow list (the ground) as the large AABBs
    dl = torch.from_numpy(w.read(L.BUF_IS_LARGE, np.int32)).cuda()
    max_pairs = 1 << 23
    desc = L.SapDesc(0, max_pairs)
    h = ctypes.c_void_p()
    L.call("mi_rb_sap_create", ctypes.byref(desc), n, ctypes.byref(h))
    pairs = torch.zeros((max_pairs, 2), dtype=torch.int32, device="cuda")
    num = torch.zeros(1, dtype=torch.int32, device="cuda")
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    ms = np.zeros(4, np.float32)
    acc = np.zeros(4)
    for k in range(steps + 2):
        L.call("mi_rb_sap_find_pairs", h, P(dmn), P(dmx), P(dl), n, P(pairs), P(num), None)
        L.call("mi_rb_sap_last_ms", h, ms.ctypes.data_as(ctypes.c_void_p))
        if k >= 2:
            acc += ms
    acc /= steps
    k = int(num.cpu()[0])
    same = bool(np.array_equal(pairs[:k].cpu().numpy(), w.pairs()))
    L.lib().mi_rb_sap_destroy(h)
    return {"aabbs": n, "axis": "x", "large": int(dl.sum()), "pairs": k,
            "equal_to_grid_pairs": same,
            "phase_ms": {"keys": round(acc[0], 4), "sort": round(acc[1], 4),
                         "sweep_count_scan_write": round(acc[2], 4), "pair_order": round(acc[3], 4)},
            "ms": round(float(acc.sum()), 4)}


def broadphase_1m(steps=100, warmup=5):
    """Broadphase stress (1,048,576 spheres and boxes, radii log-uniform in
    [0.25, 2], 400^3 domain, level-0 cell = 2 x the median radius, 4 grid
    levels): the hierarchical grid API on AABBs that move with their velocity
    every step.  Per-phase times from the grid's HIP events; hash + sort +
    scan bytes against the HBM peak."""
    import ctypes
    import torch
    from experiments_amd import _lib as L
    from experiments_amd import scenes
    mn, mx, v, r = scenes.broadphase_aabbs()
    n = len(mn)
    cell = float(2 * np.median(r))
    max_pairs = 1 << 23
    desc = L.GridDesc(cell, (ctypes.c_int * 3)(256, 256, 256), max_pairs, 1, 4)
    g = ctypes.c_void_p()
    L.call("mi_rb_grid_create", ctypes.byref(desc), n, ctypes.byref(g))
    dmn = torch.from_numpy(mn).cuda()
    dmx = torch.from_numpy(mx).cuda()
    dv = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    dv[:, :3] = torch.from_numpy(v).cuda() * (1.0 / 60.0)
    pairs = torch.zeros((max_pairs, 2), dtype=torch.int32, device="cuda")
    num = torch.zeros(1, dtype=torch.int32, device="cuda")
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    acc = np.zeros(4)
    ms = np.zeros(4, np.float32)
    npairs = 0
    for k in range(warmup + steps):
        dmn += dv
        dmx += dv
        L.call("mi_rb_grid_find_pairs", g, P(dmn), P(dmx), None, n, P(pairs), P(num), stream)
        L.call("mi_rb_grid_last_ms", g, ms.ctypes.data_as(ctypes.c_void_p))
        if k >= warmup:
            acc += ms
            npairs += int(num.cpu()[0])
    flags = ctypes.c_int()
    L.call("mi_rb_grid_overflow", g, ctypes.byref(flags))
    L.lib().mi_rb_grid_destroy(g)
    acc /= steps
    key_bits = 24 + 2
    passes = (key_bits + 7) // 8
    hss_bytes = (40 + (2 * passes + 1) * 8 + 8) * n      # hash + sort + count scan
    hss_ms = acc[0] + acc[1]                              # hash, sort (the scan is inside "pairs")
    total_ms = float(acc.sum())
    return {"bodies": n, "steps": steps, "pairs_per_step": npairs // steps,
            "phase_ms": {"hash": round(acc[0], 4), "sort": round(acc[1], 4),
                         "cells_count_scan_write": round(acc[2], 4), "pair_order": round(acc[3], 4)},
            "broadphase_ms_per_step": round(total_ms, 4),
            "body_steps_per_s": round(n / (total_ms / 1e3), 1),
            "key_bits": key_bits, "sort_passes": passes,
            "hash_sort_hbm": {"bytes_per_body": (40 + (2 * passes + 1) * 8), "ms": round(hss_ms, 4),
                              "achieved_GBs": round((40 + (2 * passes + 1) * 8) * n / (hss_ms / 1e3) / 1e9, 1),
                              "frac_of_8TBs": round((40 + (2 * passes + 1) * 8) * n / (hss_ms / 1e3) / 8e12, 4)},
            "overflow_flags": flags.value}


def mesh_terrain(quads=100, steps=50, warmup=60):
    """Concave mesh leg: 4,096 unit cubes (a 64 x 64 grid, 1.5 apart, from
    y = 1) dropped on a static 100 x 100-quad terrain mesh (20,000
    triangles over 100 x 100 units, height 0.2 sin(0.3 x) cos(0.25 z)),
    8 solver iterations, graph replay.  Time per step after the cubes have
    landed, the readback of the bodies included."""
    import time
    from experiments_amd import scenes
    from experiments_amd.world import World, reference_config
    g = np.linspace(-50, 50, quads + 1)
    verts = np.array([(x, 0.2 * np.sin(0.3 * x) * np.cos(0.25 * z), z) for x in g for z in g],
                     np.float32)
    tris = []
    for i in range(quads):
        for j in range(quads):
            a, b = i * (quads + 1) + j, i * (quads + 1) + j + 1
            c, d = (i + 1) * (quads + 1) + j + 1, (i + 1) * (quads + 1) + j
            tris += [(a, b, c), (a, c, d)]
    w = World(reference_config(max_bodies=8192, max_shapes=2 * quads * quads + 8,
                               max_pairs=1 << 20, max_contacts=1 << 18, max_pairs_per_body=64,
                               solver_iterations=8))
    cube = w.add_hull(scenes.cube_points(0.5))
    mesh = w.add_concave_mesh(verts, np.array(tris, np.int32))
    w.add_body(0.0, (0, 0, 0), (0, 0, 0, 1), mesh)
    for i in range(64):
        for j in range(64):
            w.add_body(1.0, (-47 + 1.5 * i, 1.0, -47 + 1.5 * j), (0, 0, 0, 1), cube)
    w.upload()
    w.use_graph(True)
    w.step(warmup)
    w.get_state()
    t = time.perf_counter()
    w.step(steps)
    w.get_state()
    ms = (time.perf_counter() - t) / steps * 1e3
    st = w.stats()
    w.close()
    return {"triangles": len(tris), "cubes": 4096, "steps": steps, "ms_per_step": round(ms, 4),
            "body_steps_per_s": round(4097 / (ms / 1e3), 1), "pairs": st["num_pairs"],
            "contacts": st["num_contacts"], "overflow_flags": st["overflow_flags"]}


def batched_leg(steps=20, warmup=5, sample_every=16):
    """The 4,096 batched scenes (about 400k bodies): body-steps/s of the
    batched graph-replayed step after the piles settle, and the same scenes
    stepped one world and one captured graph per scene on a sample of every
    `sample_every`-th scene (the per-scene rate of the sample stands for all
    scenes)."""
    from experiments_amd import scenes
    scene = build_scene("batched4k")
    n = scene.num_bodies
    w = scenes.build_world(scene)
    w.use_graph(True)
    w.step(SETTLE["batched4k"] + warmup)
    w.step(steps)
    ms = w.last_step_ms()
    st = w.stats()
    value = n * steps / (ms / 1e3)
    w.close()
    subs = scenes.split_scenes(scene)[::sample_every]
    worlds, bodies = [], 0
    for sub in subs:
        bodies += sub.num_bodies
        ws = scenes.build_world(sub, max_pairs=max(4096, 16 * sub.num_bodies),
                                max_contacts=max(2048, 8 * sub.num_bodies),
                                max_hull_vertices=1 << 16)
        ws.use_graph(True)
        ws.step(SETTLE["batched4k"] + warmup)
        worlds.append(ws)
    t0 = time.perf_counter()
    for ws in worlds:
        ws.step(steps)
    host_s = time.perf_counter() - t0
    for ws in worlds:
        ws.close()
    per_scene = bodies * steps / host_s
    return {"bodies": n, "scenes": 4096, "steps": steps,
            "timed_steps": [SETTLE["batched4k"] + warmup + 1,
                            SETTLE["batched4k"] + warmup + steps],
            "ms_per_step": round(ms / steps, 4), "value": round(value, 1),
            "contacts": st["num_contacts"],
            "per_scene_launch_sets": {"value": round(per_scene, 1), "scenes": len(subs),
                                      "bodies": bodies, "sample": "every %dth scene"
                                      % sample_every, "host_s": round(host_s, 4)},
            "batched_over_per_scene": round(value / per_scene, 2),
            "overflow_flags": st["overflow_flags"]}


def warm_start_leg(n_scenes=1024, settles=(240, 600), steps=20, tol=1e-4, max_sweeps=4096):
    """Solver sweeps to a 1e-4 residual for the zero, previous-step and
    learned warm starts on held-out batched piles (another seed than the
    training scenes), each run from the same state: the piles advanced
    `settle` fixed-sweep steps (at 240 many piles are still moving, at 600
    nearly all rest).  Sweeps are counted per scene and step from the
    solver's per-scene record; the cap is high enough that scenes reaching
    it are reported, not averaged in silently.  The trained weights come
    from experiments_amd/warmstart_weights.npz."""
    from experiments_amd import scenes
    from experiments_amd.warmstart import compare_starts, load_weights
    weights = load_weights()
    scene = scenes.batched_piles(n_scenes=n_scenes, seed=scenes.SEED + 1, iterations=10)
    out = {"scenes": n_scenes, "steps": steps, "tol": tol, "max_sweeps": max_sweeps}
    if weights is None:
        out["learned"] = "no trained weights (experiments_amd/warmstart_weights.npz)"
    for settle in settles:
        out["settled_%d" % settle] = compare_starts(scene, weights, settle, steps, tol,
                                                    max_sweeps)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--settle", type=int, default=-1)
    ap.add_argument("--workload", default="hulls64k")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--per-scene-baseline", action="store_true")
    ap.add_argument("--per-scene-max", type=int, default=0)
    ap.add_argument("--no-extras", action="store_true",
                    help="only the workload's own figures (for profiler runs)")
    ap.add_argument("--profile-only", action="store_true",
                    help="stop after the timed window and the eager stage-timing steps (the "
                         "kernel-trace and counter runs use this)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of timing the candidates first")
    ap.add_argument("--eager", action="store_true",
                    help="launch the timed steps stage by stage instead of replaying the "
                         "captured graph (the counter runs use this)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB 5 times (the HBM "
                         "counter calibration for rocprofv3 --pmc) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only; --gpus %d is not supported (sharding the "
              "scenes over several GPUs is designed but not built)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.copy_probe_mib:
        import torch
        src = torch.ones(args.copy_probe_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": src.numel() * 4, "launches": 5}))
        return

    try:
        # torch provides device buffers for the 1M broadphase figures; its HIP
        # runtime must initialise before the library's first HIP call
        import torch
        torch.cuda.init()
    except Exception:
        pass

    from experiments_amd import scenes
    scene = build_scene(args.workload)
    n = scene.num_bodies
    settle = SETTLE[args.workload] if args.settle < 0 else args.settle
    w = scenes.build_world(scene)
    w.use_graph(not args.eager)
    run_ms = 0.0
    if settle:
        w.step(settle)
        run_ms += w.last_step_ms()
    if args.warmup:
        w.step(args.warmup)
        run_ms += w.last_step_ms()
    tuning, tune_ms = w.get_tuning(), None
    if not args.no_autotune and not args.eager:
        # launch shapes timed on a state from the middle of the timed window,
        # then back to the post-warmup state
        snap = w.get_state()
        w.step(max(args.steps // 2, 1))
        tuning, tune_ms = w.autotune(steps=5)
        w.set_state(**snap)
        w.use_graph(True)
    t0 = time.perf_counter()
    w.step(args.steps)           # synchronises on completion
    host_s = time.perf_counter() - t0
    gpu_ms = w.last_step_ms()
    run_ms += gpu_ms
    done = settle + args.warmup + args.steps
    st = w.stats()
    value = n * args.steps / (gpu_ms / 1e3)
    # per-stage breakdown and the solver kernel's own launch times: a few
    # eager steps with HIP events from the end state of the timed window,
    # which is then restored
    after = w.get_state()
    # the solver kernel's launch duration: two eager steps from the end state
    # of the timed window with HIP events before the first and after the last
    # of its back-to-back launches (a kernel trace gives the same duration
    # for the eager and the graph-replayed launches)
    w.use_graph(False)
    w.set_solve_timing(True)
    span_ms, span_launches = 0.0, 0
    for _ in range(2):
        w.step(1)
        s2 = w.stats()
        span_ms += s2["solve_kernel_ms"]
        span_launches += s2["solve_launches"]
    w.set_solve_timing(False)
    w.set_state(**after)
    w.use_graph(False)
    acc = np.zeros(8)
    counts = dict(pairs=0, contacts=0, tile_bodies=0, solve_ms=0.0, launches=0, events_ms=0.0,
                  face_pass=0)
    for _ in range(args.profile_steps):
        w.step(1)
        s2 = w.stats()
        acc += np.array(s2["stage_ms"])
        counts["pairs"] += s2["num_pairs"]
        counts["contacts"] += s2["num_contacts"]
        counts["face_pass"] += s2["face_pass_pairs"]
        counts["tile_bodies"] += s2["solve_bodies"]
        counts["events_ms"] += s2["solve_events_ms"]
        counts["launches"] += s2["solve_launches"]
    k = max(args.profile_steps, 1)
    stage_ms = acc / k
    pairs, contacts = counts["pairs"] / k, counts["contacts"] / k
    tile_bodies = counts["tile_bodies"] / k
    w.set_state(**after)
    w.use_graph(not args.eager)
    prm = w.params()
    digit = tuning.get("sort_digit_bits") or 8  # the body-key sort's digit width
    passes = (prm["key_bits"] + digit - 1) // digit
    sweeps = 2 * scene.iterations
    scene_tiles = prm["tile_mode"] == 1
    # the dominant kernel is the tile solve (spatial tiles: one launch per
    # sweep and parity phase; scene tiles: one launch per step); its launch
    # duration is the span of its back-to-back launches over the launch
    # count; the HIP-event spans around each single launch beside it
    launches = max(counts["launches"] / k, 1)
    launch_s = span_ms / max(span_launches, 1) / 1e3
    launch_events_s = counts["events_ms"] / max(counts["launches"], 1) / 1e3
    per_launch_bytes = solve_bytes(contacts, tile_bodies, sweeps, scene_tiles) / launches
    achieved = per_launch_bytes / launch_s / 1e9 if launch_s > 0 else 0.0
    prof = traffic_profile(args.workload)
    traffic = None
    if prof and prof.get("state_steps") == [settle + args.warmup + 1, done]:
        traffic = prof.get("k_tile_solve_bytes_per_dispatch")
    roofline = {"bound": "hbm", "kernel": "k_tile_solve", "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": round(traffic) if traffic else None,
                "traffic_source": (prof or {}).get("source") if traffic else
                ("no counter profile at this state" if prof is None else
                 "counter profile taken at steps %s, not this window" % prof.get("state_steps")),
                "algorithmic_bytes_per_launch": round(per_launch_bytes),
                "launches_per_step": round(launches, 2), "launch_us": round(launch_s * 1e6, 2),
                "launch_us_events": round(launch_events_s * 1e6, 2),
                "launch_us_source": "HIP events before the first and after the last of the "
                                    "back-to-back solver launches in two eager steps from the "
                                    "timed window's end state, over the launch count; "
                                    "launch_us_events: HIP events around each launch",
                "rows": round(contacts), "tile_bodies": round(tile_bodies)}
    stage_hbm = {}
    for i, name in enumerate(STAGES):
        sb = stage_bytes(name, n, pairs, contacts, passes, sweeps, tile_bodies, scene_tiles)
        if stage_ms[i] > 0:
            stage_hbm[name] = round(sb / (stage_ms[i] / 1e3) / 1e9, 1)
    breakdown = {STAGES[i]: round(float(stage_ms[i]), 4) for i in range(8)}
    # hash + sort of this workload's bodies against the HBM peak: the setup
    # stage (which hashes the cell keys with the rest of the body setup) and
    # the key sort, their algorithmic bytes over their eager-step time
    hs_ms = float(stage_ms[0] + stage_ms[1])
    hs_bytes = stage_bytes("setup", n, 0, 0, passes, 0, 0, 0) + \
        stage_bytes("sort", n, 0, 0, passes, 0, 0, 0)
    hash_sort = {"bytes_per_body": round(hs_bytes / n), "ms": round(hs_ms, 4),
                 "achieved_GBs": round(hs_bytes / (hs_ms / 1e3) / 1e9, 1) if hs_ms > 0 else None,
                 "frac_of_8TBs": round(hs_bytes / (hs_ms / 1e3) / (HBM_PEAK_GBS * 1e9), 4)
                 if hs_ms > 0 else None,
                 "note": "setup (gravity, inertia, AABB, cell key) + cell-key sort, eager stage times"}
    # the rest of the configured run, for the whole-run average
    whole = None
    if done < scene.steps and not args.profile_only:
        w.step(scene.steps - done)
        run_ms += w.last_step_ms()
        whole = {"steps": scene.steps, "ms_per_step": round(run_ms / scene.steps, 4),
                 "value": round(n * scene.steps / (run_ms / 1e3), 1),
                 "note": "every step of the configured run from the initial state, graph "
                         "replays, HIP-event time (the autotuner's trials excluded)"}
    cpu = None
    if args.cpu_seconds > 0 and not args.profile_only:
        w.set_state(**after)
        cpu = cpu_baseline(w, scene, args.cpu_seconds)
    per_scene = None
    if args.per_scene_baseline and args.workload == "batched4k":
        w.close()
        w = None
        per_scene = per_scene_baseline(scene, args.steps, args.warmup, args.per_scene_max or None)
        per_scene["batched_over_per_scene"] = round(value / per_scene["value"], 3)
    out = {
        "metric": "body-steps/s", "value": round(value, 1), "unit": "body-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(gpu_ms / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic: seeded PCG32 scene (random hull shapes, positions, orientations, "
                "velocities, masses)",
        "config": {"workload": args.workload, "bodies": n, "solver_iterations": scene.iterations,
                   "dt": scene.dt, "graph": not args.eager, "settle_steps": settle,
                   "timed_steps": [settle + args.warmup + 1, done],
                   "pairs": int(pairs), "face_pass_pairs": int(counts["face_pass"] / k),
                   "contacts": int(contacts), "key_bits": prm["key_bits"],
                   "sort_passes": passes, "host_s": round(host_s, 4), "tuning": tuning},
        "whole_run": whole,
        "autotune_ms_per_step": tune_ms,
        "stage_ms": breakdown,
        "stage_hbm_GBs": stage_hbm,
        "hash_sort_hbm": hash_sort,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "overflow_flags": st["overflow_flags"],
    }
    if per_scene is not None:
        out["per_scene_launch_sets"] = per_scene
    if not args.no_extras and args.workload == "hulls64k":
        if w is not None:
            out["sap_broadphase"] = sap_on_world(w)
            w.close()
            w = None
        out["broadphase_1m"] = broadphase_1m()
        out["mesh_terrain"] = mesh_terrain()
        out["batched4k"] = batched_leg()
        out["warm_start_iterations"] = warm_start_leg()
    if w is not None:
        w.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
