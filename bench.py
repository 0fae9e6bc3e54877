#!/usr/bin/env python3
"""Benchmark of the collide-and-solve path on one MI355X.

Prints one JSON line. A step is one full physics step (integrate, poses/AABBs, CCD, LBVH broad
phase, SAT narrow phase with warm start, colouring + sequential-impulse solve, islands) plus the
post-update pose pass, i.e. Simulation::update(1/60) with a 1/60 s fixed interval. Bodies are
resident in HBM when a timed region starts.

Headline (``value``): the batch of 1,024 independent scenes (log-uniform 64-4,096 bodies; box
pyramids, random piles, funnels; about 1M bodies), settled at the residual target, then stepped
with velocity sweeps until the relative contact residual is <= 1e-3: body-steps/s of the batched
solver with the learned warm start when it is available, else of the plain (cached) warm start.
Reported beside it: the plain warm start on the same batch, the plain method without batching
(a stratified sample of the scenes, each in a world of its own), the plain solver at a fixed 8
sweeps, the 65,536-body pile with 8 velocity + 3 position sweeps (the plain GPU path against the
single-threaded oracle) and the 1,048,576-collider scatter stepped collide-only with CCD
(pairs/s).
"""
import argparse
import concurrent.futures
import importlib.util
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
PKG_DIR = ROOT / "physicssimulation-agony_amd"
METRIC = "body-steps/s at ≤1e-3 contact residual: learned+batched vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
RESIDUAL_TARGET = 1e-3
# Velocity sweeps a scene may run to reach the target: at 64, 17 % of the settled batch's bodies
# (58 % of its funnels) stop short of it; at 256 more than 99 % of the bodies reach it.
MAX_VELOCITY_SWEEPS = 256


def load_package():
    if "agx" in sys.modules:
        return sys.modules["agx"]
    spec = importlib.util.spec_from_file_location("agx", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["agx"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(segment, st, counts, settings, mean_poly_verts, sweeps):
    """Bytes a segment must move per launch (each array element counted once per pass); `sweeps`
    is the mean number of velocity sweeps of the launches timed."""
    n_dyn_bodies = counts.bodies - counts.static_bodies
    n_dyn_cols = counts.colliders - counts.static_colliders
    m, p = float(st.manifolds), float(st.pairs)
    key_bits = 30
    passes = math.ceil(key_bits / 8)
    table = {
        "integrate": 80.0 * n_dyn_bodies,
        "pose": (56.0 + 8.0 * mean_poly_verts) * n_dyn_cols,
        "tree_build": (24.0 + 16.0 * passes + 36.0 + 84.0) * n_dyn_cols,
        "pair_traversal": 16.0 * n_dyn_cols + 8.0 * p,
        "narrow": (48.0 + 2 * 8.0 * mean_poly_verts) * p + (64.0 + 32.0) * m,
        # constraint prep 290 B + warm start 156 B per manifold; colouring 8 B per manifold and pass
        "colour_prep": (290.0 + 156.0 + 8.0) * m,
        # velocity sweeps 180 B and position sweeps 64 B per manifold and sweep
        "solve": m * (180.0 * sweeps + 64.0 * settings.position_iterations),
        "solve_finish": 32.0 * m,
        "islands": 16.0 * m + 8.0 * n_dyn_bodies,
        "finish": 48.0 * m,
        "post_update": 40.0 * n_dyn_cols,
    }
    return table.get(segment, 0.0)


def calibrate_copy():
    """A plain device-to-device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE readings."""
    import torch
    n = 128 * 1024 * 1024                      # 512 MiB of float32
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"copied {n * 4} bytes three times")


def measured_traffic(kernel, workload):
    """HBM bytes per launch of a kernel from the committed PMC summary, if one exists."""
    path = ROOT / "profiles" / f"{kernel}_{workload}_hbm_traffic.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None



def _cpu_flags():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("flags"):
                return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


_CPU_INFO = None


def cpu_info():
    """Host CPU model, the CPUs this process may use, and the threads the CPU baseline runs on (at
    most 16: the host-core share of one GPU here); read once, before any pool pins a thread."""
    global _CPU_INFO
    if _CPU_INFO is not None:
        return _CPU_INFO
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0))
    _CPU_INFO = {"model": model, "cpus_allowed": allowed, "threads": max(1, min(16, allowed))}
    return _CPU_INFO


def _oracle():
    """The oracle as the CPU baseline: the -O3 AVX2/FMA build when the host has both (the same
    results as the checker build), with the threaded step of the reference's structure."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as orc
    fast = ROOT / "oracle" / "liboracle_fast.so"
    if orc._LIB is None and fast.exists() and {"avx2", "fma"} <= _cpu_flags():
        orc.LIB_PATH = fast
    return orc


BODY_STATE = (("com_x", "comX"), ("com_y", "comY"), ("rotation", "rot"), ("vel_x", "velX"), ("vel_y", "velY"),
              ("ang_vel", "angVel"), ("cos", "cos"), ("sin", "sin"))


def oracle_from_world(orc, scene, world, bsel, csel, v_it, p_it):
    """Oracle world of one scene holding the GPU world's state for that scene's slices."""
    ow = orc.world_from_scene(scene, dict(v=v_it, p=p_it))
    for g, o in BODY_STATE:
        ow.set(o, world.body(g)[bsel])
    ow.set("wX", world.collider("world_x")[csel])
    ow.set("wY", world.collider("world_y")[csel])
    prev = world.manifolds(1)
    c0, c1 = int(csel.start), int(csel.stop)
    pm = prev[(prev["a"] >= c0) & (prev["a"] < c1)].copy()
    pm["a"] -= c0
    pm["b"] -= c0
    ow.set_prev_manifolds(pm)
    return ow


def time_oracle(worlds_and_sizes, budget_s, pool=None):
    """Steps oracle worlds round-robin (one update(1/60) each, through the threaded step when a
    pool is given) for about budget_s of wall time."""
    dt = 1.0 / 60.0
    bodies = steps = 0
    t0 = time.perf_counter()
    while True:
        for ow, n in worlds_and_sizes:
            if pool is None:
                ow.update(dt, 1)
            else:
                ow.update_threaded(pool, dt, 1)
            bodies += n
            steps += 1
            if time.perf_counter() - t0 > budget_s:
                return bodies / (time.perf_counter() - t0), steps, time.perf_counter() - t0


def time_oracle_scenes(worlds_and_sizes, budget_s, threads):
    """Independent scenes stepped side by side: each of `threads` host threads steps its own share
    of the worlds round-robin (the oracle's calls release the interpreter lock), for about budget_s
    of wall time; body-steps/s over all threads."""
    dt = 1.0 / 60.0
    shares = [worlds_and_sizes[k::threads] for k in range(threads)]
    deadline = time.perf_counter() + budget_s

    def run(share):
        bodies = steps = 0
        while time.perf_counter() < deadline:
            for ow, n in share:
                ow.update(dt, 1)
                bodies += n
                steps += 1
        return bodies, steps

    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(run, [s for s in shares if s]))
    el = time.perf_counter() - t0
    return sum(r[0] for r in res) / el, sum(r[1] for r in res), el


USE_GRAPHS = True   # --no-graphs: every step as ordinary launches (for kernel tracing)


def advance(world, dt, steps):
    """`steps` steps replayed from captured step graphs (the first step of a call runs as ordinary
    launches), or all as ordinary launches without graphs."""
    if USE_GRAPHS:
        world.step_graph(dt, steps)
    else:
        world.step(dt, steps)


def step_timed(world, steps, dt):
    """Wall time of `steps` steps."""
    world.sync()
    t0 = time.perf_counter()
    advance(world, dt, steps)
    world.sync()
    return time.perf_counter() - t0


def dynamic_bodies(world):
    return max(world.counts.bodies - world.counts.static_bodies, 1)


def at_target(world, st0, st, steps, elapsed):
    """Body-steps/s of dynamic bodies whose scene's velocity sweeps met the residual target, and
    the fractions of bodies (over the timed steps) and scenes (last step) that met it."""
    conv = st.total_converged_bodies - st0.total_converged_bodies
    return {"value": round(conv / elapsed, 1), "residual_target": RESIDUAL_TARGET,
            "max_velocity_sweeps": MAX_VELOCITY_SWEEPS,
            "fraction_of_body_steps": round(conv / (steps * dynamic_bodies(world)), 5),
            "fraction_of_scenes_last_step": round(st.converged_scenes / max(world.counts.scenes, 1), 5)}


def mean_sweeps(world, stats0, steps):
    """Velocity sweeps per dynamic body and step since `stats0` (scenes solved apart may stop at
    different sweeps; each body counts its scene's)."""
    return (world.stats().total_body_sweeps - stats0.total_body_sweeps) / (steps * dynamic_bodies(world))


def profiled(world, steps, dt):
    """Per-segment GPU ms per launch (HIP events on the world's stream) over `steps` steps, and the
    mean velocity sweeps per dynamic body of those steps."""
    world.set_profiling(True)
    world.profile()
    s0 = world.stats().total_body_sweeps
    world.step(dt, steps)
    world.sync()
    prof = world.profile()
    sweeps = (world.stats().total_body_sweeps - s0) / (steps * dynamic_bodies(world))
    world.set_profiling(False)
    return {k: v[0] / max(v[1], 1) for k, v in prof.items() if v[1]}, sweeps


def roofline(seg_ms, st, counts, settings, mean_v, workload, sweeps):
    dominant = max(seg_ms, key=seg_ms.get)
    per_launch_ms = seg_ms[dominant]
    alg = algorithmic_bytes(dominant, st, counts, settings, mean_v, sweeps)
    achieved = alg / (per_launch_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(dominant, workload),
            "algorithmic_bytes_per_launch": int(alg), "avg_launch_ms": round(per_launch_ms, 4)}


def errors_of(st):
    """computeConstraintErrors of the last step (Simulation.cpp:506-515): summed approach velocity
    and penetration over all contact points (the drift the sweeps left)."""
    return {"velocity": round(float(st.velocity_error), 6), "penetration": round(float(st.position_error), 6)}


def mean_poly_verts(scenes):
    v = np.concatenate([s["bodies"]["vcount"] for s in scenes])
    v = v[v > 0]
    return float(v.mean()) if len(v) else 0.0


def build_world(agx, scenes, settings):
    b = agx.Builder(len(scenes))
    for i, sc in enumerate(scenes):
        agx.build_scene(b, sc, i)
    return agx.World(b, settings)


def scene_ranges(world, num_scenes):
    """Body and collider index ranges of every scene of a batch world (scenes are concatenated)."""
    cs = world.collider("scene")
    cb = world.collider("body")
    bscene = np.empty(world.counts.bodies, np.int64)
    bscene[cb] = cs
    bstart = np.searchsorted(bscene, np.arange(num_scenes + 1))
    cstart = np.searchsorted(cs, np.arange(num_scenes + 1))
    return bstart, cstart


def world_for_scene(agx, scene, batch_world, bsel, csel, settings, prev_manifolds):
    """A one-scene world holding the batch world's current state of that scene: body state, the
    warm-start manifolds (collider indices made local) and the post-update collider poses."""
    w = build_world(agx, [scene], settings)
    for g, _ in BODY_STATE:
        w.set_body(g, batch_world.body(g)[bsel])
    c0, c1 = int(csel.start), int(csel.stop)
    pm = prev_manifolds[(prev_manifolds["a"] >= c0) & (prev_manifolds["a"] < c1)].copy()
    pm["a"] -= c0
    pm["b"] -= c0
    w.set_warm_start_source(pm)
    w.post_update()
    return w


def unbatched_sample(agx, world, scenes, settings, dt, stride, steps):
    """The plain method without batching: every stride-th scene of the batch in a world of its own,
    started from the batch world's current state and stepped `steps` times with ordinary launches
    (Simulation::update per scene, one scene after the other). Returns the sample's bodies, summed
    ms per step and velocity sweeps per dynamic body and step."""
    bstart, cstart = scene_ranges(world, len(scenes))
    prev = world.manifolds(1)
    bodies = dyn = 0
    ms = 0.0
    sweeps = 0.0
    converged = 0.0
    for i in range(0, len(scenes), stride):
        w = world_for_scene(agx, scenes[i], world, slice(int(bstart[i]), int(bstart[i + 1])),
                            slice(int(cstart[i]), int(cstart[i + 1])), settings, prev)
        n = w.counts.bodies
        st0 = w.stats()
        w.sync()
        t0 = time.perf_counter()
        w.step(dt, steps)
        w.sync()
        ms += 1e3 * (time.perf_counter() - t0) / steps
        st1 = w.stats()
        sweeps += (st1.total_body_sweeps - st0.total_body_sweeps) / steps
        converged += (st1.total_converged_bodies - st0.total_converged_bodies) / steps
        bodies += n
        dyn += dynamic_bodies(w)
        w.close()
    return bodies, ms, sweeps / max(dyn, 1), converged, dyn


def collect_training_rows(agx, scenes_mod, mlp, dt, settings, target_rows=10_000_000):
    """Training rows of the learned warm start from the piles of the first two workloads (the
    2,560-body pile over its 600 steps, then the 65,536-body pile from rest until about
    target_rows rows), solved the way the batch is solved: every contact point without a cached
    impulse, with the impulse the solver reached."""
    out = []
    for name, scene, max_steps in (("pile2560", scenes_mod.pile_small(seed=1), 600),
                                   ("pile65536", scenes_mod.pile_large(seed=2), 600)):
        w = build_world(agx, [scene], settings)
        w.set_learned(mlp, agx.LEARN_COLLECT | agx.LEARN_NEW_CONTACTS)
        before, steps = mlp.sample_count(), 0
        while steps < max_steps and (name == "pile2560" or mlp.sample_count() < target_rows):
            w.step(dt, 20)
            steps += 20
        w.sync()
        out.append({"workload": name, "steps": steps, "rows": int(mlp.sample_count() - before)})
        w.set_learned(None, agx.LEARN_OFF)
        w.close()
        log(f"network rows: {out[-1]}")
    return out


def run_batch(agx, scenes_mod, args, dt, learned=True):
    """Plain and learned warm starts on two copies of the batch stepped in lockstep to the 1e-3
    residual target from the same settled state. The network is trained on the GPU on the learned
    copy's untimed steps (rows of contact points without a cached impulse, with the corrections the
    solver itself reached as targets) and then warm-starts those points, behind the residual gate.
    Beside them: the plain method without batching on a stratified sample of the scenes, and the
    plain solver at a fixed 8 velocity sweeps."""
    t0 = time.perf_counter()
    scenes = scenes_mod.batch(1024)
    base = dict(position_iterations=3, track_residual=1, track_errors=1)
    fixed8 = agx.default_settings(velocity_iterations=8, **base)
    to_target = agx.default_settings(velocity_iterations=MAX_VELOCITY_SWEEPS, residual_target=RESIDUAL_TARGET, **base)
    worlds = [build_world(agx, scenes, to_target) for _ in range(2 if learned else 1)]
    plain_w = worlds[0]
    n = plain_w.counts.bodies
    log(f"batch: {len(scenes)} scenes, {n} bodies built in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    for w in worlds:
        for k in range(0, args.settle, 40):
            advance(w, dt, min(40, args.settle - k))
        w.sync()
    log(f"batch: settled {args.settle} steps to the residual target in {time.perf_counter() - t0:.1f} s")
    out = {}
    # launch shape of the dataflow kernels, measured on the plain copy and applied to both
    if args.tuning:
        sb, cb = (int(v) for v in args.tuning.split(","))
        for w in worlds:
            w.set_tuning(sb, cb)
        out["tuning"] = {"solve_blocks_per_cu": sb, "colour_blocks_per_cu": cb, "from_cache": 0, "steps": 0, "given": 1}
    else:
        n0 = plain_w.stats().solved_steps
        tune = plain_w.autotune(dt, 4)
        tuned_steps = plain_w.stats().solved_steps - n0
        for w in worlds[1:]:
            w.set_tuning(tune.solve_blocks_per_cu, tune.colour_blocks_per_cu, tune.scene_whole)
            w.step_graph(dt, tuned_steps)
        out["tuning"] = {"scene_whole": int(tune.scene_whole), "from_cache": int(tune.from_cache),
                         "steps": int(tuned_steps), "ms_per_step": round(float(tune.ms_per_step), 4)}
    log(f"batch: tuned per-scene launch: heaviest {out['tuning'].get('scene_whole')} scenes on whole blocks, "
        f"{out['tuning']['steps']} tuning steps")
    mode = agx.LEARN_PREDICT | agx.LEARN_NEW_CONTACTS
    if learned:
        lw = worlds[1]
        mlp = agx.Mlp(seed=1, sample_capacity=16 << 20)
        t0 = time.perf_counter()
        sources = collect_training_rows(agx, scenes_mod, mlp, dt, to_target)
        t1 = time.perf_counter()
        iterations, loss = 10000, 0.0
        for e in range(iterations // 1000):
            loss = mlp.train(1000, 4096, 1e-3 if e < 7 else 2e-4, seed=e + 1)
        tr = time.perf_counter() - t1
        net = {"rows": mlp.sample_count(), "train_iterations": iterations, "batch": 4096,
               "final_batch_loss": round(loss, 4), "train_s": round(tr, 2),
               "collect_and_train_s": round(time.perf_counter() - t0, 2), "mode": "new contacts, gated",
               "data": "every contact point without a cached impulse in runs of the 2,560-body pile (600 steps) and "
                       "the 65,536-body pile (from rest until the rows reach about 10M), solved to the 1e-3 target "
                       "(features, and the impulses the solver reached, in units of the point's weight impulse)",
               "sources": sources}
        log(f"network: {mlp.sample_count()} rows, loss {loss:.3f}, train {tr:.1f} s")
        lw.set_learned(mlp, mode)
    # timed: each scene's velocity sweeps run until its residual meets the target (at most
    # MAX_VELOCITY_SWEEPS); plain then learned, each from the same step index after the warm-up
    if args.warmup > 0:
        for w in worlds:
            advance(w, dt, args.warmup)
    st0 = plain_w.stats()
    el = step_timed(plain_w, args.steps, dt)
    st = plain_w.stats()
    mean_plain = mean_sweeps(plain_w, st0, args.steps)
    conv = at_target(plain_w, st0, st, args.steps, el)
    out["plain"] = {"value": conv["value"], "all_body_steps_per_s": round(n * args.steps / el, 1),
                    "ms_per_step": round(1e3 * el / args.steps, 4), "at_target": conv,
                    "mean_velocity_sweeps": round(mean_plain, 2), "velocity_sweeps_last_step": int(st.velocity_sweeps),
                    "constraint_errors_last_step": errors_of(st),
                    "residual_per_sweep_last_step": [round(float(st.residual[k]), 6) for k in range(st.velocity_sweeps)]}
    if learned:
        st0 = lw.stats()
        el = step_timed(lw, args.steps, dt)
        st = lw.stats()
        g = lw.gate()
        mean_learned = mean_sweeps(lw, st0, args.steps)
        conv_l = at_target(lw, st0, st, args.steps, el)
        out["learned"] = {"value": conv_l["value"], "all_body_steps_per_s": round(n * args.steps / el, 1),
                          "ms_per_step": round(1e3 * el / args.steps, 4), "at_target": conv_l,
                          "mean_velocity_sweeps": round(mean_learned, 2), "overflow": int(st.overflow),
                          "constraint_errors_last_step": errors_of(st),
                          "sweep_reduction": round(1.0 - mean_learned / max(mean_plain, 1e-9), 4),
                          "gate": {"cost_cached": round(g.cost_cached, 3), "cost_learned": round(g.cost_learned, 3),
                                   "steps_cached": int(g.steps_cached), "steps_learned": int(g.steps_learned)},
                          "network": net}
    else:
        out["learned"] = None
    seg, sweeps = profiled(plain_w, max(5, min(args.steps, 20)), dt)
    st = plain_w.stats()
    out["plain"]["segments_ms_per_step"] = {k: round(v, 4) for k, v in seg.items()}
    out["roofline"] = roofline(seg, st, plain_w.counts, to_target, mean_poly_verts(scenes), "batch1024", sweeps)
    out["roofline"]["mean_velocity_sweeps"] = round(sweeps, 2)
    out["stats"] = {"manifolds": int(st.manifolds), "pairs": int(st.pairs), "colours": int(st.colours),
                    "islands": int(st.islands), "overflow": int(st.overflow)}
    if learned:
        lseg, _ = profiled(lw, max(5, min(args.steps, 20)), dt)
        out["learned"]["segments_ms_per_step"] = {k: round(v, 4) for k, v in lseg.items()}
        for w in worlds[1:]:
            w.close()
    # the plain method without batching, from the same state
    out["plain_unbatched"] = None
    if args.unbatched_stride > 0:
        t0 = time.perf_counter()
        sb, sms, ssw, sconv, sdyn = unbatched_sample(agx, plain_w, scenes, to_target, dt, args.unbatched_stride, 10)
        out["plain_unbatched"] = {
            "value": round(sconv / (sms * 1e-3), 1), "all_body_steps_per_s": round(sb / (sms * 1e-3), 1),
            "fraction_of_body_steps_at_target": round(sconv / max(sdyn, 1), 5),
            "ms_per_batch_step_of_sample": round(sms, 3),
            "body_weighted_mean_velocity_sweeps": round(ssw, 2),
            "sample": f"every {args.unbatched_stride}th scene ({len(range(0, len(scenes), args.unbatched_stride))} "
                      f"scenes, {sb} bodies), each in a world of its own started from the batch's state, 10 steps "
                      f"each with ordinary launches, one scene after the other"}
        log(f"unbatched sample: {sb} bodies, {sms:.1f} ms per step of all sampled scenes, in {time.perf_counter() - t0:.1f} s")
    # the plain solver at the reference's fixed 8 velocity sweeps (last: it leaves the state less converged)
    plain_w.set_settings(fixed8)
    plain_w.step(dt, 2)
    el = step_timed(plain_w, args.steps, dt)
    out["plain_fixed8"] = {"value": round(n * args.steps / el, 1), "ms_per_step": round(1e3 * el / args.steps, 4),
                           "velocity_sweeps": 8}
    plain_w.set_settings(to_target)
    out["bodies"] = n
    out["world"], out["scenes"] = plain_w, scenes
    return out


def batch_cpu_baseline(world, scenes, sweeps, budget_s):
    """The oracle on the host cores: 64 of the batch's settled scenes (a fixed random pick), each
    stepped with update(1/60) at the plain GPU solver's mean sweep count, first on one thread, then
    the scenes side by side on every host thread this process may use (the reference has no batch;
    independent simulations on separate threads are how its CPU path would step many scenes)."""
    orc = _oracle()
    cs = world.collider("scene")
    cb = world.collider("body")
    bscene = np.empty(world.counts.bodies, np.int64)
    bscene[cb] = cs
    rng = np.random.default_rng(0)
    pick = rng.permutation(len(scenes))[:64]
    worlds = []
    for i in pick:
        bi = np.flatnonzero(bscene == i)
        ci = np.flatnonzero(cs == i)
        worlds.append((oracle_from_world(orc, scenes[i], world, slice(int(bi[0]), int(bi[-1]) + 1),
                                         slice(int(ci[0]), int(ci[-1]) + 1), sweeps, 3), len(bi)))
    info = cpu_info()
    one, steps1, t1 = time_oracle(worlds, budget_s / 3)
    value, steps, total = time_oracle_scenes(worlds, budget_s, info["threads"])
    return {"value": round(value, 1), "unit": "body-steps/s", "cores": info["threads"], "kind": "port",
            "one_thread_value": round(one, 1), "cpu": info["model"], "cpus_allowed": info["cpus_allowed"],
            "build": orc.LIB_PATH.name,
            "sample": f"{steps} oracle scene-steps (update(1/60), {sweeps} velocity + 3 position sweeps, the plain GPU "
                      f"solver's mean sweep count at the 1e-3 target) over 64 of the batch's settled scenes, the "
                      f"scenes side by side on {info['threads']} threads, {total:.1f} s; one thread: {steps1} "
                      f"scene-steps in {t1:.1f} s"}


def run_pile(agx, scenes_mod, args, dt, cpu):
    t0 = time.perf_counter()
    scene = scenes_mod.pile_large(seed=2)
    settings = agx.default_settings(**scene["settings"], track_residual=1, track_errors=1)
    world = build_world(agx, [scene], settings)
    n = world.counts.bodies
    for k in range(0, 240, 40):
        world.step(dt, 40)
    world.sync()
    tune = world.autotune(dt, 4) if USE_GRAPHS else world.tuning()
    advance(world, dt, 3)   # captures the step graphs of the chosen tuning outside the timed region
    steps = 100
    el = step_timed(world, steps, dt)
    st = world.stats()
    seg, sweeps = profiled(world, 50, dt)
    st = world.stats()
    out = {"workload": f"pile_large_{n}_bodies", "value": round(n * steps / el, 1), "unit": "body-steps/s",
           "ms_per_step": round(1e3 * el / steps, 4), "velocity_iterations": 8, "position_iterations": 3,
           "segments_ms_per_step": {k: round(v, 4) for k, v in seg.items()},
           "tuning": {"solve_blocks_per_cu": tune.solve_blocks_per_cu, "colour_blocks_per_cu": tune.colour_blocks_per_cu},
           "roofline": roofline(seg, st, world.counts, settings, mean_poly_verts([scene]), "pile65536", sweeps),
           "constraint_errors_last_step": errors_of(st),
           "stats": {"pairs": int(st.pairs), "manifolds": int(st.manifolds), "colours": int(st.colours),
                     "islands": int(st.islands), "overflow": int(st.overflow),
                     "residual_per_sweep": [round(float(st.residual[k]), 6) for k in range(st.velocity_sweeps)]}}
    log(f"pile: {out['ms_per_step']} ms/step in {time.perf_counter() - t0:.1f} s")
    if cpu:
        orc = _oracle()
        info = cpu_info()
        ow = oracle_from_world(orc, scene, world, slice(0, n), slice(0, world.counts.colliders), 8, 3)
        one, steps1, t1 = time_oracle([(ow, n)], args.cpu_seconds / 3)
        ow = oracle_from_world(orc, scene, world, slice(0, n), slice(0, world.counts.colliders), 8, 3)
        pool = orc.Pool(info["threads"])
        value, steps_c, total = time_oracle([(ow, n)], args.cpu_seconds, pool)
        pool.close()
        out["cpu_baseline"] = {"value": round(value, 1), "unit": "body-steps/s", "cores": info["threads"],
                               "kind": "port", "one_thread_value": round(one, 1), "cpu": info["model"],
                               "cpus_allowed": info["cpus_allowed"], "build": orc.LIB_PATH.name,
                               "sample": f"{steps_c} oracle steps (update(1/60)) of the same settled {n}-body pile "
                                         f"with the reference's threading (narrow-phase chunks, wave-planned "
                                         f"solver) on {info['threads']} threads, {total:.1f} s; one thread: "
                                         f"{steps1} steps in {t1:.1f} s"}
        out["speedup_vs_cpu_baseline"] = round(out["value"] / value, 2)
        out["speedup_vs_one_cpu_thread"] = round(out["value"] / one, 2)
    return out


def _state(mass, vx, vy, depth, errors):
    """Momentum, kinetic energy and penetration of one state (dynamic bodies)."""
    dyn = mass > 0
    m, vx, vy = mass[dyn].astype(np.float64), vx[dyn].astype(np.float64), vy[dyn].astype(np.float64)
    d = depth.astype(np.float64)
    return {"px": float(np.sum(m * vx)), "py": float(np.sum(m * vy)), "kinetic": float(0.5 * np.sum(m * (vx * vx + vy * vy))),
            "abs_momentum": float(np.sum(m * np.hypot(vx, vy))), "manifolds": int(len(d)),
            "max_depth": float(d.max()) if len(d) else 0.0, "mean_depth": float(d.mean()) if len(d) else 0.0,
            "velocity_error": float(errors[0]), "position_error": float(errors[1])}


def _gpu_state(world):
    st = world.stats()
    return _state(world.body("mass"), world.body("vel_x"), world.body("vel_y"), world.manifolds(0)["depth"],
                  (st.velocity_error, st.position_error))


def _oracle_state(ow):
    return _state(ow.get("mass"), ow.get("velX"), ow.get("velY"), ow.manifolds()["depth"], ow.errors())


def _drift(series):
    """Linear momentum drift relative to the run's initial momentum magnitude sum."""
    p0 = series[0]
    scale = max(p0["abs_momentum"], 1e-12)
    return [round(math.hypot(s["px"] - p0["px"], s["py"] - p0["py"]) / scale, 9) for s in series]


def run_small(agx, scenes_mod, args, dt, cpu):
    """The 2,560-body pile (2,048 polygons, 512 circles in a 64 x 64 m box), 600 steps from the
    drop at 8 + 3 sweeps, timed in blocks of 50 steps; momentum, kinetic energy and penetration
    after each block. Beside it the same shapes as a free swarm without walls or gravity, where the
    solver must conserve linear momentum: its drift over 600 steps. Both also on the host CPU."""
    out = {"workload": "pile_small_2560_bodies_600_steps", "unit": "body-steps/s", "steps": 600}
    for label, scene in (("pile", scenes_mod.pile_small(seed=1)), ("free_swarm", scenes_mod.free_swarm())):
        settings = agx.default_settings(**scene["settings"], track_errors=1)
        world = build_world(agx, [scene], settings)
        ndyn = dynamic_bodies(world)
        series = [_gpu_state(world)]
        total = 0.0
        for _ in range(12):
            total += step_timed(world, 50, dt)
            series.append(_gpu_state(world))
        res = {"value": round(ndyn * 600 / total, 1), "ms_per_step": round(1e3 * total / 600, 4),
               "series_every_50_steps": series}
        if label == "free_swarm":
            res["momentum_drift"] = _drift(series)
        world.close()
        if cpu:
            orc = _oracle()
            info = cpu_info()
            pool = orc.Pool(info["threads"])
            ow = orc.world_from_scene(scene, dict(v=8, p=3, gy=float(scene["settings"]["gravity_y"])))
            cser = [_oracle_state(ow)]
            ctotal = 0.0
            for _ in range(12):
                t0 = time.perf_counter()
                ow.update_threaded(pool, dt, 50)
                ctotal += time.perf_counter() - t0
                cser.append(_oracle_state(ow))
            pool.close()
            res["cpu_baseline"] = {"value": round(ndyn * 600 / ctotal, 1), "unit": "body-steps/s",
                                   "cores": info["threads"], "kind": "port", "cpu": info["model"],
                                   "build": orc.LIB_PATH.name,
                                   "sample": f"the same 600 steps with the oracle's threaded step ({ctotal:.1f} s)",
                                   "series_every_50_steps": cser}
            if label == "free_swarm":
                res["cpu_baseline"]["momentum_drift"] = _drift(cser)
            res["speedup_vs_cpu_baseline"] = round(res["value"] / res["cpu_baseline"]["value"], 2)
        out[label] = res
        log(f"{label}: {res['ms_per_step']} ms/step")
    out["value"] = out["pile"]["value"]
    return out


def run_scatter(agx, scenes_mod, args, dt, cpu):
    t0 = time.perf_counter()
    scene = scenes_mod.scatter_large(seed=3)
    settings = agx.default_settings(gravity_y=0.0, collide_only=1)
    world = build_world(agx, [scene], settings)
    n = world.counts.bodies - world.counts.static_bodies
    world.step(dt, 5)
    advance(world, dt, 3)   # captures the step graphs outside the timed region
    steps = 20
    el = step_timed(world, steps, dt)
    st = world.stats()
    seg, _ = profiled(world, 10, dt)
    st2 = world.stats()
    out = {"workload": f"scatter_{n}_dynamic_colliders_{world.counts.static_bodies}_static", "unit": "pairs/s",
           "value": round(st.pairs * steps / el, 1), "body_steps_per_s": round(n * steps / el, 1),
           "ms_per_step": round(1e3 * el / steps, 4), "pairs_per_step": int(st.pairs),
           "manifolds_per_step": int(st.manifolds), "ccd_hits_per_step": int(st2.ccd_hits),
           "segments_ms_per_step": {k: round(v, 4) for k, v in seg.items()},
           "roofline": roofline(seg, st2, world.counts, settings, mean_poly_verts([scene]), "scatter1m", 0),
           "overflow": int(st2.overflow)}
    log(f"scatter: {out['ms_per_step']} ms/step in {time.perf_counter() - t0:.1f} s")
    if cpu:
        orc = _oracle()
        info = cpu_info()
        pool = orc.Pool(info["threads"])
        res = {}
        for label, p in (("one", None), ("all", pool)):
            ow = orc.world_from_scene(scene, dict(v=8, p=3))
            t1 = time.perf_counter()
            ow.snapshot(); ow.integrate(dt); ow.wrap_cossin(); ow.poses(); ow.aabbs(); ow.ccd()
            pa, pb = ow.broad_bvh()
            if p is None:
                ow.narrow(pa, pb)
            else:
                ow.narrow_threaded(p, pa, pb)
            res[label] = (len(pa), time.perf_counter() - t1)
            del ow
        pool.close()
        (pairs, total), (_, t_one) = res["all"], res["one"]
        out["cpu_baseline"] = {"value": round(pairs / total, 1), "unit": "pairs/s", "cores": info["threads"],
                               "kind": "port", "one_thread_value": round(pairs / t_one, 1), "cpu": info["model"],
                               "cpus_allowed": info["cpus_allowed"], "build": orc.LIB_PATH.name,
                               "sample": f"one oracle collide-only step (integrate, poses, AABBs, CCD, BVH pairs single-"
                                         f"threaded as in the reference, narrow phase in 1,024-pair chunks on "
                                         f"{info['threads']} threads) of the initial scatter, {total:.1f} s; one "
                                         f"thread {t_one:.1f} s"}
        out["speedup_vs_cpu_baseline"] = round(out["value"] / out["cpu_baseline"]["value"], 2)
    return out


def main():
    cpu_info()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60, help="timed steps of the headline batch")
    ap.add_argument("--settle", type=int, default=240,
                    help="settling steps of the headline batch (part of the workload: the settled batch)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps right before each timed region")
    ap.add_argument("--workload", default="all", choices=["all", "batch1024", "pile65536", "scatter1m", "pile2560"])
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--unbatched-stride", type=int, default=16,
                    help="every n-th scene of the batch is stepped in a world of its own (plain, unbatched); 0: none")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="plain warm start only")
    ap.add_argument("--no-graphs", action="store_true",
                    help="step with ordinary launches instead of captured step graphs (for kernel tracing)")
    ap.add_argument("--tuning", default="",
                    help="'sweep blocks per CU,colouring blocks per CU' for the batch instead of autotuning")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 512 MiB device copy three times (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); how the path would shard is "
              f"described in DESIGN.md and is not run.", file=sys.stderr)
        sys.exit(2)
    global USE_GRAPHS
    USE_GRAPHS = not args.no_graphs
    if args.no_graphs and not args.tuning:
        args.tuning = "1,2"
    agx = load_package()
    import agx.scenes  # noqa: F401
    scenes_mod = sys.modules["agx.scenes"]
    dt = 1.0 / 60.0
    cpu = not args.no_cpu
    if args.workload in ("pile65536", "scatter1m", "pile2560"):
        fn = {"pile65536": run_pile, "scatter1m": run_scatter, "pile2560": run_small}[args.workload]
        print(json.dumps(fn(agx, scenes_mod, args, dt, cpu)), flush=True)
        return
    # the pile and the scatter first: measured after the learned batch in the same process, the
    # pile's steps ran at about half speed (graph-replayed steps only; not seen when the batch ran
    # alone or without graphs), so each workload is measured from a clean process state
    side = {}
    if args.workload == "all":
        side["pile65536"] = run_pile(agx, scenes_mod, args, dt, cpu)
        side["scatter1m"] = run_scatter(agx, scenes_mod, args, dt, cpu)
        side["pile2560"] = run_small(agx, scenes_mod, args, dt, cpu)
    b = run_batch(agx, scenes_mod, args, dt, learned=not args.no_learned)
    head = b["learned"] if b["learned"] else b["plain"]
    out = {
        "metric": METRIC, "value": head["value"], "unit": "body-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batch1024_scenes_{b['bodies']}_bodies_settled_{args.settle}_steps_to_residual_1e-3",
                   "scenes": 1024, "bodies": b["bodies"], "settled_steps": args.settle,
                   "residual_target": RESIDUAL_TARGET, "max_velocity_sweeps": MAX_VELOCITY_SWEEPS,
                   "value_counts": "dynamic bodies of scenes whose sweeps met the residual target",
                   "position_iterations": 3, "dt": dt, "warm_start": "learned" if b["learned"] else "cached",
                   "parallelism": "single GPU"},
        "roofline": b["roofline"],
        "plain": b["plain"], "learned": b["learned"], "plain_unbatched": b["plain_unbatched"],
        "plain_fixed8": b["plain_fixed8"], "stats": b["stats"], "tuning": b["tuning"],
    }
    if b["plain_unbatched"]:
        out["batched_over_unbatched_plain"] = round(head["value"] / b["plain_unbatched"]["value"], 2)
    if b["learned"]:
        out["learned_over_plain_batched"] = round(b["learned"]["value"] / b["plain"]["value"], 3)
    if cpu:
        out["cpu_baseline"] = batch_cpu_baseline(b["world"], b["scenes"], max(int(round(b["plain"]["mean_velocity_sweeps"])), 1),
                                                 args.cpu_seconds)
        out["speedup_vs_cpu_baseline"] = round(head["value"] / out["cpu_baseline"]["value"], 2)
    del b
    out.update(side)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
