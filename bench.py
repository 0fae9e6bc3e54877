#!/usr/bin/env python3
"""Benchmark of the collide-and-solve path on one MI355X.

Prints one JSON line. A step is one full physics step (integrate, poses/AABBs, CCD, LBVH
broad phase, SAT narrow phase with warm start, colouring + sequential-impulse solve, islands)
plus the post-update pose pass, i.e. Simulation::update(1/60) with a 1/60 s fixed interval.
Workload (default): the 65,536-body pile (70 % convex polygons with 3-8 vertices, 30 % circles,
lognormal(ln 0.3, 0.3) radii, densities U[0.5, 2], 512 x 256 m static box, friction 0.6,
restitution 0, dt 1/60, 8 velocity + 3 position iterations), settled by the warm-up steps.
``value`` is body-steps/s with the bodies already resident in HBM.
"""
import argparse
import importlib.util
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
PKG_DIR = ROOT / "physicssimulation-agony_amd"
METRIC = "body-steps/s at ≤1e-3 contact residual: learned+batched vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0   # MI355X HBM3E


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


def algorithmic_bytes(segment, st, counts, settings, mean_poly_verts):
    """Bytes a segment must move per launch (each array element counted once per pass)."""
    n_dyn_bodies = counts.bodies - counts.static_bodies
    n_dyn_cols = counts.colliders - counts.static_colliders
    m, p = float(st.manifolds), float(st.pairs)
    sweeps = max(st.velocity_sweeps, 0)
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


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed PMC summary, if one exists."""
    path = ROOT / "profiles" / f"{kernel}_pile65536_hbm_traffic.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def cpu_baseline(agx, scene, world, budget_s):
    """The oracle (single-threaded CPU restatement) stepping the same settled pile."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as orc
    s = scene["settings"]
    ow = orc.world_from_scene(scene, dict(v=s["velocity_iterations"], p=s["position_iterations"]))
    for g, o in (("com_x", "comX"), ("com_y", "comY"), ("rotation", "rot"), ("vel_x", "velX"), ("vel_y", "velY"),
                 ("ang_vel", "angVel"), ("cos", "cos"), ("sin", "sin")):
        ow.set(o, world.body(g))
    ow.set("wX", world.collider("world_x"))
    ow.set("wY", world.collider("world_y"))
    ow.set_prev_manifolds(world.manifolds(1))
    dt = 1.0 / 60.0
    t0 = time.perf_counter()
    ow.update(dt, 1)
    first = time.perf_counter() - t0
    steps = 1
    extra = max(0, min(200, int(budget_s / max(first, 1e-6)) - 1))
    t1 = time.perf_counter()
    for _ in range(extra):
        ow.update(dt, 1)
    steps += extra
    total = first + (time.perf_counter() - t1)
    n = world.counts.bodies
    return {"value": n * steps / total, "unit": "body-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} oracle steps (update(1/60), single thread) of the same settled {n}-body pile; "
                      f"{total:.1f} s of CPU time", "ms_per_step": 1e3 * total / steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=240)
    ap.add_argument("--workload", default="pile65536", choices=["pile65536", "pile2560"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 512 MiB device copy three times (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU design is described "
              f"in DESIGN.md and is not run.", file=sys.stderr)
        sys.exit(2)

    agx = load_package()
    import agx.scenes as scenes  # noqa: F401
    scenes = sys.modules["agx.scenes"]
    t0 = time.perf_counter()
    scene = scenes.pile_large(seed=2) if args.workload == "pile65536" else scenes.pile_small(seed=1)
    builder = agx.Builder(1)
    agx.build_scene(builder, scene)
    settings = agx.default_settings(**scene["settings"], track_residual=1, track_errors=1)
    world = agx.World(builder, settings)
    log(f"scene {scene['name']}: {world.counts.bodies} bodies built in {time.perf_counter() - t0:.1f} s")
    dt = 1.0 / 60.0
    t0 = time.perf_counter()
    for k in range(0, args.warmup, 20):
        world.step(dt, min(20, args.warmup - k))
    world.sync()
    log(f"warm-up {args.warmup} steps in {time.perf_counter() - t0:.1f} s")

    try:
        import torch
        sync_torch = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync_torch = lambda: None  # noqa: E731
    # timed region: plain stepping, no profiling events
    world.sync()
    sync_torch()
    t0 = time.perf_counter()
    world.step(dt, args.steps)
    world.sync()
    sync_torch()
    elapsed = time.perf_counter() - t0
    st = world.stats()
    n = world.counts.bodies
    value = n * args.steps / elapsed
    log(f"timed {args.steps} steps in {elapsed:.3f} s: {1e3 * elapsed / args.steps:.3f} ms/step, "
        f"{st.pairs} pairs, {st.manifolds} manifolds, {st.colours} colours, overflow={st.overflow}")
    # profiled pass: per-segment GPU time from HIP events on the world's stream
    prof_steps = max(10, min(args.steps, 50))
    world.set_profiling(True)
    world.profile()
    world.step(dt, prof_steps)
    world.sync()
    prof = world.profile()
    world.set_profiling(False)
    st = world.stats()

    vcount = scene["bodies"]["vcount"]
    poly = vcount[vcount > 0]
    mean_v = float(poly.mean()) if len(poly) else 0.0
    seg_ms = {k: v[0] / max(v[1], 1) for k, v in prof.items() if v[1]}
    dominant = max(seg_ms, key=seg_ms.get)
    per_launch_ms = seg_ms[dominant]
    alg = algorithmic_bytes(dominant, st, world.counts, settings, mean_v)
    achieved = alg / (per_launch_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(dominant),
                "algorithmic_bytes_per_launch": int(alg), "avg_launch_ms": round(per_launch_ms, 4)}
    residual = [float(st.residual[k]) for k in range(st.velocity_sweeps)]
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "body-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{scene['name']}_{n}_bodies_plain_solver", "bodies": n,
                   "velocity_iterations": settings.velocity_iterations,
                   "position_iterations": settings.position_iterations, "dt": dt, "warm_start": True,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "segments_ms_per_step": {k: round(v, 4) for k, v in seg_ms.items()},
        "segments_profiled_steps": prof_steps,
        "step_stats": {"pairs": int(st.pairs), "manifolds": int(st.manifolds), "contacts": int(st.contacts),
                       "colours": int(st.colours), "islands": int(st.islands),
                       "residual_per_sweep": [round(r, 6) for r in residual],
                       "penetration_sum": float(st.position_error), "overflow": int(st.overflow)},
    }
    if not args.no_cpu:
        cb = cpu_baseline(agx, scene, world, args.cpu_seconds)
        out["cpu_baseline"] = {k: (round(v, 1) if isinstance(v, float) else v) for k, v in cb.items()}
        out["speedup_vs_cpu_baseline"] = round(value / cb["value"], 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
