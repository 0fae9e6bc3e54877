"""Capsules on the device against the oracle: a pile of every shape (capsules, polygons, boxes,
circles) on static capsule and box ramps, with fast bullets fired at the ramps.
- CCD of bullets against the static capsules (CCD/ContinuousCollisionDetector.cpp:222-240): the
  same bodies move back, to the same time of impact within 3e-5;
- after 600 steps, one step phase by phase: poses and capsule AABBs bit for bit, the pair set, every
  manifold field of all ten shape combinations (CollisionMethods/*Capsule.cpp among them) bit for
  bit, the solve in the exported order within 1e-4, islands.
"""
import numpy as np
import pytest

from parity_steps import check_step, oracle_from_gpu

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def make_world(agx, scene, **settings):
    b = agx.Builder(1)
    agx.build_scene(b, scene)
    return agx.World(b, agx.default_settings(**{**scene.get("settings", {}), **settings}))


def test_bullets_into_static_capsules_match_oracle_ccd(agx, oracle, scenes):
    scene = scenes.capsule_pile(seed=7)
    w = make_world(agx, scene, track_errors=1)
    types = w.collider("type")
    assert np.count_nonzero(types == agx.SHAPE_CAPSULE) > 300
    hits = 0
    for step in range(4):
        ow = oracle_from_gpu(oracle, scene, w, dict(v=8, p=3))
        ow.snapshot()
        w.run_phase(agx.PHASE_INTEGRATE, DT)
        w.run_phase(agx.PHASE_POSE, DT)
        ow.integrate(DT)
        ow.wrap_cossin()
        ow.poses()
        ow.aabbs()
        com0 = np.stack([w.body("com_x"), w.body("com_y")], 1)
        assert np.array_equal(com0[:, 0], ow.get("comX")) and np.array_equal(com0[:, 1], ow.get("comY"))
        for g, o in (("aabb_min_x", "minX"), ("aabb_min_y", "minY"), ("aabb_max_x", "maxX"), ("aabb_max_y", "maxY")):
            assert np.array_equal(w.collider(g), ow.get(o)), g
        prev = np.stack([ow.get("prevComX"), ow.get("prevComY")], 1)
        w.run_phase(agx.PHASE_CCD, DT)
        _, t_orc = ow.ccd()
        com1 = np.stack([w.body("com_x"), w.body("com_y")], 1)
        moved = np.any(com1 != com0, axis=1)
        hit_orc = t_orc < 1.0
        assert np.array_equal(moved, hit_orc), step
        if hit_orc.any():
            disp = (com0 - prev)[hit_orc].astype(np.float64)
            back = (com1 - com0)[hit_orc].astype(np.float64)
            t_gpu = 1.0 + np.sum(back * disp, 1) / np.sum(disp * disp, 1)
            np.testing.assert_allclose(t_gpu, t_orc[hit_orc], rtol=0, atol=3e-5)
        hits += int(hit_orc.sum())
        for p in range(agx.PHASE_BROAD, agx.PHASE_FINISH + 1):
            w.run_phase(p, DT)
        w.post_update()
    print(f"ccd hits over 4 steps: {hits}", flush=True)
    assert hits >= 32


def test_capsule_pile_step_matches_oracle_phase_by_phase(agx, oracle, scenes):
    scene = scenes.capsule_pile(seed=7)
    w = make_world(agx, scene, track_errors=1)
    w.step(DT, 600)
    m = w.manifolds(1)
    types = w.collider("type")
    ta, tb = types[m["a"]], types[m["b"]]
    combos = set(zip(np.minimum(ta, tb).tolist(), np.maximum(ta, tb).tolist()))
    assert combos == {(i, j) for j in range(4) for i in range(j + 1)}, combos
    print(f"settled: {len(m)} manifolds, all ten shape combinations", flush=True)
    check_step(agx, oracle, scene, w, DT, min_manifolds=1500)
