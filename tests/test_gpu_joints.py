"""Joints on the GPU against the oracle: a rig with every joint type beside contacts, stepped phase
by phase from the GPU's state with the oracle replaying the GPU's joint pass order (integration with
the joints between the velocity-Verlet halves bit for bit, then the collide-and-solve checks of
every step, with the joints' island edges); the motors' warm-start impulses; a saturated motor's
closed form; graph-replayed steps and snapshots with joints."""
import numpy as np
import pytest

from parity_steps import check_step, oracle_from_gpu

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def make_world(agx, scene, **settings):
    b = agx.Builder(1)
    agx.build_scene(b, scene)
    return agx.World(b, agx.default_settings(**{**scene.get("settings", {}), **settings}))


def test_joint_rig_steps_match_oracle_phase_by_phase(agx, oracle, scenes):
    scene = scenes.joint_rig()
    w = make_world(agx, scene, track_errors=1)
    order, starts, types = w.joint_order()
    recs, imp = w.joints()
    assert len(order) == len(scene["joints"]) == len(recs) and np.array_equal(np.sort(order), np.arange(len(order)))
    for t in range(5):   # colours of one type share no dynamic body
        for c in range(types[t], types[t + 1]):
            js = recs[order[starts[c]:starts[c + 1]]]
            inv = w.body("inv_mass")
            bodies = [int(b) for j in js for b in ((j["body_a"],) if t == 4 else (j["body_a"], j["body_b"]))
                      if inv[b] != 0]
            assert len(bodies) == len(set(bodies))
    for k, warm in enumerate((30, 40, 50)):
        w.step(DT, warm)
        out = check_step(agx, oracle, scene, w, DT, 8, 3, min_manifolds=5)
        print(f"step check {k}: {out}", flush=True)
    ow = oracle_from_gpu(oracle, scene, w, dict(v=8, p=3))
    motors = recs["type"] == 4
    assert np.all(np.abs(w.joints()[1][motors]) > 0)


def test_joint_integration_bit_for_bit_over_many_steps(agx, oracle, scenes):
    scene = scenes.joint_rig(seed=13)
    w = make_world(agx, scene)
    w.step(DT, 20)
    fields = (("com_x", "comX"), ("com_y", "comY"), ("rotation", "rot"), ("vel_x", "velX"), ("vel_y", "velY"),
              ("ang_vel", "angVel"))
    for _ in range(25):
        ow = oracle_from_gpu(oracle, scene, w, dict(v=8, p=3))
        ow.snapshot()
        w.run_phase(agx.PHASE_INTEGRATE, DT)
        ow.integrate(DT)
        ow.wrap_cossin()
        for g, o in fields:
            assert np.array_equal(w.body(g), ow.get(o)), g
        assert np.array_equal(w.joints()[1], ow.motor_impulses())
        for p in range(agx.PHASE_POSE, agx.PHASE_FINISH + 1):
            w.run_phase(p, DT)
        w.post_update()


def test_saturated_motor_closed_form(agx, scenes):
    import agx.scenes as sc  # noqa: F401
    acc = scenes._Acc()
    acc.add(0, 0.0, 0.0, mass=2.0, p0=0.5)
    scene = {"materials": [(0.0, 0.5, 0.5)], "bodies": acc.bodies(),
             "joints": [dict(type=4, body_a=0, target_speed=50.0, max_torque=0.3)]}
    w = make_world(agx, scene, gravity_y=0.0)
    inertia = 2.0 * 0.25 / 2
    for k in range(1, 5):
        w.step(DT, 1)
        assert w.body("ang_vel")[0] == pytest.approx(k * 0.3 * DT / inertia, rel=1e-5)


def test_graph_steps_and_snapshot_with_joints(agx, scenes):
    scene = scenes.joint_rig(seed=14)
    a = make_world(agx, scene)
    b = make_world(agx, scene)
    a.step(DT, 60)
    b.step_graph(DT, 60)
    for f in ("com_x", "vel_x", "ang_vel"):
        assert a.body(f).tobytes() == b.body(f).tobytes(), f
    assert a.joints()[1].tobytes() == b.joints()[1].tobytes()
    snap = a.save()
    a.step(DT, 40)
    c = make_world(agx, scene)
    c.load(snap)
    c.step(DT, 40)
    assert c.body("com_x").tobytes() == a.body("com_x").tobytes()
    assert c.joints()[1].tobytes() == a.joints()[1].tobytes()
    x = a.body("com_x")
    assert np.all(np.isfinite(x))
