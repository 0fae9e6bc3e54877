"""World snapshots: a world loaded from a snapshot steps to the same bits as the world that saved
it (bodies, manifolds with their warm-start impulses, islands, statistics, collision events),
whether it is a fresh world built from the same scenes or the saving world rewound; graph-replayed
steps agree too; snapshots of other worlds and damaged snapshots are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0
FIELDS = ("com_x", "com_y", "rotation", "vel_x", "vel_y", "ang_vel", "cos", "sin", "mass", "inv_mass")


def make_world(agx, scene, **settings):
    b = agx.Builder(1)
    agx.build_scene(b, scene)
    return agx.World(b, agx.default_settings(**{**scene.get("settings", {}), **settings}))


def state(w):
    out = {f: w.body(f).tobytes() for f in FIELDS}
    out["manifolds"] = w.manifolds(1).tobytes()
    out["islands"] = w.islands()[0].tobytes()
    out["aabb"] = w.collider("aabb_min_x").tobytes() + w.collider("aabb_max_y").tobytes()
    st = w.stats()
    out["stats"] = (st.manifolds, st.pairs, st.colours, st.islands, st.solved_steps, st.velocity_error, st.position_error)
    out["events"] = w.events()[0].tobytes()
    return out


def test_snapshot_round_trip_steps_to_the_same_bits(agx, scenes):
    scene = scenes.pile_small(seed=2)
    a = make_world(agx, scene, track_errors=1)
    a.step(DT, 50)
    for c in range(0, 300, 7):
        a.set_collider_callback(c, 3, 0.05)
    dyn = np.flatnonzero(a.body("inv_mass") != 0)
    a.queue_body(agx.SET_VELOCITY, int(dyn[10]), 1.0, 4.0)      # queued, not yet applied: part of the state
    snap = a.save()
    print(f"snapshot {len(snap)} bytes", flush=True)
    a.step(DT, 20)
    ref = state(a)
    b = make_world(agx, scene, track_errors=1)
    b.load(snap)
    b.step(DT, 20)
    assert state(b) == ref
    a.load(snap)                                                  # rewind the saving world
    a.step(DT, 20)
    assert state(a) == ref
    c = make_world(agx, scene, track_errors=1)
    c.load(snap)
    c.step_graph(DT, 20)
    assert state(c) == ref
    assert b.save() == a.save()


def test_snapshot_refuses_other_worlds_and_damage(agx, scenes):
    a = make_world(agx, scenes.pile_small(seed=2))
    a.step(DT, 5)
    snap = a.save()
    other = make_world(agx, scenes.pile_small(seed=3))
    with pytest.raises(agx.AgxError):
        other.load(snap)
    with pytest.raises(agx.AgxError):
        a.load(snap[: len(snap) // 2])
    bad = bytearray(snap)
    bad[0] ^= 0xFF
    with pytest.raises(agx.AgxError):
        a.load(bytes(bad))
    a.load(snap)                                                  # still loads after the refusals
    a.step(DT, 1)
