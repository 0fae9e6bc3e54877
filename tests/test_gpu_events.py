"""Collision events on the GPU (the event recording of computeConstraintData,
Solving/BodyCollisionSolver.cpp:300-388) against the oracle's recording on the same pre-solve
state, bit for bit; every physics step of one update call kept in step order; graph-replayed steps
record the same events as launched steps; events beyond the capacity are counted, not kept."""
import numpy as np
import pytest

from parity_steps import oracle_from_gpu

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def make_world(agx, scene, **settings):
    b = agx.Builder(1)
    agx.build_scene(b, scene)
    return agx.World(b, agx.default_settings(**{**scene.get("settings", {}), **settings}))


def _callbacks(w, n_colliders, seed=5):
    """Callbacks on about a third of the colliders, thresholds from 0 to about a typical impact."""
    rng = np.random.default_rng(seed)
    picks = np.flatnonzero(rng.random(n_colliders) < 0.35)
    ids = rng.integers(1, 40, len(picks))
    thr = np.where(rng.random(len(picks)) < 0.5, 0.0, rng.uniform(0.0, 0.2, len(picks))).astype(np.float32)
    for c, i, t in zip(picks.tolist(), ids.tolist(), thr.tolist()):
        w.set_collider_callback(c, i, t)
    return picks, ids, thr


def test_events_match_oracle_on_the_same_state(agx, oracle, scenes):
    scene = scenes.pile_small(seed=3)
    w = make_world(agx, scene)
    w.step(DT, 60)                        # bodies falling into the pile: many impacts
    n_col = len(w.collider("world_x"))
    picks, ids, thr = _callbacks(w, n_col)
    for p in range(agx.PHASE_NARROW + 1):
        w.run_phase(p, DT)
    mg = w.manifolds(0)
    ow = oracle_from_gpu(oracle, scene, w, dict(v=8, p=3))
    for c, i, t in zip(picks.tolist(), ids.tolist(), thr.tolist()):
        ow.set_callback(c, i, t)
    eo = ow.collision_events(mg)
    for p in range(agx.PHASE_SOLVE, agx.PHASE_FINISH + 1):
        w.run_phase(p, DT)
    eg, dropped = w.events()
    print(f"manifolds {len(mg)}, events {len(eg)}", flush=True)
    assert dropped == 0 and len(eg) > 20
    assert len(eg) == len(eo)
    for f in ("self", "other", "self_body", "other_body", "px", "py", "nx", "ny", "impulse", "callback", "step"):
        assert np.array_equal(eg[f], eo[f]), f
    assert np.all(eg["impulse"] > 0)
    thr_of = np.full(n_col, np.inf, np.float32)
    thr_of[picks] = thr
    assert np.all(eg["impulse"] >= thr_of[eg["self"]])


def test_every_step_of_a_call_in_order_and_graph_replay(agx, scenes):
    scene = scenes.pile_small(seed=4)
    ws = [make_world(agx, scene) for _ in range(2)]
    for w in ws:
        w.step(DT, 30)
        _callbacks(w, len(w.collider("world_x")))
    ws[0].step(DT, 6)
    ws[1].step_graph(DT, 6)
    e0, d0 = ws[0].events()
    e1, d1 = ws[1].events()
    assert d0 == d1 == 0 and len(e0) > 0
    assert set(np.unique(e0["step"]).tolist()) <= set(range(6)) and np.all(np.diff(e0["step"].astype(int)) >= 0)
    assert len(np.unique(e0["step"])) > 1
    assert e0.tobytes() == e1.tobytes()
    # the next call starts a new list; one step at a time gives the same first step
    ws[0].step(DT, 1)
    e2, _ = ws[0].events()
    assert len(e2) == 0 or np.all(e2["step"] == 0)


def test_capacity_counts_dropped_events_and_switches(agx, scenes):
    scene = scenes.pile_small(seed=3)
    w = make_world(agx, scene, event_capacity=8)
    w.step(DT, 60)
    n_col = len(w.collider("world_x"))
    for c in range(n_col):
        w.set_collider_callback(c, 1, 0.0)
    w.step(DT, 2)
    ev, dropped = w.events()
    assert len(ev) == 8 and dropped > 0
    s = w.settings()
    s.collision_events = 0
    w.set_settings(s)
    w.step(DT, 1)
    assert len(w.events()[0]) == 0
    with pytest.raises(agx.AgxError):
        w.set_collider_callback(n_col, 1, 0.0)
    with pytest.raises(agx.AgxError):
        w.set_collider_callback(0, 1, -1.0)
