"""Pins the oracle's collision-event recording (Solving/BodyCollisionSolver.cpp:300-388) against
closed forms: the impact impulse of a contact point is its approach speed times its normal effective
mass, 1 / (1/mA + 1/mB + (rA x n)^2 / IA + (rB x n)^2 / IB), computed here in float64 from the
textbook rigid-body formulas rather than from the oracle's own expressions.
"""
import numpy as np
import pytest

from test_oracle_pinning import acc_bodies, scene_of

MAT = ((0.0, 0.0, 0.0),)


def _head_on(oracle, scenes, m1=1.0, m2=3.0, v=2.0):
    w = scene_of(oracle, acc_bodies(scenes, [dict(t=0, px=0.0, py=0.0, vx=v, mass=m1, p0=0.5),
                                              dict(t=0, px=0.99, py=0.0, vx=-v, mass=m2, p0=0.5)]), MAT)
    return w, w.narrow([0], [1])


def test_head_on_impact_impulse_closed_form(oracle, scenes):
    w, m = _head_on(oracle, scenes)
    w.set_callback(0, 7, 0.0)
    w.set_callback(1, 9, 0.0)
    ev = w.collision_events(m)
    assert len(ev) == 2
    expect = 4.0 * (1.0 * 3.0) / (1.0 + 3.0)          # closing speed 4 times the reduced mass
    assert ev["impulse"] == pytest.approx([expect, expect], rel=1e-6)
    a, b = int(m["a"][0]), int(m["b"][0])
    # side A first, its normal from A towards B; side B with the normal reversed
    assert (ev["self"][0], ev["other"][0]) == (a, b) and (ev["self"][1], ev["other"][1]) == (b, a)
    assert (ev["nx"][0], ev["ny"][0]) == (m["nx"][0], m["ny"][0])
    assert (ev["nx"][1], ev["ny"][1]) == (-m["nx"][0], -m["ny"][0])
    assert list(ev["callback"]) == [7 if a == 0 else 9, 9 if a == 0 else 7]
    assert list(ev["self_body"]) == [a, b]                 # one collider per body, same dense order
    assert (ev["px"][0], ev["py"][0]) == (m["p0x"][0], m["p0y"][0])


def test_threshold_callback_and_switch(oracle, scenes):
    w, m = _head_on(oracle, scenes)
    imp = 4.0 * 3.0 / 4.0
    assert len(w.collision_events(m)) == 0                 # no callbacks: nothing recorded
    w.set_callback(1, 5, imp * 1.01)
    assert len(w.collision_events(m)) == 0                 # above the impact impulse
    w.set_callback(1, 5, 0.0)
    exact = float(w.collision_events(m)["impulse"][0])
    w.set_callback(1, 5, exact)                             # exactly at it fires (>=)
    ev = w.collision_events(m)
    assert len(ev) == 1 and ev["self"][0] == 1
    w.set_events_enabled(False)
    assert len(w.collision_events(m)) == 0
    w.set_events_enabled(True)
    w.set_callback(1, 0, 0.0)                               # id 0 clears the callback
    assert len(w.collision_events(m)) == 0


def test_separating_bodies_record_nothing(oracle, scenes):
    w = scene_of(oracle, acc_bodies(scenes, [dict(t=0, px=0.0, py=0.0, vx=-1.0, mass=1.0, p0=0.5),
                                              dict(t=0, px=0.99, py=0.0, vx=1.0, mass=1.0, p0=0.5)]), MAT)
    m = w.narrow([0], [1])
    w.set_callback(0, 1, 0.0)
    w.set_callback(1, 1, 0.0)
    assert len(w.collision_events(m)) == 0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_spinning_box_on_floor_matches_rigid_body_formula(oracle, scenes, seed):
    rng = np.random.default_rng(seed)
    rot = float(rng.uniform(0.1, 0.5))
    vx, vy, av = float(rng.uniform(-1, 1)), float(rng.uniform(-3, -1)), float(rng.uniform(-4, 4))
    w = scene_of(oracle, acc_bodies(scenes, [
        dict(t=1, px=0.0, py=-1.0, mass=0.0, p0=20.0, p1=2.0),                        # static floor, top y = 0
        dict(t=1, px=0.3, py=0.45, rot=rot, vx=vx, vy=vy, av=av, mass=2.0, p0=1.0, p1=1.0)]), MAT)
    m = w.narrow([1], [0])
    assert len(m) == 1
    w.set_callback(1, 3, 0.0)
    ev = w.collision_events(m)
    d = m[0]
    n = np.array([d["nx"], d["ny"]], np.float64)
    a, b = int(d["a"]), int(d["b"])
    com = np.stack([w.get("comX"), w.get("comY")], 1).astype(np.float64)
    inv_m = w.get("invMass").astype(np.float64)
    inv_i = w.get("invInertia").astype(np.float64)
    vel = np.stack([w.get("velX"), w.get("velY")], 1).astype(np.float64)
    om = w.get("angVel").astype(np.float64)
    body = w.get("colBody")
    ba, bb = int(body[a]), int(body[b])
    best = 0.0
    for p in ([d["p0x"], d["p0y"]], [d["p1x"], d["p1y"]])[:int(d["count"])]:
        p = np.array(p, np.float64)
        ra, rb = p - com[ba], p - com[bb]
        cross = lambda r: r[0] * n[1] - r[1] * n[0]
        point_v = lambda bi, r: vel[bi] + om[bi] * np.array([-r[1], r[0]])
        approach = -np.dot(point_v(bb, rb) - point_v(ba, ra), n)
        k = inv_m[ba] + inv_m[bb] + cross(ra) ** 2 * inv_i[ba] + cross(rb) ** 2 * inv_i[bb]
        best = max(best, approach / k)
    if best <= 0:
        assert len(ev) == 0
    else:
        assert len(ev) == 1 and ev["self"][0] == 1
        assert ev["impulse"][0] == pytest.approx(best, rel=2e-5)
        if d["count"] == 2:
            assert ev["px"][0] == pytest.approx(0.5 * (d["p0x"] + d["p1x"]), abs=1e-6)


def test_update_keeps_every_step_in_order(oracle, scenes):
    # a circle dropping onto a static box: one event per step while it is in contact
    w = scene_of(oracle, acc_bodies(scenes, [dict(t=1, px=0.0, py=-1.0, mass=0.0, p0=20.0, p1=2.0),
                                              dict(t=0, px=0.0, py=0.52, vy=-3.0, mass=1.0, p0=0.5)]),
                 ((0.5, 0.0, 0.0),), settings=dict(v=8, p=3))
    w.set_callback(1, 2, 0.0)
    w.update(1.0 / 60.0, 4)
    ev = w.events()
    assert len(ev) >= 1
    assert np.all(np.diff(ev["step"].astype(int)) >= 0) and ev["step"].max() < 4
    assert np.all(ev["self"] == 1) and np.all(ev["callback"] == 2)
    first = ev[ev["step"] == ev["step"].min()][0]
    assert first["ny"] < -0.99                              # from the circle towards the floor
    w.update(1.0 / 60.0, 1)
    assert len(w.events()) <= 1 and (len(w.events()) == 0 or w.events()["step"][0] == 0)
