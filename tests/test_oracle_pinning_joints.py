"""Pins the oracle's joints (Solving/*Solver.cpp under ConstraintScheduler.cpp:9-84) against closed
forms and physics the constraints must reproduce: a motor reaches its target speed in one pass when
its torque allows and otherwise gains maxTorque * dt / I per step; one spring pass is the soft-
constraint map v <- v - (v + beta C) / (1 + m gamma); a stiff revolute pendulum swings with the
physical pendulum's period; a prismatic joint lets a body slide down its axis at g sin(theta) with
its rotation held, a slider leaves the rotation free, and both stop at their limits; jointed bodies
form one island."""
import math

import numpy as np
import pytest

from test_oracle_pinning import acc_bodies, scene_of

DT = 1.0 / 60.0
MAT = ((0.0, 0.0, 0.0),)


def world(oracle, scenes, items, joints, gy=0.0, it=6):
    w = scene_of(oracle, acc_bodies(scenes, items), MAT, settings=dict(gy=gy, v=8, p=3, ccd=0))
    w.add_joints(joints)
    w.set_joint_iterations([it] * 5)
    return w


@pytest.mark.parametrize("torque,saturated", [(100.0, False), (0.3, True)])
def test_motor_reaches_target_or_gains_its_torque(oracle, scenes, torque, saturated):
    m, r, target = 2.0, 0.5, 5.0
    inertia = m * r * r / 2
    w = world(oracle, scenes, [dict(t=0, px=0.0, py=0.0, mass=m, p0=r)],
              [dict(type=4, body_a=0, target_speed=target, max_torque=torque)])
    for k in range(1, 6):
        w.update(DT, 1)
        omega = float(w.get("angVel")[0])
        if saturated:
            assert omega == pytest.approx(k * torque * DT / inertia, rel=1e-5)
            assert w.motor_impulses()[0] == pytest.approx(torque * DT, rel=1e-6)
        else:
            assert omega == pytest.approx(target, rel=1e-6)
    if not saturated:
        assert abs(w.motor_impulses()[0]) < 1e-5   # after the first step the warm start carries it


@pytest.mark.parametrize("iterations", [1, 3, 6])
def test_spring_pass_is_the_soft_constraint_map(oracle, scenes, iterations):
    m, L, rest, k, c, v0 = 0.7, 2.0, 1.5, 40.0, 0.9, 0.3
    w = world(oracle, scenes, [dict(t=1, px=0.0, py=0.0, mass=0.0, p0=0.2, p1=0.2),
                               dict(t=0, px=L, py=0.0, vx=v0, mass=m, p0=0.2)],
              [dict(type=0, body_a=0, body_b=1, rest_length=rest, stiffness=k, damping=c)], it=iterations)
    gamma = 1.0 / (DT * (c + DT * k))
    beta = DT * k * gamma
    C = L - rest
    v = v0
    for _ in range(iterations):
        v = v - (v + beta * C) / (1.0 + m * gamma)
    w.integrate(DT)
    assert w.get("velX")[1] == pytest.approx(v, rel=1e-5, abs=1e-7)
    assert w.get("comX")[1] == pytest.approx(L + v * DT, rel=1e-6)
    assert w.get("velY")[1] == 0.0


def test_revolute_pendulum_period(oracle, scenes):
    m, rad, L, theta0, g = 1.0, 0.1, 2.0, 0.12, 9.81
    px, py = 0.0, 10.0
    bx, by = px + L * math.sin(theta0), py - L * math.cos(theta0)
    w = world(oracle, scenes, [dict(t=1, px=px, py=py, mass=0.0, p0=0.1, p1=0.1),
                               dict(t=0, px=bx, py=by, mass=m, p0=rad)],
              [dict(type=1, body_a=0, body_b=1, anchor_b_x=px - bx, anchor_b_y=py - by, stiffness=1e5, damping=200.0)],
              gy=-g, it=8)
    xs, pivot_err = [], 0.0
    for _ in range(600):
        w.update(DT, 1)
        xs.append(float(w.get("comX")[1]))
        # the bob's anchor, turned with the bob, stays on the pivot
        rot = float(w.get("rot")[1])
        ax, ay = px - bx, py - by
        wx = w.get("comX")[1] + math.cos(rot) * ax - math.sin(rot) * ay
        wy = w.get("comY")[1] + math.sin(rot) * ax + math.cos(rot) * ay
        pivot_err = max(pivot_err, math.hypot(wx - px, wy - py))
    xs = np.array(xs)
    assert pivot_err < 5e-3
    # downward zero crossings of x (the bob passes under the pivot moving left)
    cross = [i + xs[i] / (xs[i] - xs[i + 1]) for i in range(len(xs) - 1) if xs[i] > 0 >= xs[i + 1]]
    period = (cross[-1] - cross[0]) / (len(cross) - 1) * DT
    inertia_pivot = 0.5 * m * rad * rad + m * L * L
    expect = 2 * math.pi * math.sqrt(inertia_pivot / (m * g * L)) * (1 + theta0 ** 2 / 16)
    assert len(cross) >= 3
    assert period == pytest.approx(expect, rel=0.01)


@pytest.mark.parametrize("kind", [2, 3])
def test_axis_joints_slide_at_g_sin_theta_and_stop_at_the_limit(oracle, scenes, kind):
    theta, g = 0.5, 9.81
    axis = (math.cos(theta), math.sin(theta))
    # the static body sits well away; its anchor is the moving body's starting point
    w = world(oracle, scenes, [dict(t=1, px=0.0, py=8.0, mass=0.0, p0=0.2, p1=0.2),
                               dict(t=1, px=0.0, py=0.0, av=2.0, mass=1.0, p0=0.4, p1=0.4)],
              [dict(type=kind, body_a=0, body_b=1, anchor_a_y=-8.0, axis_x=axis[0], axis_y=axis[1], stiffness=1e5,
                    damping=300.0, lower_limit=-3.0, upper_limit=3.0, enable_limit=1)], gy=-g, it=8)
    n = 30
    for _ in range(n):
        w.update(DT, 1)
    x, y = float(w.get("comX")[1]), float(w.get("comY")[1])
    along = x * axis[0] + y * axis[1]
    across = -x * axis[1] + y * axis[0]
    t = n * DT
    assert along == pytest.approx(-0.5 * g * math.sin(theta) * t * t, rel=0.02)
    assert abs(across) < 2e-3
    omega = float(w.get("angVel")[1])
    if kind == 2:
        assert abs(omega) < 0.05                 # rotation held at the reference angle
    else:
        assert omega == pytest.approx(2.0, rel=1e-5)   # a slider leaves it free
    for _ in range(120):
        w.update(DT, 1)
    x, y = float(w.get("comX")[1]), float(w.get("comY")[1])
    assert x * axis[0] + y * axis[1] == pytest.approx(-3.0, abs=0.03)   # resting on the lower limit


def test_jointed_bodies_form_one_island(oracle, scenes):
    items = [dict(t=0, px=float(3 * k), py=0.0, mass=1.0, p0=0.3) for k in range(5)]
    joints = [dict(type=0, body_a=k, body_b=k + 1, rest_length=3.0, stiffness=10.0, damping=0.1) for k in range(2)]
    joints += [dict(type=1, body_a=3, body_b=4, anchor_a_x=1.5, anchor_b_x=-1.5, stiffness=100.0, damping=1.0),
               dict(type=4, body_a=2, target_speed=1.0, max_torque=1.0)]
    w = world(oracle, scenes, items, joints)
    w.update(DT, 1)
    ids = w.get("islandId")
    assert ids[0] == ids[1] == ids[2] and ids[3] == ids[4] and ids[0] != ids[3]
    assert w.counts()["islands"] == 2
