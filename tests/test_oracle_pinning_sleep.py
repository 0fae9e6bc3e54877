"""Pins the oracle's sleeping and waking (ObjectManager.cpp:1314-1491, Simulation.cpp:255-351,
NarrowPhaseCollisionDetector.cpp:299-414) against what the rules fix by themselves: a body at rest
in zero gravity falls asleep on the first step its float32 timer reaches the threshold; the speed it
is judged by includes the angular term times the bound radius; a whole island sleeps together and
wakes together when an awake body touches it, with its stored manifolds back in the step; commands
wake a body's island; sleeping bodies do not move and pair only with awake dynamic bodies."""
import numpy as np
import pytest

from test_oracle_pinning import acc_bodies, scene_of

DT = np.float32(1.0 / 60.0)
MAT = ((0.0, 0.6, 0.6),)


def steps_to_sleep(threshold):
    """Quiet steps until the float32 timer (t + dt) reaches the threshold."""
    t, k = np.float32(0.0), 0
    while True:
        t = np.float32(t + DT)
        k += 1
        if t >= np.float32(threshold):
            return k


def _free(oracle, scenes, items, gy=0.0):
    w = scene_of(oracle, acc_bodies(scenes, items), MAT, settings=dict(gy=gy, v=8, p=3))
    w.set_sleep(0.1, 0.5)
    return w


def test_body_at_rest_sleeps_when_its_timer_reaches_the_threshold(oracle, scenes):
    w = _free(oracle, scenes, [dict(t=0, px=0.0, py=0.0, vx=0.06, mass=1.0, p0=0.5)])
    k = steps_to_sleep(0.5)
    for i in range(k - 1):
        w.update(DT, 1)
        assert w.sleep_state()[0][0] == 0, i
    x_before = w.get("comX")[0]
    w.update(DT, 1)
    ids, timers = w.sleep_state()
    assert ids[0] == 2 + 0 and timers[0] == 0.0          # its own island: id 2 + smallest body index
    x_asleep = w.get("comX")[0]
    assert x_asleep == pytest.approx(x_before + 0.06 * DT, rel=1e-6)   # the step it fell asleep still moved it
    for _ in range(5):
        w.update(DT, 1)
    assert w.get("comX")[0] == x_asleep                   # asleep: not integrated
    assert w.get("velX")[0] == np.float32(0.06)           # and it keeps its velocity


@pytest.mark.parametrize("reach_speed,sleeps", [(0.095, True), (0.105, False)])
def test_angular_speed_counts_through_the_bound_radius(oracle, scenes, reach_speed, sleeps):
    # a 2 x 1 box: bound radius = |offset| + the larger half extent (ObjectManager.cpp:30-43, :136), 1.0
    radius = 1.0
    w = _free(oracle, scenes, [dict(t=1, px=0.0, py=0.0, av=float(reach_speed / radius), mass=1.0, p0=2.0, p1=1.0)])
    assert w.get("boundRadius")[0] == pytest.approx(radius, rel=1e-6)
    for _ in range(steps_to_sleep(0.5) + 2):
        w.update(DT, 1)
    assert (w.sleep_state()[0][0] != 0) == sleeps


def _stack(oracle, scenes, ball_y=None):
    items = [dict(t=1, px=0.0, py=-0.5, mass=0.0, p0=40.0, p1=1.0),            # floor, top at y = 0
             dict(t=1, px=0.0, py=0.5, mass=1.0, p0=1.0, p1=1.0),              # box on the floor
             dict(t=1, px=0.1, py=1.5, mass=1.0, p0=1.0, p1=1.0),              # box on the box
             dict(t=1, px=8.0, py=0.5, mass=1.0, p0=1.0, p1=1.0)]              # a lone box elsewhere
    if ball_y is not None:
        items.append(dict(t=0, px=0.1, py=ball_y, mass=0.5, p0=0.3))
    w = scene_of(oracle, acc_bodies(scenes, items), MAT, settings=dict(gy=-9.81, v=8, p=3))
    w.set_sleep(0.1, 0.5)
    return w


def test_island_sleeps_together_stashes_and_wakes_on_contact(oracle, scenes):
    w = _stack(oracle, scenes, ball_y=30.0)
    body_of = w.get("colBody")
    for _ in range(150):
        w.update(DT, 1)
        ids = w.sleep_state()[0]
        if ids[1] and ids[2] and ids[3]:
            break
    ids = w.sleep_state()[0]
    stack_id = 2 + min(1, 2)
    assert ids[1] == ids[2] == stack_id and ids[3] == 2 + 3 and ids[4] == 0 and ids[0] == 0
    stash, isl = w.stash()
    assert len(stash) >= 3                                  # box-floor, box-box, lone box-floor
    assert set(isl.tolist()) <= {stack_id, 2 + 3}
    for m, i in zip(stash, isl):
        a, b = body_of[m["a"]], body_of[m["b"]]
        assert i == (ids[a] if ids[a] else ids[b])
    assert w.counts()["prev_manifolds"] == 0                # every manifold went with its island
    frozen = (w.get("comX")[:4].copy(), w.get("comY")[:4].copy())
    stash_stack = stash[isl == stack_id].copy()
    # let the ball fall until it touches the stack
    woke = False
    for _ in range(200):
        w.update(DT, 1)
        ids = w.sleep_state()[0]
        if ids[1] == 0:
            woke = True
            break
        assert np.array_equal(w.get("comX")[:4], frozen[0]) and np.array_equal(w.get("comY")[:4], frozen[1])
    assert woke and ids[2] == 0 and ids[3] == 2 + 3       # the stack woke together, the lone box did not
    stash2, isl2 = w.stash()
    assert set(isl2.tolist()) == {2 + 3}
    # the stack's stored manifolds were solved in that step (warm-started from their stored impulses)
    cur = w.manifolds()
    keys = {(int(m["a"]), int(m["b"])) for m in cur}
    assert all((int(m["a"]), int(m["b"])) in keys for m in stash_stack)


def test_wake_contacts_appends_the_stash_in_order(oracle, scenes):
    w = _stack(oracle, scenes)
    for _ in range(150):
        w.update(DT, 1)
    stash, isl = w.stash()
    assert len(stash) >= 3
    body_of = w.get("colBody")
    # a made-up contact between the lone box and the floor wakes only the lone box's island
    lone = [k for k in range(len(stash)) if isl[k] == 2 + 3]
    probe = stash[lone[:1]].copy()
    probe["jn0"] = 0.0
    out = w.wake_contacts(probe)
    assert len(out) == 1 + len(lone)
    assert out[1:].tobytes() == stash[lone].tobytes()
    assert w.sleep_state()[0][3] == 0 and w.sleep_state()[0][1] != 0
    rest, rest_isl = w.stash()
    assert len(rest) == len(stash) - len(lone) and np.all(rest_isl != 2 + 3)
    assert np.all(body_of[rest["a"]] != 3) and np.all(body_of[rest["b"]] != 3)


def test_commands_wake_islands_and_static_edits_wake_what_rests_on_them(oracle, scenes):
    w = _stack(oracle, scenes)
    for _ in range(150):
        w.update(DT, 1)
    ids = w.sleep_state()[0]
    assert ids[1] and ids[2] and ids[3]
    n_stash = len(w.stash()[0])
    w.queue_body(3, 2, 0.5)                                 # angular velocity on the top box
    w.flush()
    ids = w.sleep_state()[0]
    assert ids[1] == 0 and ids[2] == 0 and ids[3] != 0      # its whole island woke
    assert w.counts()["prev_manifolds"] == n_stash - len(w.stash()[0]) > 0   # stored manifolds restored
    assert w.get("angVel")[2] == np.float32(0.5)
    # moving the static floor drops the lone box's stored floor contact and wakes it
    w.queue_body(0, 0, 0.0, -0.6)
    w.flush()
    assert w.sleep_state()[0][3] == 0 and len(w.stash()[0]) == 0


def test_broad_phase_pairs_sleeping_bodies_only_with_awake_dynamic_ones(oracle, scenes):
    rng = np.random.default_rng(4)
    items = [dict(t=1, px=0.0, py=-0.5, mass=0.0, p0=40.0, p1=1.0)]
    for _ in range(120):
        items.append(dict(t=int(rng.integers(0, 2)), px=float(rng.uniform(-8, 8)), py=float(rng.uniform(0, 6)),
                          rot=float(rng.uniform(0, 3)), mass=1.0, p0=0.6, p1=0.5))
    w = scene_of(oracle, acc_bodies(scenes, items), MAT)
    n = w.counts()["bodies"]
    ids = np.where(rng.random(n) < 0.5, 2 + np.arange(n), 0).astype(np.uint32)
    ids[:w.counts()["static_bodies"]] = 0
    w.set_sleep_state(ids, np.zeros(n, np.float32))
    a, b = w.broad_brute()
    body = w.get("colBody")
    inv = w.get("invMass")
    asleep = ids[body] != 0
    static = inv[body] == 0
    assert not np.any(asleep[a] & (asleep[b] | static[b])) and not np.any(asleep[b] & static[a])
    ab, bb = w.broad_bvh()
    key = lambda x, y: sorted(zip(np.minimum(x, y).tolist(), np.maximum(x, y).tolist()))
    assert key(ab, bb) == key(a, b)
    w.set_sleep_state(np.zeros(n, np.uint32), np.zeros(n, np.float32))
    a0, b0 = w.broad_brute()
    pairs0 = set(key(a0, b0))
    kept = set(key(a, b))
    # exactly the awake-awake, awake-static and awake-sleeping pairs of the full set remain
    expect = {(x, y) for x, y in pairs0
              if not (asleep[x] and (asleep[y] or static[y])) and not (asleep[y] and static[x])}
    assert kept == expect and len(expect) < len(pairs0)
