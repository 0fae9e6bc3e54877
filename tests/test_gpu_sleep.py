"""Sleeping and waking on the GPU against the oracle, one step at a time from the GPU's own state:
the pair set without sleeping pairs, the manifolds with a woken island's stash appended in stash
order, the sleep ids and timers after the wake and after the sleep pass, the manifolds kept and
stashed, and the islands, bit for bit; solver impulses within 1e-4 in the GPU's colour order. Also:
sleeping bodies stay put, deferred commands wake islands like the oracle's flush, graph-replayed
steps match launched ones, and a snapshot carries the sleeping state."""
import numpy as np
import pytest

from parity_steps import BODY_MAP, COLLIDER_MAP, assert_manifolds_equal, assert_solved_close, canon_islands, \
    oracle_from_gpu, pair_keys

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def make_world(agx, scene, **settings):
    b = agx.Builder(1)
    agx.build_scene(b, scene)
    return agx.World(b, agx.default_settings(**{**scene.get("settings", {}), **settings}))


def oracle_state(oracle, scene, w):
    ow = oracle_from_gpu(oracle, scene, w, dict(v=8, p=3))
    st = scene["settings"]
    ow.set_sleep(st["sleep_speed_threshold"], st["sleep_time_threshold"])
    ids, t = w.sleep_state()
    ow.set_sleep_state(ids, t)
    ow.set_stash(*w.stash())
    return ow


def sleep_step(agx, oracle, scene, w):
    """One step phase by phase against the oracle; returns (bodies woken, bodies put to sleep)."""
    ow = oracle_state(oracle, scene, w)
    ids0 = w.sleep_state()[0].copy()
    ow.snapshot()
    ns_col = ow.counts()["static_colliders"]
    w.run_phase(agx.PHASE_INTEGRATE, DT)
    ow.integrate(DT)
    ow.wrap_cossin()
    for g, o in BODY_MAP:
        assert np.array_equal(w.body(g), ow.get(o)), g
    w.run_phase(agx.PHASE_POSE, DT)
    ow.poses()
    ow.aabbs()
    for g, o in COLLIDER_MAP:
        assert np.array_equal(w.collider(g)[ns_col:], ow.get(o)[ns_col:]), g
    w.run_phase(agx.PHASE_CCD, DT)
    if ow.ccd()[0]:
        ow.aabbs()
    w.run_phase(agx.PHASE_BROAD, DT)
    ga, gb = w.pairs()
    oa, ob = ow.broad_brute()
    assert np.array_equal(pair_keys(ga, gb, ns_col), pair_keys(oa, ob, ns_col))
    w.run_phase(agx.PHASE_NARROW, DT)
    mg = w.manifolds(0)
    generated = ow.warm_start(ow.narrow(ga, gb))
    mo = ow.wake_contacts(generated)
    assert_manifolds_equal(mg, mo, "narrow + wake")
    restored = len(mo) - len(generated)
    if restored:   # the woken islands' stashed manifolds, after the step's own, in stash order
        assert mg[-restored:].tobytes() == mo[-restored:].tobytes()
    ids_w = w.sleep_state()
    ids_wo = ow.sleep_state()
    assert np.array_equal(ids_w[0], ids_wo[0]) and np.array_equal(ids_w[1], ids_wo[1])
    w.run_phase(agx.PHASE_SOLVE, DT)
    order, _ = w.solve_order()
    solved_o = ow.solve(mg, order=order, v_it=8, p_it=3)
    solved_g = w.manifolds(0)
    assert_solved_close(solved_g, solved_o, "solve")
    # the sleep pass from the GPU's solved state, so its decisions see the same velocities
    for g, o in BODY_MAP:
        ow.set(o, w.body(g))
    w.run_phase(agx.PHASE_ISLANDS, DT)
    ids_g, k_g = w.islands()
    ids_o, k_o = ow.islands(solved_g)
    assert k_g == k_o and canon_islands(ids_g) == canon_islands(ids_o)
    kept_o = ow.update_sleeping(DT, solved_g)
    w.run_phase(agx.PHASE_FINISH, DT)
    sg, tg = w.sleep_state()
    so, to = ow.sleep_state()
    assert np.array_equal(sg, so) and np.array_equal(tg, to)
    assert w.manifolds(1).tobytes() == kept_o.tobytes()
    stash_g, isl_g = w.stash()
    stash_o, isl_o = ow.stash()
    assert stash_g.tobytes() == stash_o.tobytes() and np.array_equal(isl_g, isl_o)
    woke = int(np.count_nonzero((ids0 != 0) & (ids_w[0] == 0)))
    slept = int(np.count_nonzero((ids_w[0] == 0) & (sg != 0)))
    return woke, slept


def test_sleep_and_wake_match_oracle_step_by_step(agx, oracle, scenes):
    scene = scenes.sleepy_stacks()
    w = make_world(agx, scene)
    w.step(DT, 40)
    woke = slept = 0
    for k in range(420):
        ids_before = w.sleep_state()[0]
        pos_before = (w.body("com_x").copy(), w.body("com_y").copy(), w.body("rotation").copy())
        a, b = sleep_step(agx, oracle, scene, w)
        woke += a
        slept += b
        # a body asleep through a whole step has not moved
        ids_after = w.sleep_state()[0]
        still = (ids_before != 0) & (ids_after != 0)
        for arr, f in zip(pos_before, ("com_x", "com_y", "rotation")):
            assert np.array_equal(arr[still], w.body(f)[still])
        if woke and slept and k > 200:
            break
    print(f"bodies woken {woke}, put to sleep {slept}", flush=True)
    assert slept > 0 and woke > 0


def test_commands_wake_islands_like_the_oracle_flush(agx, oracle, scenes):
    scene = scenes.sleepy_stacks()
    w = make_world(agx, scene)
    for _ in range(40):
        w.step(DT, 10)
        if np.count_nonzero(w.sleep_state()[0]) > 6:
            break
    ids = w.sleep_state()[0]
    sleeping = np.flatnonzero(ids)
    assert len(sleeping) > 6
    ow = oracle_state(oracle, scene, w)
    body = int(sleeping[len(sleeping) // 2])
    w.queue_body(agx.SET_VELOCITY, body, 0.5, 1.0)
    ow.queue_body(2, body, 0.5, 1.0)
    w.queue_body(agx.SET_POSITION, 0, 0.0, -1.02)     # nudge the static floor: everything on it wakes
    ow.queue_body(0, 0, 0.0, -1.02)
    w.flush()
    ow.flush()
    sg, tg = w.sleep_state()
    so, to = ow.sleep_state()
    assert np.array_equal(sg, so) and np.array_equal(tg, to)
    assert np.count_nonzero(sg) < len(sleeping)
    assert w.stash()[0].tobytes() == ow.stash()[0].tobytes()
    assert w.manifolds(1).tobytes() == ow.prev_manifolds().tobytes()   # the warm-start source with the restored
    for g, o in BODY_MAP:
        assert np.array_equal(w.body(g), ow.get(o)), g


def test_graph_steps_and_snapshot_keep_sleeping_state(agx, scenes):
    scene = scenes.sleepy_stacks(seed=10)
    a = make_world(agx, scene)
    b = make_world(agx, scene)
    a.step(DT, 90)
    b.step_graph(DT, 90)
    assert a.sleep_state()[0].tobytes() == b.sleep_state()[0].tobytes()
    assert np.count_nonzero(a.sleep_state()[0]) > 0
    snap = a.save()
    a.step(DT, 150)
    b.step_graph(DT, 150)
    for f in ("com_x", "com_y", "vel_x", "ang_vel"):
        assert a.body(f).tobytes() == b.body(f).tobytes(), f
    assert a.stash()[0].tobytes() == b.stash()[0].tobytes()
    c = make_world(agx, scene)
    c.load(snap)
    c.step(DT, 150)
    assert c.body("com_x").tobytes() == a.body("com_x").tobytes()
    assert c.sleep_state()[0].tobytes() == a.sleep_state()[0].tobytes()


def test_creation_after_upload_keeps_sleeping_islands(agx, oracle, scenes):
    """A body created while islands sleep: the rebuilt world keeps them asleep (ids 2 + their smallest
    new body index, timers, stashed contacts moved with their colliders), and the following steps
    sleep and wake like the oracle built from the extended scene, phase by phase."""
    from parity_steps import extend_scene
    scene = scenes.sleepy_stacks()
    w = make_world(agx, scene)
    for _ in range(40):
        w.step(DT, 10)
        if np.count_nonzero(w.sleep_state()[0]) > 6:
            break
    ids0, t0 = w.sleep_state()
    stash0, _ = w.stash()
    assert np.count_nonzero(ids0) > 6 and len(stash0) > 0
    n = w.counts.bodies
    extra = {"type": np.array([agx.SHAPE_CIRCLE], np.uint8), "px": np.array([0.0], np.float32),
             "py": np.array([6.0], np.float32), "rot": np.zeros(1, np.float32), "vx": np.zeros(1, np.float32),
             "vy": np.zeros(1, np.float32), "av": np.zeros(1, np.float32), "mass": np.array([1.0], np.float32),
             "material": np.ones(1, np.uint32), "p0": np.array([0.4], np.float32), "p1": np.zeros(1, np.float32)}
    idx = w.create_bodies(0, extra)
    assert int(idx[0]) == n   # a dynamic body joins the end of the dynamic group: nothing moves
    ids1, t1 = w.sleep_state()
    assert np.array_equal(ids1[:n], ids0) and np.array_equal(t1[:n], t0) and ids1[n] == 0
    for v in np.unique(ids1[ids1 != 0]):
        assert v == 2 + np.flatnonzero(ids1 == v).min()
    assert w.stash()[0].tobytes() == stash0.tobytes()
    ext = extend_scene(scene, extra)
    for _ in range(90):
        sleep_step(agx, oracle, ext, w)
