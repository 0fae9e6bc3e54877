"""The drop-in boundary beyond stepping: deferred body modifications (ObjectAPI::setBody*,
SimulationAPIs/ObjectAPI.h:76-81, applied by ObjectManager::flushCommands,
SimulationImpl/ObjectManager.cpp:638-865) against the oracle's flush, bit for bit, followed by a
step checked phase by phase; an edited static body's collider, AABB and static tree are rebuilt in
the next step; refused commands leave the world unchanged."""
import numpy as np
import pytest

from parity_steps import BODY_MAP, MASS_MAP, check_step, extend_scene, oracle_from_gpu

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def make_world(agx, scene, **settings):
    b = agx.Builder(1)
    agx.build_scene(b, scene)
    return agx.World(b, agx.default_settings(**{**scene.get("settings", {}), **settings}))


def _commands(agx, w):
    dyn = np.flatnonzero(w.body("inv_mass") != 0)
    b = [int(x) for x in dyn[[3, 50, 400, 901, 1500, 2000]]]
    return [
        (agx.SET_POSITION, b[0], 5.0, 40.0),           # teleport a body above the pile
        (agx.SET_ROTATION, b[0], 7.5, 0.0),            # wrapped to [0, 2 pi)
        (agx.SET_VELOCITY, b[1], 2.5, -1.0),
        (agx.SET_ANGULAR_VELOCITY, b[2], 3.0, 0.0),
        (agx.SET_MASS, b[3], 4.0, 0.0),                # inertia scales with the mass
        (agx.SET_INERTIA, b[4], 0.75, 0.0),
        (agx.SET_VELOCITY, b[5], 0.0, 6.0),
        (agx.SET_VELOCITY, b[5], 1.0, 3.0),            # the later command wins
    ]


def test_deferred_body_commands_match_oracle_flush_and_step(agx, oracle, scenes):
    scene = scenes.pile_small(seed=1)
    w = make_world(agx, scene, track_errors=1)
    w.step(DT, 120)
    ow = oracle_from_gpu(oracle, scene, w, dict(v=8, p=3))
    for kind, body, a, b in _commands(agx, w):
        w.queue_body(kind, body, a, b)
        ow.queue_body(kind, body, a, b)
    before = w.body("com_x").copy()
    w.flush()
    ow.flush()
    for g, o in BODY_MAP + MASS_MAP:
        assert np.array_equal(w.body(g), ow.get(o)), g
    assert not np.array_equal(w.body("com_x"), before)
    rot = w.body("rotation")
    assert 0.0 <= rot.min() and rot.max() < 2 * np.pi
    # the next step, phase by phase, from the edited state
    out = check_step(agx, oracle, scene, w, DT, 8, 3, min_manifolds=300)
    assert out["pairs"] > 300


def test_static_body_edit_rebuilds_static_collider_and_tree(agx, oracle, scenes):
    scene = scenes.pile_small(seed=1)
    w = make_world(agx, scene)
    w.step(DT, 30)
    statics = np.flatnonzero(w.body("inv_mass") == 0)
    assert len(statics) == 4
    wall = int(statics[0])
    ow = oracle_from_gpu(oracle, scene, w, dict(v=8, p=3))
    w.queue_body(agx.SET_POSITION, wall, 1.5, 3.0)
    ow.queue_body(agx.SET_POSITION, wall, 1.5, 3.0)
    w.flush()
    ow.flush()
    for p in range(agx.PHASE_BROAD + 1):
        w.run_phase(p, DT)
    ow.snapshot()
    ow.integrate(DT)
    ow.wrap_cossin()
    ow.poses(True)
    ow.aabbs(True)
    for g, o in (("world_x", "wX"), ("world_y", "wY"), ("aabb_min_x", "minX"), ("aabb_min_y", "minY"),
                 ("aabb_max_x", "maxX"), ("aabb_max_y", "maxY")):
        assert np.array_equal(w.collider(g), ow.get(o)), g
    ga, gb = w.pairs()
    for f, o in (("aabb_min_x", "minX"), ("aabb_min_y", "minY"), ("aabb_max_x", "maxX"), ("aabb_max_y", "maxY")):
        ow.set(o, w.collider(f))
    oa, ob = ow.broad_brute()
    from parity_steps import pair_keys
    ns = w.counts.static_colliders
    assert np.array_equal(pair_keys(ga, gb, ns), pair_keys(oa, ob, ns))
    for p in range(agx.PHASE_NARROW, agx.PHASE_FINISH + 1):
        w.run_phase(p, DT)
    assert w.stats().overflow == 0


def test_refused_body_commands(agx, scenes):
    scene = scenes.pile_small(seed=1)
    w = make_world(agx, scene)
    static = int(np.flatnonzero(w.body("inv_mass") == 0)[0])
    dyn = int(np.flatnonzero(w.body("inv_mass") != 0)[0])
    for kind, body, a in ((agx.SET_MASS, static, 2.0), (agx.SET_MASS, dyn, 0.0), (agx.SET_MASS, dyn, -1.0),
                          (99, dyn, 0.0), (agx.SET_VELOCITY, w.counts.bodies, 0.0)):
        with pytest.raises(Exception):
            w.queue_body(kind, body, a, 0.0)
    w.step(DT, 1)
    assert w.stats().overflow == 0


def test_standalone_narrowphase_on_injected_pairs(agx, oracle, scenes):
    """agx_world_narrowphase on the last step's pairs at the post-update poses equals the oracle's
    narrow phase plus warm start on the same poses, field by field; the next step is unchanged."""
    from parity_steps import assert_manifolds_equal
    scene = scenes.pile_small(seed=1)
    w = make_world(agx, scene)
    twin = make_world(agx, scene)
    w.step(DT, 120)
    twin.step(DT, 120)
    ga, gb = w.pairs()
    mg = w.narrowphase(ga, gb)
    ow = oracle_from_gpu(oracle, scene, w, dict(v=8, p=3))
    mo = ow.warm_start(ow.narrow(ga, gb))
    assert len(mg) > 300
    assert_manifolds_equal(mg, mo, "standalone narrow")
    assert np.count_nonzero(mg["jn0"]) > len(mg) // 2
    # reversed pair order gives the same manifolds (canonical A/B), in the new order
    mr = w.narrowphase(gb[::-1], ga[::-1])
    assert_manifolds_equal(mr, mo, "reversed pairs")
    w.step(DT, 1)
    twin.step(DT, 1)
    for f in ("com_x", "com_y", "vel_x", "vel_y", "ang_vel", "rotation"):
        assert np.array_equal(w.body(f), twin.body(f)), f
    with pytest.raises(Exception):
        w.narrowphase(np.array([0], np.uint32), np.array([0], np.uint32))


def _levels(w, mans, order):
    """Levels of a sequential order: one more than the last level at either dynamic body."""
    body_of = w.collider("body")
    dyn = w.body("inv_mass") != 0
    nxt = np.zeros(w.counts.bodies, np.int64)
    top = 0
    for m in order:
        a, b = body_of[mans["a"][m]], body_of[mans["b"][m]]
        lv = max(nxt[a] if dyn[a] else 0, nxt[b] if dyn[b] else 0)
        if dyn[a]:
            nxt[a] = lv + 1
        if dyn[b]:
            nxt[b] = lv + 1
        top = max(top, lv + 1)
    return top


def test_standalone_solve_in_a_given_order(agx, oracle, scenes):
    """agx_world_solve_contacts on the last step's manifolds from the post-update bodies, in a
    sequential order the caller gives (as the reference planner's waves would; here a seeded
    shuffle and the reverse of the world's own order) and in the world's own colouring, against the
    oracle solving the same manifolds from the same bodies in that order: impulses and body
    velocities and positions within 1e-4. The next step starts from the solved bodies; an order that
    is not a permutation is refused and changes nothing."""
    from parity_steps import assert_solved_close
    scene = scenes.pile_small(seed=1)
    w = make_world(agx, scene)
    w.step(DT, 120)
    mans = w.manifolds(1)
    assert len(mans) > 300
    own, _ = w.solve_order()
    rng = np.random.default_rng(5)
    for label, order in (("shuffle", rng.permutation(len(mans)).astype(np.uint32)),
                         ("reverse", own[::-1].copy()), ("own colouring", None)):
        if order is not None:
            print(label, "levels", _levels(w, mans, order), flush=True)
        ow = oracle_from_gpu(oracle, scene, w, dict(v=8, p=3))
        before = {g: w.body(g).copy() for g, _ in BODY_MAP}
        solved_g = w.solve_contacts(mans, order)
        if order is None:
            order, _ = w.solve_order()
        solved_o = ow.solve(mans, order=order, v_it=8, p_it=3)
        assert_solved_close(solved_g, solved_o, label)
        for g, o in (("vel_x", "velX"), ("vel_y", "velY"), ("ang_vel", "angVel"), ("com_x", "comX"), ("com_y", "comY")):
            np.testing.assert_allclose(w.body(g), ow.get(o), rtol=1e-4, atol=1e-4, err_msg=f"{label} {g}")
        assert not np.array_equal(before["vel_y"], w.body("vel_y"))
    bad = np.zeros(len(mans), np.uint32)
    snap = {g: w.body(g).copy() for g, _ in BODY_MAP}
    with pytest.raises(agx.AgxError):
        w.solve_contacts(mans, bad)
    for g in snap:
        assert np.array_equal(snap[g], w.body(g)), g
    w.step(DT, 1)
    assert w.stats().overflow == 0


def _bodies(agx, rows):
    """Body arrays from (shape, x, y, mass, p0, p1) rows."""
    return {"type": np.array([r[0] for r in rows], np.uint8), "px": np.array([r[1] for r in rows], np.float32),
            "py": np.array([r[2] for r in rows], np.float32), "rot": np.zeros(len(rows), np.float32),
            "vx": np.zeros(len(rows), np.float32), "vy": np.zeros(len(rows), np.float32),
            "av": np.zeros(len(rows), np.float32), "mass": np.array([r[3] for r in rows], np.float32),
            "material": np.ones(len(rows), np.uint32), "p0": np.array([r[4] for r in rows], np.float32),
            "p1": np.array([r[5] for r in rows], np.float32)}


def test_rebuild_keeps_the_world_bit_for_bit_and_new_bodies_fall(agx, scenes):
    """Creation after the upload rebuilds the world and moves everything it carried: with no body
    created, the rebuilt world steps bit for bit like a twin that was never rebuilt (bodies,
    warm-start manifolds, islands). A dynamic body created later (ObjectAPI::createCircle after
    the first update) goes to the end of the dynamic group and falls freely. (Bit-for-bit stepping
    of the other bodies is not expected once a body is added: the dynamic colliders' bounds, and
    with them every Morton key, the pair order and the sequential solve order, change.)"""
    scene = scenes.pile_small(seed=1)
    w = make_world(agx, scene)
    twin = make_world(agx, scene)
    w.step(DT, 60)
    twin.step(DT, 60)
    n = w.counts.bodies
    none = _bodies(agx, [])
    assert len(w.create_bodies(0, none)) == 0 and w.counts.bodies == n
    w.step(DT, 30)
    twin.step(DT, 30)
    for f in ("com_x", "com_y", "rotation", "vel_x", "vel_y", "ang_vel"):
        assert np.array_equal(w.body(f), twin.body(f)), f
    assert np.array_equal(w.manifolds(1), twin.manifolds(1))
    assert np.array_equal(w.islands()[0], twin.islands()[0])
    idx = w.create_bodies(0, _bodies(agx, [(agx.SHAPE_CIRCLE, 500.0, 80.0, 2.0, 0.5, 0.0)]))
    assert int(idx[0]) == n and w.counts.bodies == n + 1
    w.step(DT, 30)
    g = w.settings().gravity_y
    assert w.body("vel_y")[n] == pytest.approx(30 * DT * g, rel=1e-4)
    assert w.body("com_x")[n] == 500.0
    assert w.stats().overflow == 0


def test_static_and_dynamic_bodies_created_mid_run_match_the_oracle(agx, oracle, scenes):
    """A static ledge and a box above it created in a settled pile: the static body joins the static
    group (the first dynamic body moves to the end, as ObjectManager's group moves do), every old
    body keeps its state, and the following steps agree with the oracle built from the extended
    scene phase by phase."""
    scene = scenes.pile_small(seed=1)
    w = make_world(agx, scene, track_errors=1)
    w.step(DT, 60)
    old = {f: w.body(f).copy() for f in ("com_x", "com_y", "rotation", "vel_x", "vel_y", "ang_vel")}
    n, ns = w.counts.bodies, w.counts.static_bodies
    rows = [(agx.SHAPE_BOX, 10.0, 55.0, 0.0, 6.0, 0.5), (agx.SHAPE_BOX, 10.0, 57.0, 1.5, 0.8, 0.6)]
    idx = w.create_bodies(0, _bodies(agx, rows))
    assert int(idx[0]) == ns and w.counts.static_bodies == ns + 1 and w.counts.bodies == n + 2
    key_old = {(x, y, r): i for i, (x, y, r) in enumerate(zip(old["com_x"], old["com_y"], old["rotation"]))}
    new_of = np.full(n, -1)
    for j, k in enumerate(zip(w.body("com_x"), w.body("com_y"), w.body("rotation"))):
        if k in key_old:
            new_of[key_old[k]] = j
    assert (new_of >= 0).all() and len(np.unique(new_of)) == n
    for f, v in old.items():
        assert np.array_equal(w.body(f)[new_of], v), f
    assert new_of[ns] == n   # the first dynamic body moved to the end of the dynamic group
    assert int(idx[1]) == n + 1
    ext = extend_scene(scene, _bodies(agx, rows))
    for _ in range(2):
        check_step(agx, oracle, ext, w, DT, min_manifolds=300)
        w.step(DT, 20)
    assert w.body("com_y")[n + 1] > 55.0   # the box rests on the ledge


def test_body_with_several_colliders(agx):
    """ObjectAPI::createBody + create*Collider: a dumbbell (one body, two circles at +-1 and a third
    circle overlapping the right one) keeps its colliders at body pose + rotated offset, never
    collides with itself (no pair while it falls), lands tilted on a static floor, turns flat and
    rests on all three circles, one manifold each, its mass and inertia as given."""
    b = agx.Builder(1)
    mat = b.create_material(0.0, 0.6, 0.6)
    floor = b.create_body(0, 0.0, -0.5, 0.0, 0.0)
    b.add_collider(0, floor, agx.SHAPE_BOX, 40.0, 1.0, material=mat)
    body = b.create_body(0, 0.0, 4.0, 2.0, 1.5, rot=0.3)
    for off in ((1.0, 0.0), (-1.0, 0.0), (0.6, 0.0)):
        b.add_collider(0, body, agx.SHAPE_CIRCLE, 0.5, offset=off, material=mat)
    w = agx.World(b, agx.default_settings(velocity_iterations=8, position_iterations=3))
    assert w.counts.bodies == 2 and w.counts.colliders == 4
    assert w.body("mass")[1] == 2.0 and w.body("inertia")[1] == 1.5
    w.step(1.0 / 60.0, 5)
    assert len(w.pairs()[0]) == 0   # in the air: the body's own overlapping circles make no pair
    x, y, r = w.body("com_x")[1], w.body("com_y")[1], w.body("rotation")[1]
    cx, cy = w.collider("world_x"), w.collider("world_y")
    for c, (ox, oy) in zip(range(1, 4), ((1.0, 0.0), (-1.0, 0.0), (0.6, 0.0))):
        assert cx[c] == pytest.approx(x + np.cos(r) * ox - np.sin(r) * oy, abs=1e-5)
        assert cy[c] == pytest.approx(y + np.sin(r) * ox + np.cos(r) * oy, abs=1e-5)
    w.step(1.0 / 60.0, 300)
    rot = float(w.body("rotation")[1]) % (2 * np.pi)
    assert min(rot, 2 * np.pi - rot) < 1e-2
    assert w.body("com_y")[1] == pytest.approx(0.5, abs=2e-2)
    assert abs(w.body("vel_y")[1]) < 1e-2 and abs(w.body("ang_vel")[1]) < 1e-2
    m = w.manifolds(1)
    assert len(m) == 3 and set(np.minimum(m["a"], m["b"]).tolist()) == {0}
    assert w.stats().overflow == 0


def _removal_order(n, n_static, removed):
    """Dense order after ObjectManager::destroyBodyImmediate of each body in `removed` (old dense
    indices): a static body swaps with the last static body and the static group shrinks, then the
    body swaps with the last body and is dropped."""
    order = list(range(n))
    for r in removed:
        p = order.index(r)
        if p < n_static:
            order[p], order[n_static - 1] = order[n_static - 1], order[p]
            n_static -= 1
            p = n_static
        order[p], order[-1] = order[-1], order[p]
        order.pop()
    return order, n_static


def test_bodies_removed_mid_run_match_the_oracle(agx, oracle, scenes):
    """ObjectAPI::destroyBody after the upload: the bodies leave with their colliders, each moved to
    the end of its arrays through its group as the reference moves it; every other body keeps its
    state at its new index, and the following steps agree with the oracle built from the remaining
    bodies in the new dense order, phase by phase."""
    scene = scenes.pile_small(seed=1)
    w = make_world(agx, scene, track_errors=1)
    w.step(DT, 60)
    n, ns = w.counts.bodies, w.counts.static_bodies
    old = {f: w.body(f).copy() for f in ("com_x", "com_y", "rotation", "vel_x", "vel_y", "ang_vel")}
    removed = [ns + 5, n - 1, 1, ns + 700, 900]   # body 1 is a static wall
    w.remove_bodies(removed)
    order, ns2 = _removal_order(n, ns, removed)
    assert w.counts.bodies == n - len(removed) and w.counts.static_bodies == ns2 == ns - 1
    for f, v in old.items():
        assert np.array_equal(w.body(f), v[order]), f
    bodies = {k: (v[order] if k not in ("verts",) else v) for k, v in scene["bodies"].items()}
    reduced = dict(scene, bodies=bodies)
    for _ in range(2):
        check_step(agx, oracle, reduced, w, DT, min_manifolds=300)
        w.step(DT, 10)
    with pytest.raises(agx.AgxError):
        w.remove_bodies([w.counts.bodies])


def test_removing_a_sleeping_body_wakes_its_island(agx, scenes):
    """A body removed from a sleeping island wakes the island (ObjectManager.cpp:394-395); other
    islands sleep on."""
    scene = scenes.sleepy_stacks()
    w = make_world(agx, scene)
    for _ in range(40):
        w.step(DT, 10)
        if len(np.unique(w.sleep_state()[0][w.sleep_state()[0] != 0])) >= 2:
            break
    ids = w.sleep_state()[0]
    islands = np.unique(ids[ids != 0])
    assert len(islands) >= 2
    victim = int(np.flatnonzero(ids == islands[0])[0])
    others = np.flatnonzero((ids != 0) & (ids != islands[0]))
    keep_pos = w.body("com_x")[others].copy()
    w.remove_bodies([victim])
    ids2 = w.sleep_state()[0]
    assert np.count_nonzero(ids2) == len(others)   # the victim's island woke, every other sleeper sleeps on
    assert set(np.flatnonzero(ids2 != 0).tolist()) <= set(range(w.counts.bodies))
    assert np.array_equal(np.sort(w.body("com_x")[ids2 != 0]), np.sort(keep_pos))
    w.step(DT, 30)
    assert w.stats().overflow == 0
