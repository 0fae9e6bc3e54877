"""Shared GPU-against-oracle checks of one collide-and-solve step, phase by phase.

Bars: integration, poses and AABBs bit for bit; broad-phase pair sets bit for bit; every manifold
field bit for bit (the second point's fields where a manifold has two); solver impulses and body
state within 1e-4 of the oracle stepped in the GPU's exported colour order; the step's
computeConstraintErrors sums within 1e-4 relative of the oracle's on that state; island
partitions bit for bit.
"""
import numpy as np

BODY_MAP = (("com_x", "comX"), ("com_y", "comY"), ("rotation", "rot"), ("vel_x", "velX"), ("vel_y", "velY"),
            ("ang_vel", "angVel"), ("cos", "cos"), ("sin", "sin"))
COLLIDER_MAP = (("world_x", "wX"), ("world_y", "wY"), ("cos", "wCos"), ("sin", "wSin"), ("world_rot", "wRot"),
                ("aabb_min_x", "minX"), ("aabb_min_y", "minY"), ("aabb_max_x", "maxX"), ("aabb_max_y", "maxY"))
SECOND_POINT = ("p1x", "p1y", "id1", "jn1", "jt1")
MASS_MAP = (("mass", "mass"), ("inv_mass", "invMass"), ("inertia", "inertia"), ("inv_inertia", "invInertia"))


def oracle_from_gpu(oracle, scene, w, settings=None):
    """Oracle world with the GPU world's current state (bodies, collider poses, warm-start source)."""
    ow = oracle.world_from_scene(scene, settings)
    for g, o in BODY_MAP + MASS_MAP:
        ow.set(o, w.body(g))
    for g, o in COLLIDER_MAP:
        ow.set(o, w.collider(g))
    ow.set_prev_manifolds(w.manifolds(1))
    if scene.get("joints"):   # the joints' pass order and the motors' warm-start impulses
        ow.set_joint_order(*w.joint_order())
        ow.set_motor_impulses(w.joints()[1])
        ow.set_joint_iterations(list(w.settings().joint_iterations))
    return ow


def sort_key(m):
    return np.argsort((m["a"].astype(np.uint64) << np.uint64(32)) | m["b"].astype(np.uint64), kind="stable")


def pair_keys(a, b, n_static):
    """Order-free keys of a pair list: dynamic x dynamic pairs unordered, dynamic x static as (dyn, static)."""
    a = a.astype(np.uint64)
    b = b.astype(np.uint64)
    both = (a >= n_static) & (b >= n_static)
    lo = np.where(both, np.minimum(a, b), np.where(a >= n_static, a, b))
    hi = np.where(both, np.maximum(a, b), np.where(a >= n_static, b, a))
    return np.sort((lo << np.uint64(32)) | hi)


def assert_manifolds_equal(mg, mo, label=""):
    """Every field bit for bit; the second point's fields where the manifold has two points."""
    assert len(mg) == len(mo), (label, len(mg), len(mo))
    g, o = mg[sort_key(mg)], mo[sort_key(mo)]
    two = g["count"] == 2
    for f in mg.dtype.names:
        sel = two if f in SECOND_POINT else np.ones(len(g), bool)
        assert np.array_equal(g[f][sel], o[f][sel]), (label, f)
    return int(np.count_nonzero(two))


def assert_solved_close(solved_g, solved_o, label=""):
    two = solved_g["count"] == 2
    for f in ("jn0", "jt0", "jn1", "jt1"):
        sel = two if f in SECOND_POINT else np.ones(len(solved_g), bool)
        np.testing.assert_allclose(solved_g[f][sel], solved_o[f][sel], rtol=1e-4, atol=1e-4, err_msg=f"{label} {f}")


def canon_islands(lab):
    return {frozenset(np.flatnonzero(lab == v).tolist()) for v in np.unique(lab) if v != 0xFFFFFFFF}


def check_step(agx, oracle, scene, w, dt, v_it=8, p_it=3, oracle_settings=None, min_manifolds=1000):
    """Runs one step of w phase by phase and checks each phase against the oracle from the same
    state. The world must track constraint errors. Returns a few counts of the step."""
    settings = dict(v=v_it, p=p_it)
    settings.update(oracle_settings or {})
    ow = oracle_from_gpu(oracle, scene, w, settings)
    ow.snapshot()
    ns_col = ow.counts()["static_colliders"]
    # integrate + wrap + cos/sin
    w.run_phase(agx.PHASE_INTEGRATE, dt)
    ow.integrate(dt)
    ow.wrap_cossin()
    for g, o in BODY_MAP:
        assert np.array_equal(w.body(g), ow.get(o)), g
    # poses + AABBs (+ the CCD sweep origin)
    w.run_phase(agx.PHASE_POSE, dt)
    ow.poses()
    ow.aabbs()
    for g, o in COLLIDER_MAP + (("prev_x", "prevWX"), ("prev_y", "prevWY")):
        assert np.array_equal(w.collider(g)[ns_col:], ow.get(o)[ns_col:]), g
    w.run_phase(agx.PHASE_CCD, dt)
    hit, _ = ow.ccd()
    if hit:
        ow.aabbs()
    assert w.collider("world_x").tobytes() == ow.get("wX").tobytes()
    # broad phase: identical pair set
    w.run_phase(agx.PHASE_BROAD, dt)
    ga, gb = w.pairs()
    oa, ob = ow.broad_brute()
    kg = pair_keys(ga, gb, ns_col)
    assert len(np.unique(kg)) == len(kg), "duplicate pairs"
    assert np.array_equal(kg, pair_keys(oa, ob, ns_col))
    # narrow phase + warm start on the GPU's pairs
    w.run_phase(agx.PHASE_NARROW, dt)
    mg = w.manifolds(0)
    mo = ow.warm_start(ow.narrow(ga, gb))
    assert len(mg) >= min_manifolds
    two_point = assert_manifolds_equal(mg, mo, "narrow")
    # solve in the GPU's colour order, from the same pre-solve state
    w.run_phase(agx.PHASE_SOLVE, dt)
    order, starts = w.solve_order()
    assert len(order) == len(mg) and np.array_equal(np.sort(order), np.arange(len(mg)))
    body_of = w.collider("body")
    inv_mass = w.body("inv_mass")
    for k in range(len(starts) - 1):
        ms = mg[order[starts[k]:starts[k + 1]]]
        bodies = np.concatenate([body_of[ms["a"]], body_of[ms["b"]]])
        bodies = bodies[inv_mass[bodies] != 0]
        assert len(np.unique(bodies)) == len(bodies), f"colour {k} shares a dynamic body"
    solved_o = ow.solve(mg, order=order, v_it=v_it, p_it=p_it)
    solved_g = w.manifolds(0)
    assert_solved_close(solved_g, solved_o, "solve")
    for g, o in (("vel_x", "velX"), ("vel_y", "velY"), ("ang_vel", "angVel"), ("com_x", "comX"), ("com_y", "comY")):
        np.testing.assert_allclose(w.body(g), ow.get(o), rtol=1e-4, atol=1e-4, err_msg=g)
    # islands
    w.run_phase(agx.PHASE_ISLANDS, dt)
    ids_g, k_g = w.islands()
    ids_o, k_o = ow.islands(solved_g)
    assert k_g == k_o
    assert canon_islands(ids_g) == canon_islands(ids_o)
    w.run_phase(agx.PHASE_FINISH, dt)
    st = w.stats()
    assert st.overflow == 0 and st.manifolds == len(mg)
    # computeConstraintErrors of the solved state (BodyCollisionSolver.cpp:132-183)
    ev, ep = ow.errors()
    assert abs(st.velocity_error - ev) <= 1e-4 * abs(ev) + 1e-6, (st.velocity_error, ev)
    assert abs(st.position_error - ep) <= 1e-4 * abs(ep) + 1e-6, (st.position_error, ep)
    return dict(pairs=len(ga), manifolds=len(mg), two_point=two_point, colours=len(starts) - 1, islands=k_g)


def extend_scene(scene, extra):
    """The scene with bodies appended in creation order (non-polygon entries of the scene format)."""
    out = dict(scene)
    bodies = dict(scene["bodies"])
    n = len(extra["type"])
    for k in ("type", "px", "py", "rot", "vx", "vy", "av", "mass", "material", "p0", "p1"):
        bodies[k] = np.concatenate([bodies[k], np.asarray(extra[k], bodies[k].dtype)])
    bodies["vstart"] = np.concatenate([bodies["vstart"], np.zeros(n, np.uint32)])
    bodies["vcount"] = np.concatenate([bodies["vcount"], np.zeros(n, np.uint32)])
    out["bodies"] = bodies
    return out
