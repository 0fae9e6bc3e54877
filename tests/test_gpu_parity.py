"""GPU parity: the HIP path against the oracle, phase by phase, on identical inputs.

Bars: integration, poses and AABBs bit for bit (same operation order, IEEE division and
square root, no contraction); broad-phase pair sets bit for bit; manifolds and warm-start
impulses bit for bit (checked at the 1e-5 relative bar as well); solver impulses within 1e-4
of the oracle stepped in the GPU's colour order; island partitions bit for bit.
"""
import numpy as np
import pytest

from parity_steps import assert_manifolds_equal, check_step

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def make_world(agx, scene, **settings):
    b = agx.Builder(1)
    agx.build_scene(b, scene)
    s = agx.default_settings(**{**scene.get("settings", {}), **settings})
    return agx.World(b, s)


def oracle_from_gpu(oracle, scene, w, settings=None):
    """Oracle world with the GPU world's current state (bodies, collider poses, warm-start source)."""
    ow = oracle.world_from_scene(scene, settings)
    for g, o in (("com_x", "comX"), ("com_y", "comY"), ("rotation", "rot"), ("vel_x", "velX"), ("vel_y", "velY"),
                 ("ang_vel", "angVel"), ("cos", "cos"), ("sin", "sin")):
        ow.set(o, w.body(g))
    for g, o in (("world_x", "wX"), ("world_y", "wY"), ("cos", "wCos"), ("sin", "wSin"), ("world_rot", "wRot"),
                 ("aabb_min_x", "minX"), ("aabb_min_y", "minY"), ("aabb_max_x", "maxX"), ("aabb_max_y", "maxY")):
        ow.set(o, w.collider(g))
    ow.set_prev_manifolds(w.manifolds(1))
    return ow


def sort_key(m):
    return np.argsort((m["a"].astype(np.uint64) << np.uint64(32)) | m["b"].astype(np.uint64), kind="stable")


def pair_set(a, b, is_dyn):
    s = set()
    for x, y in zip(a.tolist(), b.tolist()):
        s.add((min(x, y), max(x, y)) if is_dyn[x] and is_dyn[y] else (x, y))
    return s


def brute_pairs(aabb, owner, dyn, scene=None):
    """Vectorised exact pair set: strict overlap, different owner, at least one dynamic, same scene."""
    n = len(owner)
    order = np.argsort(aabb[:, 0], kind="stable")
    mnx = aabb[order, 0]
    out = set()
    hi = np.searchsorted(mnx, aabb[order, 2], side="left")
    for k in range(n):
        i = order[k]
        js = order[k + 1:hi[k]]
        if len(js) == 0:
            continue
        ok = (aabb[i, 0] < aabb[js, 2]) & (aabb[i, 2] > aabb[js, 0]) & (aabb[i, 1] < aabb[js, 3]) & (aabb[i, 3] > aabb[js, 1])
        ok &= owner[js] != owner[i]
        ok &= dyn[js] | dyn[i]
        if scene is not None:
            ok &= scene[js] == scene[i]
        for j in js[ok].tolist():
            if dyn[i] and dyn[j]:
                out.add((min(i, j), max(i, j)))
            elif dyn[i]:
                out.add((i, j))
            else:
                out.add((j, i))
    return out


@pytest.mark.parametrize("n,scenes_n", [(3000, 1), (20000, 1), (5000, 7)])
def test_broadphase_pairs_match_brute_force(agx, n, scenes_n):
    rng = np.random.default_rng(n + scenes_n)
    size = np.sqrt(n) * 1.2
    c = rng.uniform(0, size, (n, 2)).astype(np.float32)
    e = rng.uniform(0.1, 1.2, (n, 2)).astype(np.float32)
    aabb = np.concatenate([c - e, c + e], axis=1).astype(np.float32)
    owner = np.arange(n, dtype=np.uint32)
    owner[1::17] = owner[0::17][: len(owner[1::17])]          # some bodies own two colliders
    dyn = (rng.random(n) > 0.1).astype(np.uint8)
    scene = rng.integers(0, scenes_n, n).astype(np.uint32) if scenes_n > 1 else None
    a, b = agx.broadphase_pairs(aabb, owner, dyn, scene)
    got = pair_set(a, b, dyn.astype(bool))
    assert len(got) == len(a), "duplicate pairs"
    assert got == brute_pairs(aabb, owner, dyn.astype(bool), scene)


def test_broadphase_edge_cases(agx):
    # empty input, one collider, touching (not strictly overlapping) boxes, identical boxes.
    a, b = agx.broadphase_pairs(np.zeros((0, 4), np.float32), np.zeros(0, np.uint32), np.zeros(0, np.uint8))
    assert len(a) == 0
    one = np.array([[0, 0, 1, 1]], np.float32)
    a, b = agx.broadphase_pairs(one, np.zeros(1, np.uint32), np.ones(1, np.uint8))
    assert len(a) == 0
    boxes = np.array([[0, 0, 1, 1], [1, 0, 2, 1], [0, 0, 1, 1], [0, 0, 1, 1]], np.float32)
    a, b = agx.broadphase_pairs(boxes, np.array([0, 1, 2, 2], np.uint32), np.ones(4, np.uint8))
    assert pair_set(a, b, np.ones(4, bool)) == {(0, 2), (0, 3)}   # touching edges do not overlap; 2,3 share a body


def test_step_phases_match_oracle(agx, oracle, scenes):
    scene = scenes.pile_small(seed=1)
    w = make_world(agx, scene, track_errors=1)
    w.step(DT, 90)
    out = check_step(agx, oracle, scene, w, DT, 8, 3)
    assert out["pairs"] > 1000 and out["two_point"] > 100


def test_long_run_matches_oracle_statistics(agx, oracle, scenes):
    # After 300 steps the GPU (colour order) and the oracle (narrow-phase order) have diverged
    # body by body, but the settled pile must agree in aggregate.
    scene = scenes.pile_small(seed=3)
    w = make_world(agx, scene, track_errors=1)
    w.step(DT, 300)
    ow = oracle.world_from_scene(scene, dict(v=8, p=3))
    ow.update(DT, 300)
    ns = w.counts.static_bodies
    y, x = w.body("com_y")[ns:], w.body("com_x")[ns:]
    assert np.all(np.isfinite(y)) and np.all(np.isfinite(x))
    assert np.all(y > -0.5) and np.all(np.abs(x) < 33.0)          # nothing escapes the box
    yo = ow.get("comY")[ns:]
    assert abs(float(y.mean()) - float(yo.mean())) < 0.05 * float(yo.mean())
    st = w.stats()
    assert st.overflow == 0
    pen_gpu = st.position_error / max(st.manifolds, 1)
    pen_orc = float(ow.errors()[1]) / max(len(ow.manifolds()), 1)
    assert pen_gpu < 2.0 * pen_orc + 1e-3
    ke = lambda vx, vy, m: float(np.sum(0.5 * m * (vx * vx + vy * vy)))
    mass = w.body("mass")[ns:]
    assert ke(w.body("vel_x")[ns:], w.body("vel_y")[ns:], mass) < 4.0 * ke(ow.get("velX")[ns:], ow.get("velY")[ns:], mass) + 1.0


def test_islands_match_oracle_every_step(agx, oracle, scenes):
    scene = scenes.pile_small(seed=5)
    w = make_world(agx, scene)
    w.step(DT, 60)
    ow = oracle.world_from_scene(scene)
    canon = lambda lab: {frozenset(np.flatnonzero(lab == v).tolist()) for v in np.unique(lab)}
    for _ in range(8):
        w.step(DT, 1)
        m = w.manifolds(0)
        ids_g, k_g = w.islands()
        ids_o, k_o = ow.islands(m)
        assert k_g == k_o
        assert canon(ids_g) == canon(ids_o)
        ns = w.counts.static_bodies
        assert np.all(ids_g[:ns] == 0xFFFFFFFF)
        # ids number the components by their smallest body index
        first = [int(np.flatnonzero(ids_g == v)[0]) for v in range(k_g)]
        assert first == sorted(first)


def test_ccd_matches_oracle(agx, oracle, scenes):
    # fast circles and polygons shot at thin static walls
    acc = scenes._Acc()
    acc.add(1, 20.0, 0.0, p0=0.3, p1=40.0)
    acc.add(2, -20.0, 0.0, rot=0.3, verts=scenes.polygon_outline(np.random.default_rng(0), 6, 1.5))
    rng = np.random.default_rng(4)
    for i in range(64):
        y = float(rng.uniform(-15, 15))
        if i % 2:
            acc.add(0, 0.0, y, vx=900.0 * (1 if i % 4 == 1 else -1), mass=1.0, p0=0.2)
        else:
            acc.add(2, 0.0, y, rot=float(rng.uniform(0, 6)), vx=-900.0 if i % 4 == 0 else 900.0, mass=1.0,
                    verts=scenes.polygon_outline(rng, 5, 0.3))
    scene = {"materials": [(0.0, 0.6, 0.6)], "bodies": acc.bodies(), "settings": {"gravity_y": 0.0}}
    w = make_world(agx, scene)
    w.step(DT, 1)
    ow = oracle_from_gpu(oracle, scene, w, dict(gy=0.0))
    ow.snapshot()
    for p in (agx.PHASE_INTEGRATE, agx.PHASE_POSE):
        w.run_phase(p, DT)
    ow.integrate(DT)
    ow.wrap_cossin()
    ow.poses()
    ow.aabbs()
    w.run_phase(agx.PHASE_CCD, DT)
    hit, t = ow.ccd()
    assert hit
    ow.aabbs()          # Simulation.cpp:311-315 rebuilds dynamic AABBs after a CCD hit
    np.testing.assert_allclose(w.body("com_x"), ow.get("comX"), rtol=0, atol=1e-4)
    np.testing.assert_allclose(w.collider("aabb_max_x"), ow.get("maxX"), rtol=0, atol=1e-4)
    for p in range(agx.PHASE_BROAD, agx.PHASE_FINISH + 1):
        w.run_phase(p, DT)
    assert w.stats().ccd_hits == int(np.count_nonzero(t < 1.0))


def test_batched_scenes_match_single_scenes(agx, oracle, scenes):
    # S independent scenes in one world give each scene's own pairs and manifolds.
    parts = [scenes.pile(120 + 40 * s, 30, 22.0, 22.0, seed=20 + s) for s in range(5)]
    b = agx.Builder(len(parts))
    for s, sc in enumerate(parts):
        agx.build_scene(b, sc, s)
    w = agx.World(b, agx.default_settings(velocity_iterations=8))
    w.step(DT, 60)
    for p in range(agx.PHASE_NARROW + 1):
        w.run_phase(p, DT)
    ga, gb = w.pairs()
    scene_of = w.collider("scene")
    assert np.all(scene_of[ga] == scene_of[gb])
    mg = w.manifolds(0)
    c_off = 0
    for s, sc in enumerate(parts):
        nc = int(np.count_nonzero(scene_of == s))
        sel = (ga >= c_off) & (ga < c_off + nc)
        ow = oracle.world_from_scene(sc)
        for g, o in (("world_x", "wX"), ("world_y", "wY"), ("cos", "wCos"), ("sin", "wSin"),
                     ("aabb_min_x", "minX"), ("aabb_min_y", "minY"), ("aabb_max_x", "maxX"), ("aabb_max_y", "maxY")):
            ow.set(o, w.collider(g)[c_off:c_off + nc])
        oa, ob = ow.broad_brute()
        is_dyn = np.zeros(nc, bool)
        is_dyn[ow.counts()["static_colliders"]:] = True
        assert pair_set(ga[sel] - c_off, gb[sel] - c_off, is_dyn) == pair_set(oa, ob, is_dyn)
        ms = mg[(mg["a"] >= c_off) & (mg["a"] < c_off + nc)].copy()
        mo = ow.narrow(ga[sel] - c_off, gb[sel] - c_off)
        ms["a"] -= c_off
        ms["b"] -= c_off
        ms, mo = ms[sort_key(ms)], mo[sort_key(mo)]
        assert len(ms) == len(mo)
        for f in ("jn0", "jt0", "jn1", "jt1"):   # warm-start impulses are checked elsewhere
            ms[f] = mo[f]
        assert_manifolds_equal(ms, mo, f"scene {s}")
        c_off += nc
    for p in range(agx.PHASE_SOLVE, agx.PHASE_FINISH + 1):
        w.run_phase(p, DT)
    assert w.stats().overflow == 0


def test_batch_scenes_match_oracle_scene_by_scene(agx, oracle, scenes):
    # A batched world (pyramids, piles, funnels) against the oracle stepping each scene alone.
    parts = [scenes.batch_scene(i, seed=4, n_min=64, n_max=400) for i in range(6)]
    b = agx.Builder(len(parts))
    for s, sc in enumerate(parts):
        agx.build_scene(b, sc, s)
    w = agx.World(b, agx.default_settings(velocity_iterations=8, position_iterations=3))
    w.step(DT, 90)
    cscene = w.collider("scene")
    bscene = np.empty(w.counts.bodies, np.uint32)
    bscene[w.collider("body")] = cscene          # one collider per body here
    bodies0 = {f: w.body(f) for f in ("com_x", "com_y", "rotation", "vel_x", "vel_y", "ang_vel", "cos", "sin")}
    cols0 = {f: w.collider(f) for f in ("world_x", "world_y", "cos", "sin", "world_rot", "aabb_min_x", "aabb_min_y",
                                         "aabb_max_x", "aabb_max_y")}
    prev = w.manifolds(1)
    for p in range(agx.PHASE_NARROW + 1):
        w.run_phase(p, DT)
    ga, gb = w.pairs()
    mg = w.manifolds(0)
    b1 = {f: w.body(f) for f in ("com_x", "com_y", "rotation", "cos")}
    aabb1 = w.collider("aabb_min_x")
    w.run_phase(agx.PHASE_SOLVE, DT)
    order, _ = w.solve_order()
    solved = w.manifolds(0)
    b2 = {f: w.body(f) for f in ("vel_x", "vel_y", "ang_vel", "com_x", "com_y")}
    pos_of = np.empty(len(order), np.int64)
    pos_of[order] = np.arange(len(order))
    renames = {"com_x": "comX", "com_y": "comY", "rotation": "rot", "vel_x": "velX", "vel_y": "velY",
               "ang_vel": "angVel", "cos": "cos", "sin": "sin"}
    two_point = 0
    col_names = {"world_x": "wX", "world_y": "wY", "cos": "wCos", "sin": "wSin", "world_rot": "wRot",
                 "aabb_min_x": "minX", "aabb_min_y": "minY", "aabb_max_x": "maxX", "aabb_max_y": "maxY"}
    for s, sc in enumerate(parts):
        bs = np.flatnonzero(bscene == s)
        cs = np.flatnonzero(cscene == s)
        b_off, c_off = int(bs[0]), int(cs[0])
        ow = oracle.world_from_scene(sc, dict(v=8, p=3))
        for f, o in renames.items():
            ow.set(o, bodies0[f][bs])
        for f, o in col_names.items():
            ow.set(o, cols0[f][cs])
        pm = prev[(prev["a"] >= c_off) & (prev["a"] < c_off + len(cs))].copy()
        pm["a"] -= c_off
        pm["b"] -= c_off
        ow.set_prev_manifolds(pm)
        ow.snapshot()
        ow.integrate(DT)
        ow.wrap_cossin()
        ow.poses()
        ow.aabbs()
        for f in ("com_x", "com_y", "rotation", "cos"):
            assert np.array_equal(b1[f][bs], ow.get(renames[f])), (s, f)
        assert np.array_equal(aabb1[cs], ow.get("minX")), s
        sel = (ga >= c_off) & (ga < c_off + len(cs))
        pa, pb = ga[sel] - c_off, gb[sel] - c_off
        oa, ob = ow.broad_brute()
        is_dyn = np.zeros(len(cs), bool)
        is_dyn[ow.counts()["static_colliders"]:] = True
        assert pair_set(pa, pb, is_dyn) == pair_set(oa, ob, is_dyn), s
        msel = np.flatnonzero((mg["a"] >= c_off) & (mg["a"] < c_off + len(cs)))
        ms = mg[msel].copy()
        ms["a"] -= c_off
        ms["b"] -= c_off
        mo = ow.warm_start(ow.narrow(pa, pb))
        assert len(ms) == len(mo) > 0, s
        two_point += assert_manifolds_equal(ms, mo, f"scene {s}")
        # the GPU's solve order restricted to this scene
        local_order = np.argsort(pos_of[msel], kind="stable").astype(np.uint32)
        so = ow.solve(ms, order=local_order, v_it=8, p_it=3)
        sg = solved[msel]
        two = sg["count"] == 2
        for f in ("jn0", "jt0", "jn1", "jt1"):
            sel = two if f in ("jn1", "jt1") else np.ones(len(sg), bool)
            np.testing.assert_allclose(sg[f][sel], so[f][sel], rtol=1e-4, atol=1e-4, err_msg=f"{s} {f}")
        for f in ("vel_x", "vel_y", "ang_vel", "com_x", "com_y"):
            np.testing.assert_allclose(b2[f][bs], ow.get(renames[f]), rtol=1e-4, atol=1e-4, err_msg=f"{s} {f}")
    assert two_point > 50   # box stacks: both points of box-box manifolds checked
    for p in range(agx.PHASE_ISLANDS, agx.PHASE_FINISH + 1):
        w.run_phase(p, DT)
    assert w.stats().overflow == 0


def test_scatter_pairs_match_oracle_at_scale(agx, oracle, scenes):
    # A quarter-size scatter (262,144 dynamic colliders, bullets, 4,096 statics): the GPU pair set
    # after CCD equals the oracle's exact sweep-and-prune set on the same AABBs.
    scene = scenes.scatter(1 << 18, 2048.0, seed=3, n_static=4096)
    w = make_world(agx, scene)
    w.step(DT, 2)
    for p in range(agx.PHASE_BROAD + 1):
        w.run_phase(p, DT)
    ga, gb = w.pairs()
    ow = oracle.world_from_scene(scene)
    for g, o in (("aabb_min_x", "minX"), ("aabb_min_y", "minY"), ("aabb_max_x", "maxX"), ("aabb_max_y", "maxY")):
        ow.set(o, w.collider(g))
    oa, ob = ow.broad_brute()
    ns = ow.counts()["static_colliders"]

    def keys(a, b):
        a = a.astype(np.uint64)
        b = b.astype(np.uint64)
        both_dyn = (a >= ns) & (b >= ns)
        lo = np.where(both_dyn, np.minimum(a, b), a)
        hi = np.where(both_dyn, np.maximum(a, b), b)
        return np.sort((lo << np.uint64(32)) | hi)

    kg, ko = keys(ga, gb), keys(oa, ob)
    assert len(kg) > 1000 and len(np.unique(kg)) == len(kg)
    assert np.array_equal(kg, ko)
    for p in range(agx.PHASE_NARROW, agx.PHASE_FINISH + 1):
        w.run_phase(p, DT)
    st = w.stats()
    assert st.overflow == 0 and st.manifolds > 0
