"""Pins the oracle (CPU restatement) against mathematics the reference's behaviour fixes.

The reference ships no tests, golden vectors or fixtures and cannot be built here, so each
oracle function is checked against closed forms, brute force or invariants chosen so that a
plausible slip (dropped term, wrong sign, swapped operand, off-by-one index) fails a check.
"""
import math

import numpy as np
import pytest

DT = 1.0 / 60.0


def scene_of(oracle, bodies, materials=((0.0, 0.6, 0.6),), settings=None):
    w = oracle.World(settings)
    for m in materials:
        w.add_material(*m)
    from agx.scenes import _Acc  # noqa: F401
    w.add_bodies(bodies)
    w.refresh()
    return w


def acc_bodies(scenes, items):
    acc = scenes._Acc()
    for it in items:
        acc.add(**it)
    return acc.bodies()


# ---------------------------------------------------------------- FastCosSin / wrap
def test_fast_cossin_error_bound(oracle):
    x = np.linspace(0.0, 2 * np.pi, 200001, endpoint=False).astype(np.float32)
    c, s = oracle.fast_cossin(x)
    xd = x.astype(np.float64)
    assert np.max(np.abs(s - np.sin(xd))) < 2e-6
    assert np.max(np.abs(c - np.cos(xd))) < 2e-3   # cos = sqrt(1 - sin^2) loses accuracy near sin = 1
    # quadrant signs: exact at the quadrant interiors
    q = (x / (np.pi / 2)).astype(int)
    mid = np.abs(x - (q + 0.5) * np.pi / 2) < 0.5
    assert np.all(np.sign(s[mid]) == np.sign(np.sin(xd[mid])))
    assert np.all(np.sign(c[mid]) == np.sign(np.cos(xd[mid])))
    assert np.all(s <= 1.0) and np.all(c >= -1.0)


def test_wrap_rotation(oracle):
    x = np.array([0.0, 1.0, 6.5, -0.5, -7.0, 100.0, 2 * np.pi], np.float32)
    y = oracle.wrap(x)
    assert np.all(y >= 0.0) and np.all(y < np.float32(2 * np.pi) + 1e-6)
    expect = np.mod(x.astype(np.float64), 2 * np.pi)
    d = np.abs(y - expect)
    d = np.minimum(d, np.abs(d - 2 * np.pi))
    assert np.max(d) < 1e-5
    assert y[1] == np.float32(1.0)   # already wrapped: untouched


# ---------------------------------------------------------------- builder closed forms
def test_builder_mass_properties(oracle, scenes):
    sq = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], np.float32) * 0.5
    b = acc_bodies(scenes, [
        dict(t=0, px=0.0, py=5.0, mass=2.0, p0=0.5),
        dict(t=1, px=3.0, py=5.0, mass=3.0, p0=2.0, p1=1.0),
        dict(t=2, px=6.0, py=5.0, mass=1.0, verts=sq + np.float32(0.25)),
    ])
    w = scene_of(oracle, b)
    inertia = w.get("inertia")
    assert inertia[0] == pytest.approx(0.5 * 0.25 * 2.0, rel=1e-6)          # circle: m r^2 / 2
    assert inertia[1] == pytest.approx(3.0 * (4.0 + 1.0) / 12.0, rel=1e-6)  # box: m (w^2 + h^2) / 12
    assert inertia[2] == pytest.approx(1.0 * (1.0 + 1.0) / 12.0, rel=1e-5)  # unit square polygon
    # polygon body sits at its centroid: position + R * (0.25, 0.25)
    assert w.get("comX")[2] == pytest.approx(6.25, abs=1e-6)
    assert w.get("comY")[2] == pytest.approx(5.25, abs=1e-6)
    v, n = w.vertices()
    assert np.allclose(v.mean(axis=0), 0.0, atol=1e-6)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-6)
    # outward normals: n_i . (v_i - centroid) > 0
    assert np.all(np.sum(n * v, axis=1) > 0)
    assert w.get("minExtent")[2] == pytest.approx(0.5, abs=1e-6)           # in-radius of the unit square
    assert w.get("maxExtent")[2] == pytest.approx(math.sqrt(0.5), abs=1e-6)


def test_builder_group_order(oracle, scenes):
    # dynamic, dynamic, static: the static body takes the first dynamic's place, which moves last.
    b = acc_bodies(scenes, [
        dict(t=0, px=0.0, py=0.0, mass=1.0, p0=0.1),
        dict(t=0, px=1.0, py=0.0, mass=2.0, p0=0.2),
        dict(t=1, px=2.0, py=0.0, mass=0.0, p0=1.0, p1=1.0),
    ])
    w = scene_of(oracle, b)
    assert w.counts()["static_bodies"] == 1
    assert list(w.get("comX")) == [2.0, 1.0, 0.0]
    assert list(w.get("colBody")) == [0, 1, 2]
    assert list(w.get("bodySlot")) == [2, 1, 0]
    assert list(w.get("colSlot")) == [2, 1, 0]


def test_builder_rejects_bad_polygons(oracle, scenes):
    concave = np.array([[0, 0], [2, 0], [1, 0.2], [2, 2], [0, 2]], np.float32)
    line = np.array([[0, 0], [1, 0], [2, 0]], np.float32)
    b = acc_bodies(scenes, [dict(t=2, px=0.0, py=0.0, mass=1.0, verts=concave),
                            dict(t=2, px=0.0, py=0.0, mass=1.0, verts=line)])
    w = oracle.World()
    idx, rc = w.add_bodies(b)
    assert rc == 2 and np.all(idx == 0xFFFFFFFF)


# ---------------------------------------------------------------- integrator
def test_integrator_closed_form(oracle, scenes):
    b = acc_bodies(scenes, [dict(t=0, px=0.0, py=10.0, vx=1.0, vy=2.0, av=0.5, mass=1.0, p0=0.5),
                            dict(t=1, px=0.0, py=0.0, vx=3.0, mass=0.0, p0=1.0, p1=1.0)])
    w = scene_of(oracle, b)
    w.integrate(DT)
    # velocity Verlet under constant gravity: x += v dt + g dt^2 / 2, v += g dt (dynamic only)
    assert w.get("comY")[1] == pytest.approx(10.0 + 2.0 * DT - 0.5 * 9.81 * DT * DT, abs=1e-5)
    assert w.get("velY")[1] == pytest.approx(2.0 - 9.81 * DT, abs=1e-5)
    assert w.get("rot")[1] == pytest.approx(0.5 * DT, abs=1e-7)
    assert w.get("comX")[0] == 0.0 and w.get("velX")[0] == 3.0   # static bodies are not integrated


# ---------------------------------------------------------------- broad phase
def random_aabb_world(oracle, scenes, seed, n=600, n_static=40, size=30.0):
    rng = np.random.default_rng(seed)
    items = []
    for i in range(n_static):
        items.append(dict(t=1, px=float(rng.uniform(-size, size)), py=float(rng.uniform(-size, size)),
                          rot=float(rng.uniform(0, 6.28)), mass=0.0, p0=float(rng.uniform(1, 4)),
                          p1=float(rng.uniform(1, 4))))
    for i in range(n):
        t = int(rng.integers(0, 3))
        r = float(rng.uniform(0.2, 1.5))
        kw = dict(px=float(rng.uniform(-size, size)), py=float(rng.uniform(-size, size)),
                  rot=float(rng.uniform(0, 6.28)), mass=1.0)
        if t == 0:
            items.append(dict(t=0, p0=r, **kw))
        elif t == 1:
            items.append(dict(t=1, p0=r, p1=float(rng.uniform(0.2, 1.5)), **kw))
        else:
            items.append(dict(t=2, verts=scenes.polygon_outline(rng, int(rng.integers(3, 9)), r), **kw))
    return scene_of(oracle, acc_bodies(scenes, items))


def pair_set(a, b, n_static):
    out = set()
    for x, y in zip(a.tolist(), b.tolist()):
        out.add((min(x, y), max(x, y)) if x >= n_static and y >= n_static else (x, y))
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bvh_pairs_equal_brute_force(oracle, scenes, seed):
    w = random_aabb_world(oracle, scenes, seed)
    ns = w.counts()["static_colliders"]
    a, b = w.broad_bvh()
    ba, bb = w.broad_brute()
    s_bvh, s_brute = pair_set(a, b, ns), pair_set(ba, bb, ns)
    assert len(a) == len(s_bvh), "duplicate pairs"
    assert s_bvh == s_brute
    assert len(s_brute) > 100
    # pure O(n^2) definition on a subset confirms the sweep-and-prune
    mnx, mny, mxx, mxy = (w.get(k) for k in ("minX", "minY", "maxX", "maxY"))
    body = w.get("colBody")
    n = len(mnx)
    ref = set()
    for i in range(n):
        for j in range(i + 1, n):
            if not (i >= ns or j >= ns):
                continue
            if body[i] == body[j]:
                continue
            if mnx[i] < mxx[j] and mxx[i] > mnx[j] and mny[i] < mxy[j] and mxy[i] > mny[j]:
                ref.add((i, j) if i >= ns else (j, i))
    assert s_brute == ref


# ---------------------------------------------------------------- narrow phase closed forms
def two_body(oracle, scenes, a, b, materials=((0.0, 0.6, 0.6),)):
    w = scene_of(oracle, acc_bodies(scenes, [a, b]), materials)
    return w


def test_circle_circle(oracle, scenes):
    w = two_body(oracle, scenes, dict(t=0, px=0.0, py=0.0, mass=1.0, p0=1.0), dict(t=0, px=1.5, py=0.0, mass=1.0, p0=1.0))
    m = w.narrow([0], [1])
    assert len(m) == 1
    # same type: the higher collider index is A, so A = 1 (at x=1.5), normal points to B (-x)
    assert (m["a"][0], m["b"][0]) == (1, 0)
    assert m["depth"][0] == pytest.approx(0.5)
    assert (m["nx"][0], m["ny"][0]) == (-1.0, 0.0)
    assert (m["p0x"][0], m["p0y"][0]) == (0.5, 0.0)
    w2 = two_body(oracle, scenes, dict(t=0, px=0.0, py=0.0, mass=1.0, p0=1.0), dict(t=0, px=0.0, py=0.0, mass=1.0, p0=1.0))
    m2 = w2.narrow([0], [1])
    assert (m2["nx"][0], m2["ny"][0]) == (1.0, 0.0) and m2["depth"][0] == 2.0   # coincident centres


def test_box_box_axis_aligned(oracle, scenes):
    w = two_body(oracle, scenes, dict(t=1, px=0.0, py=0.0, mass=0.0, p0=4.0, p1=2.0),
                 dict(t=1, px=0.5, py=1.9, mass=1.0, p0=1.0, p1=2.0))
    m = w.narrow([1], [0])
    assert len(m) == 1 and m["count"][0] == 2
    # A = higher index = 1 (the dynamic box on top): normal points down to the floor box
    assert m["a"][0] == 1
    assert m["depth"][0] == pytest.approx(0.1, abs=1e-6)
    assert m["nx"][0] == pytest.approx(0.0, abs=1e-7) and m["ny"][0] == pytest.approx(-1.0)
    pts = sorted([(m["p0x"][0], m["p0y"][0]), (m["p1x"][0], m["p1y"][0])])
    assert pts[0] == pytest.approx((0.0, 1.0), abs=1e-6) and pts[1] == pytest.approx((1.0, 1.0), abs=1e-6)


def test_polygon_polygon_matches_box_box(oracle, scenes):
    sq = lambda h: np.array([[-h, -h], [h, -h], [h, h], [-h, h]], np.float32)
    boxes = two_body(oracle, scenes, dict(t=1, px=0.0, py=0.0, rot=0.3, mass=1.0, p0=2.0, p1=2.0),
                     dict(t=1, px=1.2, py=1.0, rot=0.7, mass=1.0, p0=2.0, p1=2.0))
    polys = two_body(oracle, scenes, dict(t=2, px=0.0, py=0.0, rot=0.3, mass=1.0, verts=sq(1.0)),
                     dict(t=2, px=1.2, py=1.0, rot=0.7, mass=1.0, verts=sq(1.0)))
    mb, mp = boxes.narrow([0], [1]), polys.narrow([0], [1])
    assert len(mb) == 1 and len(mp) == 1
    assert mp["depth"][0] == pytest.approx(mb["depth"][0], abs=2e-6)
    assert mp["nx"][0] == pytest.approx(mb["nx"][0], abs=2e-6) and mp["ny"][0] == pytest.approx(mb["ny"][0], abs=2e-6)
    pb = sorted([(mb["p0x"][0], mb["p0y"][0]), (mb["p1x"][0], mb["p1y"][0])][:mb["count"][0]])
    pp = sorted([(mp["p0x"][0], mp["p0y"][0]), (mp["p1x"][0], mp["p1y"][0])][:mp["count"][0]])
    assert np.allclose(pb, pp, atol=2e-5)
    # box-polygon on the same geometry agrees too
    mixed = two_body(oracle, scenes, dict(t=1, px=0.0, py=0.0, rot=0.3, mass=1.0, p0=2.0, p1=2.0),
                     dict(t=2, px=1.2, py=1.0, rot=0.7, mass=1.0, verts=sq(1.0)))
    mx = mixed.narrow([0], [1])
    assert mx["depth"][0] == pytest.approx(mb["depth"][0], abs=2e-6)


def _convex_polygon(rng):
    """3-8 vertices on an ellipse, counter-clockwise, no sliver edges."""
    n = int(rng.integers(3, 9))
    while True:
        ang = np.sort(rng.uniform(0.0, 2 * np.pi, n))
        gaps = np.diff(np.r_[ang, ang[0] + 2 * np.pi])
        if gaps.min() > 0.35 and gaps.max() < np.pi - 0.2:
            break
    ax, ay = rng.uniform(0.6, 1.4, 2)
    return np.stack([ax * np.cos(ang), ay * np.sin(ang)], 1).astype(np.float32)


def _world_outlines(w):
    """World vertices (counter-clockwise) of colliders 0 and 1, boxes from their half extents."""
    v, _ = w.vertices()
    start, count, typ = w.get("polyStart"), w.get("polyCount"), w.types()
    wx, wy, c, s = w.get("wX"), w.get("wY"), w.get("wCos"), w.get("wSin")
    hx, hy = w.get("p0"), w.get("p1")
    out = []
    for k in range(2):
        if typ[k] == 2:
            local = v[start[k]:start[k] + count[k]].astype(np.float64)
        else:
            local = np.array([[-hx[k], -hy[k]], [hx[k], -hy[k]], [hx[k], hy[k]], [-hx[k], hy[k]]], np.float64)
        rot = np.array([[c[k], -s[k]], [s[k], c[k]]], np.float64)
        out.append(local @ rot.T + [wx[k], wy[k]])
    return out


def _best_face(p, q):
    """Separating-axis brute force: the face of p whose outward normal separates q the most."""
    best, normal = -np.inf, None
    for i in range(len(p)):
        e = p[(i + 1) % len(p)] - p[i]
        n = np.array([e[1], -e[0]]) / np.linalg.norm(e)
        sep = float(np.min((q - p[i]) @ n))
        if sep > best:
            best, normal = sep, n
    return best, normal


def test_polygon_faces_match_separating_axis_brute_force(oracle, scenes):
    """Box-box, box-polygon and polygon-polygon in general position (the two bodies' best face
    separations at least 0.05 apart, so no tie-break decides): the manifold normal is the outward
    normal of the face with the larger separation, oriented from A to B, and the depth is minus
    that separation (BoxBox.cpp:13-228, BoxPolygon.cpp:96-172, PolygonPolygon.cpp:93-147)."""
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(800):
        kinds = rng.integers(1, 3, 2)
        bodies = []
        for k in range(2):
            d = dict(px=float(k * rng.uniform(0.8, 1.8)), py=float(k * rng.uniform(-0.8, 0.8)),
                     rot=float(rng.uniform(0.0, 6.28)), mass=1.0)
            if kinds[k] == 2:
                d.update(t=2, verts=_convex_polygon(rng))
            else:
                d.update(t=1, p0=float(rng.uniform(1.0, 2.5)), p1=float(rng.uniform(1.0, 2.5)))
            bodies.append(d)
        w = two_body(oracle, scenes, *bodies)
        w.poses(True)
        m = w.narrow([0], [1])
        p0, p1 = _world_outlines(w)
        s0, n0 = _best_face(p0, p1)
        s1, n1 = _best_face(p1, p0)
        if max(s0, s1) > -0.02 or abs(s0 - s1) < 0.05:
            continue
        assert len(m) == 1
        normal = n0 if s0 > s1 else -n1          # from collider 0 to collider 1
        if m["a"][0] == 1:
            normal = -normal
        assert np.allclose([m["nx"][0], m["ny"][0]], normal, atol=2e-5)
        assert m["depth"][0] == pytest.approx(-max(s0, s1), abs=2e-5)
        checked += 1
    assert checked > 300


def test_circle_box_and_polygon_regions(oracle, scenes):
    sq = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], np.float32)
    for shape in (dict(t=1, p0=2.0, p1=2.0), dict(t=2, verts=sq)):
        # circle above the top face
        w = two_body(oracle, scenes, dict(t=0, px=0.3, py=1.4, mass=1.0, p0=0.5), dict(px=0.0, py=0.0, mass=0.0, **shape))
        m = w.narrow([1], [0])
        assert len(m) == 1
        assert m["depth"][0] == pytest.approx(0.1, abs=1e-6)
        assert m["nx"][0] == pytest.approx(0.0, abs=1e-6) and m["ny"][0] == pytest.approx(-1.0, abs=1e-6)
        assert (m["p0x"][0], m["p0y"][0]) == pytest.approx((0.3, 0.9), abs=1e-6)
        # circle centre inside, nearest face +x
        w = two_body(oracle, scenes, dict(t=0, px=0.8, py=0.1, mass=1.0, p0=0.5), dict(px=0.0, py=0.0, mass=0.0, **shape))
        m = w.narrow([1], [0])
        assert m["depth"][0] == pytest.approx(0.5 + 0.2, abs=1e-6)
        assert m["nx"][0] == pytest.approx(-1.0, abs=1e-6)
        # separated: no manifold
        w = two_body(oracle, scenes, dict(t=0, px=2.0, py=2.0, mass=1.0, p0=0.3), dict(px=0.0, py=0.0, mass=0.0, **shape))
        assert len(w.narrow([1], [0])) == 0
    # circle-box outside-corner region id: right | top = 1 | 4
    w = two_body(oracle, scenes, dict(t=0, px=1.2, py=1.2, mass=1.0, p0=0.5), dict(t=1, px=0.0, py=0.0, mass=0.0, p0=2.0, p1=2.0))
    m = w.narrow([1], [0])
    assert m["id0"][0] == 5
    d = math.sqrt(0.08)
    assert m["depth"][0] == pytest.approx(0.5 - d, abs=1e-6)


def test_manifold_invariants_random(oracle, scenes):
    w = random_aabb_world(oracle, scenes, 7, n=400, n_static=20, size=15.0)
    a, b = w.broad_brute()
    m = w.narrow(a, b)
    assert len(m) > 50
    n = np.stack([m["nx"], m["ny"]], 1)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)
    assert np.all(m["depth"] > -1e-3)
    assert np.all((m["count"] >= 1) & (m["count"] <= 2))
    # normal points from A to B for circle-circle manifolds
    wx, wy, typ = w.get("wX"), w.get("wY"), w.types()
    cc = (typ[m["a"]] == 0) & (typ[m["b"]] == 0)
    d = np.stack([wx[m["b"][cc]] - wx[m["a"][cc]], wy[m["b"][cc]] - wy[m["a"][cc]]], 1)
    assert np.all(np.sum(d * n[cc], axis=1) >= 0)
    # canonical A/B: lower type first, same type -> higher index is A
    ta, tb = typ[m["a"]], typ[m["b"]]
    assert np.all(ta <= tb)
    same = ta == tb
    assert np.all(m["a"][same] > m["b"][same])


def test_warm_start_matching(oracle, scenes):
    w = two_body(oracle, scenes, dict(t=1, px=0.0, py=0.0, mass=0.0, p0=4.0, p1=2.0),
                 dict(t=1, px=0.5, py=1.9, mass=1.0, p0=1.0, p1=2.0))
    m = w.narrow([1], [0])
    prev = m.copy()
    prev["jn0"], prev["jt0"], prev["jn1"], prev["jt1"] = 1.0, 2.0, 3.0, 4.0
    prev["id0"], prev["id1"] = m["id1"][0], m["id0"][0]          # ids swapped: impulses follow the ids
    w.set_prev_manifolds(prev)
    cur = w.warm_start(m)
    assert (cur["jn0"][0], cur["jt0"][0], cur["jn1"][0], cur["jt1"][0]) == (3.0, 4.0, 1.0, 2.0)
    prev2 = prev.copy()
    prev2["a"], prev2["b"] = prev["b"], prev["a"]                 # A/B order differs: no copy
    w.set_prev_manifolds(prev2)
    cur = w.warm_start(m)
    assert cur["jn0"][0] == 0.0 and cur["jn1"][0] == 0.0
    prev3 = prev.copy()
    prev3["jn0"] = prev3["jt0"] = prev3["jn1"] = prev3["jt1"] = 0.0  # all-zero impulses are not cached
    w.set_prev_manifolds(prev3)
    assert w.warm_start(m)["jn0"][0] == 0.0


# ---------------------------------------------------------------- solver
def test_head_on_circles_closed_form(oracle, scenes):
    # Two equal circles approaching at +-v, e = 0, no friction: one normal impulse m v stops both.
    mat = ((0.0, 0.0, 0.0),)
    w = two_body(oracle, scenes, dict(t=0, px=0.0, py=0.0, vx=2.0, mass=1.0, p0=0.5),
                 dict(t=0, px=0.99, py=0.0, vx=-2.0, mass=1.0, p0=0.5), mat)
    m = w.narrow([0], [1])
    out = w.solve(m, v_it=8, p_it=0)
    assert out["jn0"][0] == pytest.approx(2.0, rel=1e-6)     # effective mass 1/2, closing speed 4
    assert abs(out["jt0"][0]) < 1e-7
    assert np.allclose(w.get("velX"), 0.0, atol=1e-6)
    # elastic: velocities exchange
    w = two_body(oracle, scenes, dict(t=0, px=0.0, py=0.0, vx=2.0, mass=1.0, p0=0.5),
                 dict(t=0, px=0.99, py=0.0, vx=-2.0, mass=1.0, p0=0.5), ((1.0, 0.0, 0.0),))
    out = w.solve(w.narrow([0], [1]), v_it=8, p_it=0)
    assert w.get("velX") == pytest.approx([-2.0, 2.0], abs=1e-5)


def test_momentum_conserved_without_statics(oracle, scenes):
    rng = np.random.default_rng(3)
    items = [dict(t=int(rng.integers(0, 2)), px=float(rng.uniform(-3, 3)), py=float(rng.uniform(-3, 3)),
                  rot=float(rng.uniform(0, 6)), vx=float(rng.normal()), vy=float(rng.normal()),
                  av=float(rng.normal()), mass=float(rng.uniform(0.5, 2)), p0=0.6, p1=0.5) for _ in range(60)]
    w = scene_of(oracle, acc_bodies(scenes, items))
    a, b = w.broad_brute()
    m = w.narrow(a, b)
    assert len(m) > 10
    mass, vx, vy = w.get("mass"), w.get("velX"), w.get("velY")
    p0 = np.array([np.sum(mass * vx, dtype=np.float64), np.sum(mass * vy, dtype=np.float64)])
    w.solve(m, v_it=8, p_it=3)
    vx, vy = w.get("velX"), w.get("velY")
    p1 = np.array([np.sum(mass * vx, dtype=np.float64), np.sum(mass * vy, dtype=np.float64)])
    assert np.allclose(p0, p1, atol=1e-3 * (1 + np.abs(p0).max()))


def test_solver_order_invariance_within_colour(oracle, scenes):
    # Constraints sharing no dynamic body commute: reversing their order changes nothing.
    items = [dict(t=1, px=0.0, py=-1.0, mass=0.0, p0=40.0, p1=2.0)]
    items += [dict(t=0, px=float(x), py=0.45, vy=-1.0, mass=1.0, p0=0.5) for x in range(-10, 11, 2)]
    w1 = scene_of(oracle, acc_bodies(scenes, items))
    w2 = scene_of(oracle, acc_bodies(scenes, items))
    a, b = w1.broad_brute()
    m = w1.narrow(a, b)
    assert len(m) == 11
    o = np.arange(len(m), dtype=np.uint32)
    r1 = w1.solve(m, order=o, v_it=4, p_it=2)
    r2 = w2.solve(m, order=o[::-1].copy(), v_it=4, p_it=2)
    for f in ("jn0", "jt0"):
        assert np.array_equal(r1[f], r2[f])
    assert np.array_equal(w1.get("velY"), w2.get("velY"))


def brute_block_lcp(K, b):
    best = None
    for active in ((1, 1), (0, 1), (1, 0), (0, 0)):
        x = np.zeros(2)
        idx = [i for i in range(2) if active[i]]
        if idx:
            sub = K[np.ix_(idx, idx)]
            x[idx] = np.linalg.solve(sub, b[idx])
        wv = K @ x - b
        if np.all(x >= -1e-9) and np.all(wv >= -1e-9):
            best = x
            break
    return best


def test_block_lcp_matches_enumeration(oracle, scenes):
    # A box resting on the floor with two contacts: one velocity sweep of the block solver gives
    # the complementarity solution of the 2x2 LCP (w = K x - b', x >= 0, w >= 0, x.w = 0).
    w = two_body(oracle, scenes, dict(t=1, px=0.0, py=-1.0, mass=0.0, p0=10.0, p1=2.0),
                 dict(t=1, px=0.2, py=0.49, rot=0.0, vy=-1.0, av=0.8, mass=2.0, p0=1.0, p1=1.0), ((0.0, 0.0, 0.0),))
    m = w.narrow([1], [0])
    assert m["count"][0] == 2
    inv_m, inv_i = w.get("invMass")[1], w.get("invInertia")[1]
    vx, vy, av = w.get("velX")[1], w.get("velY")[1], w.get("angVel")[1]
    cx, cy = w.get("comX")[1], w.get("comY")[1]
    n = np.array([m["nx"][0], m["ny"][0]], np.float64)   # from the box (A) to the floor (B)
    rn = []
    for k in range(2):
        p = np.array([m[f"p{k}x"][0], m[f"p{k}y"][0]]) - np.array([cx, cy])
        rn.append(-p[1] * n[0] + p[0] * n[1])
    K = np.array([[inv_m + rn[0] * rn[0] * inv_i, inv_m + rn[0] * rn[1] * inv_i],
                  [inv_m + rn[0] * rn[1] * inv_i, inv_m + rn[1] * rn[1] * inv_i]])
    vn = np.array([-(vx * n[0] + vy * n[1]) + (-rn[k] * av) for k in range(2)])   # relative normal velocity B - A
    bprime = -vn
    x = brute_block_lcp(K, bprime)
    out = w.solve(m, v_it=1, p_it=0)
    assert out["jn0"][0] == pytest.approx(x[0], rel=1e-4, abs=1e-6)
    assert out["jn1"][0] == pytest.approx(x[1], rel=1e-4, abs=1e-6)


def test_position_pass_closed_form(oracle, scenes):
    # One position iteration moves the pair apart by half the penetration beyond the slop.
    w = two_body(oracle, scenes, dict(t=0, px=0.0, py=0.0, mass=1.0, p0=1.0), dict(t=0, px=1.8, py=0.0, mass=3.0, p0=1.0),
                 ((0.0, 0.0, 0.0),))
    m = w.narrow([0], [1])
    x0 = w.get("comX").copy()
    w.solve(m, v_it=0, p_it=1)
    x1 = w.get("comX")
    pen = 0.2 - 0.001
    total = pen / (1.0 + 1.0 / 3.0) * 0.5
    # A (higher index, x=1.8, inv mass 1/3) moves +x, B (x=0, inv mass 1) moves -x
    assert x1[1] - x0[1] == pytest.approx(total / 3.0, rel=1e-5)
    assert x1[0] - x0[0] == pytest.approx(-total, rel=1e-5)


# ---------------------------------------------------------------- islands
def test_islands_match_bfs(oracle, scenes):
    w = random_aabb_world(oracle, scenes, 11, n=300, n_static=10, size=12.0)
    a, b = w.broad_brute()
    m = w.narrow(a, b)
    ids, k = w.islands(m)
    ns = w.counts()["static_bodies"]
    body = w.get("colBody")
    nb = w.counts()["bodies"]
    adj = [[] for _ in range(nb)]
    for x, y in zip(body[m["a"]], body[m["b"]]):
        if x >= ns and y >= ns:
            adj[x].append(y)
            adj[y].append(x)
    comp = [-1] * nb
    c = 0
    for s in range(ns, nb):
        if comp[s] >= 0:
            continue
        stack = [s]
        comp[s] = c
        while stack:
            u = stack.pop()
            for v in adj[u]:
                if comp[v] < 0:
                    comp[v] = c
                    stack.append(v)
        c += 1
    assert k == c
    assert np.all(ids[:ns] == 0xFFFFFFFF)
    canon = lambda lab: {frozenset(np.flatnonzero(lab == v).tolist()) for v in set(lab.tolist())}
    assert canon(ids[ns:]) == canon(np.array(comp[ns:]))


# ---------------------------------------------------------------- CCD
def test_ccd_circle_into_static_box(oracle, scenes):
    # A fast circle moving +x crosses a thin static wall within one step.
    b = acc_bodies(scenes, [dict(t=1, px=5.0, py=0.0, mass=0.0, p0=0.2, p1=10.0),
                            dict(t=0, px=0.0, py=0.0, vx=600.0, mass=1.0, p0=0.2)])
    w = scene_of(oracle, b, settings=dict(gy=0.0))
    w.snapshot()
    w.integrate(DT)
    w.wrap_cossin()
    w.poses()
    w.aabbs()
    hit, t = w.ccd()
    assert hit
    # core radius 0.1 reaches the wall face x = 4.9 at x = 4.8; travel 10 m per step
    assert t[1] == pytest.approx(0.48, abs=1e-5)
    assert w.get("comX")[1] == pytest.approx(4.8, abs=1e-4)


def _distance_to_convex(p, x):
    """Distance from point x to the convex outline p (counter-clockwise); 0 inside."""
    inside = True
    best = np.inf
    for i in range(len(p)):
        a, b = p[i], p[(i + 1) % len(p)]
        e = b - a
        if e[0] * (x[1] - a[1]) - e[1] * (x[0] - a[0]) < 0:
            inside = False
        t = np.clip((x - a) @ e / (e @ e), 0.0, 1.0)
        best = min(best, float(np.linalg.norm(x - (a + t * e))))
    return 0.0 if inside else best


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_ccd_circle_into_rotated_static_polygon_and_box(oracle, scenes, seed):
    """A fast circle crossing a rotated static polygon or box within one step stops where its core
    circle (half the radius, as in the circle-into-box case) first touches the outline; the time of
    impact is found here by dense sampling and bisection of the exact point-to-outline distance
    along the straight path (CCD/ContinuousCollisionDetector.cpp:129-220)."""
    rng = np.random.default_rng(seed)
    static = dict(px=5.0, py=0.0, mass=0.0, rot=float(rng.uniform(0.0, 6.28)))
    if seed % 2:
        static.update(t=1, p0=float(rng.uniform(0.6, 1.5)), p1=float(rng.uniform(2.0, 4.0)))
    else:
        static.update(t=2, verts=(_convex_polygon(rng) * 1.5).astype(np.float32))
    vy = float(rng.uniform(-60.0, 60.0))
    b = acc_bodies(scenes, [static, dict(t=0, px=0.0, py=float(rng.uniform(-0.5, 0.5)), vx=600.0, vy=vy,
                                         mass=1.0, p0=0.2)])
    w = scene_of(oracle, b, settings=dict(gy=0.0))
    w.snapshot()
    c0 = np.array([w.get("comX")[1], w.get("comY")[1]], np.float64)
    w.integrate(DT)
    w.wrap_cossin()
    w.poses()
    w.aabbs()
    c1 = np.array([w.get("comX")[1], w.get("comY")[1]], np.float64)
    outline = _world_outlines(w)[0]
    core = 0.1
    ts = np.linspace(0.0, 1.0, 4001)
    dist = np.array([_distance_to_convex(outline, c0 + t * (c1 - c0)) for t in ts])
    k = int(np.argmax(dist <= core))
    assert dist[k] <= core and k > 0          # the path reaches the outline within the step
    lo, hi = ts[k - 1], ts[k]
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if _distance_to_convex(outline, c0 + mid * (c1 - c0)) <= core:
            hi = mid
        else:
            lo = mid
    hit, t = w.ccd()
    assert hit
    assert t[1] == pytest.approx(hi, abs=2e-5)
    assert w.get("comX")[1] == pytest.approx(c0[0] + hi * (c1[0] - c0[0]), abs=2e-4)
    assert w.get("comY")[1] == pytest.approx(c0[1] + hi * (c1[1] - c0[1]), abs=2e-4)


@pytest.mark.parametrize("mover", ["box", "polygon"])
def test_ccd_moving_box_and_polygon_core(oracle, scenes, mover):
    """A fast box or polygon sweeps its core circle (half its in-radius about the collider origin,
    ContinuousCollisionDetector.cpp:353-372) along the origin's path; the in-radius is computed
    here from the outline and the time of impact by bisection against the rotated static box."""
    rng = np.random.default_rng(5)
    static = dict(t=1, px=6.0, py=0.0, mass=0.0, rot=0.4, p0=1.0, p1=4.0)
    if mover == "box":
        body = dict(t=1, p0=0.6, p1=0.3)
    else:
        body = dict(t=2, verts=(_convex_polygon(rng) * 0.4).astype(np.float32))
    body.update(px=0.0, py=0.2, rot=0.3, vx=600.0, vy=12.0, mass=1.0)
    w = scene_of(oracle, acc_bodies(scenes, [static, body]), settings=dict(gy=0.0))
    w.snapshot()
    w.poses(True)
    o0 = np.array([w.get("wX")[1], w.get("wY")[1]], np.float64)
    com0 = np.array([w.get("comX")[1], w.get("comY")[1]], np.float64)
    w.integrate(DT)
    w.wrap_cossin()
    w.poses()
    w.aabbs()
    o1 = np.array([w.get("wX")[1], w.get("wY")[1]], np.float64)
    com1 = np.array([w.get("comX")[1], w.get("comY")[1]], np.float64)
    # in-radius about the collider origin, from the local outline
    if mover == "box":
        inradius = min(w.get("p0")[1], w.get("p1")[1])
    else:
        v, _ = w.vertices()
        s0, c = w.get("polyStart")[1], w.get("polyCount")[1]
        loc = v[s0:s0 + c].astype(np.float64)
        edges = [(loc[i], loc[(i + 1) % c] - loc[i]) for i in range(c)]
        inradius = min(abs(e[0] * a[1] - e[1] * a[0]) / np.linalg.norm(e) for a, e in edges)
    assert w.get("minExtent")[1] == pytest.approx(inradius, rel=1e-5)
    core = 0.5 * inradius
    outline = _world_outlines(w)[0]
    ts = np.linspace(0.0, 1.0, 4001)
    dist = np.array([_distance_to_convex(outline, o0 + t * (o1 - o0)) for t in ts])
    k = int(np.argmax(dist <= core))
    assert dist[k] <= core and k > 0
    lo, hi = ts[k - 1], ts[k]
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if _distance_to_convex(outline, o0 + mid * (o1 - o0)) <= core:
            hi = mid
        else:
            lo = mid
    hit, t = w.ccd()
    assert hit
    assert t[1] == pytest.approx(hi, abs=2e-5)
    assert w.get("comX")[1] == pytest.approx(com0[0] + hi * (com1[0] - com0[0]), abs=2e-4)


# ---------------------------------------------------------------- collider AABBs
def test_aabbs_match_rotated_outline_extremes(oracle, scenes):
    """Box, polygon and circle AABBs (ColliderAABBBuilder.cpp:8-136) equal the extremes of the
    collider's outline rotated with the collider's own cos/sin, computed here corner by corner in
    float64; at angle 0 a box's AABB is exactly its centre +- (half width, half height)."""
    rng = np.random.default_rng(21)
    items = []
    for k, ang in enumerate([0.0, np.pi / 2, np.pi / 4, np.pi / 6, 2.0, 3.5, 5.0, 6.0]):
        items.append(dict(t=1, px=3.0 * k, py=1.0, rot=float(ang), mass=1.0, p0=3.0, p1=1.0))
    for k in range(12):
        items.append(dict(t=2, px=3.0 * k, py=8.0, rot=float(rng.uniform(0, 6.28)), mass=1.0,
                          verts=scenes.polygon_outline(rng, int(rng.integers(3, 9)), float(rng.uniform(0.3, 1.2)))
                          + rng.uniform(-0.2, 0.2, 2).astype(np.float32)))
    items.append(dict(t=0, px=-4.0, py=-4.0, mass=1.0, p0=0.7))
    w = scene_of(oracle, acc_bodies(scenes, items))
    w.poses(True)
    w.aabbs(True)
    box = [w.get(k) for k in ("minX", "minY", "maxX", "maxY")]
    wx, wy, c, s = w.get("wX"), w.get("wY"), w.get("wCos"), w.get("wSin")
    v, _ = w.vertices()
    st, cnt, typ, p0, p1 = w.get("polyStart"), w.get("polyCount"), w.types(), w.get("p0"), w.get("p1")
    for k in range(len(items)):
        if typ[k] == 0:
            local = p0[k] * np.array([[1.0, 0.0], [-1.0, 0.0], [0.0, 1.0], [0.0, -1.0]])
            rot = np.eye(2)
        else:
            if typ[k] == 1:
                local = np.array([[p0[k], p1[k]], [-p0[k], p1[k]], [-p0[k], -p1[k]], [p0[k], -p1[k]]], np.float64)
            else:
                local = v[st[k]:st[k] + cnt[k]].astype(np.float64)
            rot = np.array([[c[k], -s[k]], [s[k], c[k]]], np.float64)
        world = local @ rot.T + [float(wx[k]), float(wy[k])]
        expect = [world[:, 0].min(), world[:, 1].min(), world[:, 0].max(), world[:, 1].max()]
        got = [float(b[k]) for b in box]
        assert np.allclose(got, expect, atol=3e-6), (k, got, expect)
    # the unrotated box: exact, and the half width lies along x
    assert [float(b[0]) for b in box] == [-1.5, 0.5, 1.5, 1.5]
    # a quarter turn swaps the box's extents (fast cos of pi/2 is within 2e-3 of 0)
    assert float(box[2][1] - box[0][1]) == pytest.approx(1.0, abs=5e-3)
    assert float(box[3][1] - box[1][1]) == pytest.approx(3.0, abs=5e-3)


# ---------------------------------------------------------------- friction law
FLOOR = dict(t=1, px=0.0, py=0.0, mass=0.0, p0=80.0, p1=1.0, material=1)


def _slide(oracle, scenes, v0, gravity, steps, mats):
    """A 1 x 1 box resting on a wide static floor (top at y = 0.5), started at velocity v0 along x."""
    box = dict(t=1, px=0.0, py=1.0, vx=float(v0), mass=1.0, p0=1.0, p1=1.0, material=2)
    w = scene_of(oracle, acc_bodies(scenes, [FLOOR, box]), materials=mats,
                 settings=dict(gx=float(gravity[0]), gy=float(gravity[1]), v=8, p=3))
    vx, jt = [], []
    for _ in range(steps):
        w.step(DT)
        vx.append(float(w.get("velX")[1]))
        m = w.manifolds()
        jt.append(float(np.sum(m["jt0"] + np.where(m["count"] == 2, m["jt1"], 0.0))) if len(m) else 0.0)
    return np.array(vx), np.array(jt), w


# static and dynamic coefficients of the box and of the floor; combined by the geometric mean
# (BodyCollisionSolver.cpp:304-305): static sqrt(0.81 * 0.25) = 0.45, dynamic sqrt(0.49 * 0.01) = 0.07
FRICTION_MATS = ((0.0, 0.25, 0.01), (0.0, 0.81, 0.49))
MU_S, MU_D = 0.45, 0.07


@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_sliding_box_decelerates_at_dynamic_friction(oracle, scenes, direction):
    """A box sliding on a level floor loses mu_d * g * dt of speed per step (Coulomb friction at the
    dynamic coefficient; a swapped static/dynamic pair or another combination rule fails), and its
    friction accumulator carries the sign of the slip along x: the tangent is (-n.y, n.x) and the
    impulse +jt t acts on B (BodyCollisionSolver.cpp:311, :469, :561-582), so jt > 0 when the box
    slides towards +x whichever of floor and box is A."""
    vx, jt, _ = _slide(oracle, scenes, 3.0 * direction, (0.0, -9.81), 60, FRICTION_MATS)
    t = np.arange(1, 61) * DT
    expect = direction * (3.0 - MU_D * 9.81 * t)
    assert np.allclose(vx[5:], expect[5:], atol=2e-3)
    assert np.all(np.sign(jt[5:]) == direction)


# combined static sqrt(0.9 * 0.9) = 0.9, dynamic sqrt(0.49 * 0.01) = 0.07
INCLINE_MATS = ((0.0, 0.9, 0.01), (0.0, 0.9, 0.49))


def test_box_on_incline_sticks_below_static_friction(oracle, scenes):
    """Gravity tilted by atan(0.3) along the floor. The static test is made per contact point
    against that point's whole slip (BodyCollisionSolver.cpp:564-580, :733-752), so a box on two
    points sticks while one point can hold the slope (0.3 < 0.9 / 2): a box at rest stays at rest;
    a box already sliding downhill speeds up at g (sin a - mu_d cos a), the dynamic coefficient
    (0.07) being below the slope. With the two coefficients swapped the sliding box would stop."""
    a = np.arctan(0.3)
    g = 9.81 * np.array([np.sin(a), -np.cos(a)])
    vx, _, w = _slide(oracle, scenes, 0.0, g, 90, INCLINE_MATS)
    assert np.max(np.abs(vx[10:])) < 2e-3
    vx, _, _ = _slide(oracle, scenes, 2.0, g, 60, INCLINE_MATS)
    t = np.arange(1, 61) * DT
    acc = 9.81 * (np.sin(a) - MU_D * np.cos(a))
    assert np.allclose(vx[5:], 2.0 + acc * t[5:], atol=5e-3)


# ---------------------------------------------------------------- CCD of spinning bodies
def test_ccd_spinning_body_casts_the_straight_origin_path(oracle, scenes):
    """A fast polygon that turns by 0.75 rad in the step. Its collider origin is the centroid, which
    is the body's centre of mass (ObjectManager.cpp:312-349: a single convex piece is recentred on
    its centroid), so the core circle (half the in-radius about the origin) is cast along the
    straight segment between the origin's previous and current positions whatever the spin, and a
    hit moves the body back by translation only, keeping the new rotation
    (ContinuousCollisionDetector.cpp:353-372, :513-529). Time of impact here: dense sampling plus
    bisection of the exact distance to the rotated static box along that segment."""
    rng = np.random.default_rng(9)
    verts = (scenes.polygon_outline(rng, 6, 0.5) + np.array([0.15, -0.1], np.float32)).astype(np.float32)
    static = dict(t=1, px=7.0, py=0.3, mass=0.0, rot=0.25, p0=1.0, p1=5.0)
    body = dict(t=2, px=0.0, py=0.0, rot=0.4, vx=640.0, vy=15.0, av=45.0, mass=1.0, verts=verts)
    w = scene_of(oracle, acc_bodies(scenes, [static, body]), settings=dict(gy=0.0))
    assert abs(float(w.get("comOffX")[1])) < 1e-6 and abs(float(w.get("comOffY")[1])) < 1e-6
    w.snapshot()
    w.poses(True)
    o0 = np.array([w.get("wX")[1], w.get("wY")[1]], np.float64)
    com0 = np.array([w.get("comX")[1], w.get("comY")[1]], np.float64)
    rot0 = float(w.get("rot")[1])
    w.integrate(DT)
    w.wrap_cossin()
    w.poses()
    w.aabbs()
    o1 = np.array([w.get("wX")[1], w.get("wY")[1]], np.float64)
    com1 = np.array([w.get("comX")[1], w.get("comY")[1]], np.float64)
    rot1 = float(w.get("rot")[1])
    assert abs(rot1 - rot0) > 0.5           # the body turns by 0.75 rad in the step
    v, _ = w.vertices()
    s0, c = w.get("polyStart")[1], w.get("polyCount")[1]
    loc = v[s0:s0 + c].astype(np.float64)
    inradius = min(abs((loc[(i + 1) % c] - loc[i])[0] * loc[i][1] - (loc[(i + 1) % c] - loc[i])[1] * loc[i][0])
                   / np.linalg.norm(loc[(i + 1) % c] - loc[i]) for i in range(c))
    assert w.get("minExtent")[1] == pytest.approx(inradius, rel=1e-5)
    core = 0.5 * inradius
    outline = _world_outlines(w)[0]
    path = lambda t: o0 + t * (o1 - o0)
    ts = np.linspace(0.0, 1.0, 8001)
    dist = np.array([_distance_to_convex(outline, path(t)) for t in ts])
    k = int(np.argmax(dist <= core))
    assert dist[k] <= core and k > 0
    lo, hi = ts[k - 1], ts[k]
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        lo, hi = (lo, mid) if _distance_to_convex(outline, path(mid)) <= core else (mid, hi)
    hit, t = w.ccd()
    assert hit
    assert t[1] == pytest.approx(hi, abs=2e-5)
    assert w.get("comX")[1] == pytest.approx(com0[0] + hi * (com1[0] - com0[0]), abs=2e-4)
    assert w.get("comY")[1] == pytest.approx(com0[1] + hi * (com1[1] - com0[1]), abs=2e-4)
    assert float(w.get("rot")[1]) == rot1   # rotation is kept
