"""Pins the oracle's capsule functions against geometry computed here in float64, independently of
the reference's formulations:
- AABB (ColliderAABBBuilder.cpp:138-175): the tight box of densely sampled boundary points;
- mass properties (PhysicsGeometry.cpp:87-125): area and polar moment by integration over a fine
  grid of the stadium;
- circle-capsule, capsule-capsule, box-capsule, polygon-capsule manifolds
  (CollisionMethods/*Capsule.cpp): depth = sum of radii - distance of the cores, normal along the
  nearest points, points within the capsule, for cores apart; the side-by-side two-point case and the
  centre-on-core fallback by construction; a box and the same box as a polygon give one manifold;
- CCD against a static capsule (ContinuousCollisionDetector.cpp:222-240): bisection of the exact
  distance to the core along the sweep.
"""
import math

import numpy as np
import pytest

DT = 1.0 / 60.0
TAG = 0x80000000
CAP_TAG = 0x40000000


def world_of(oracle, scenes, items, settings=None):
    acc = scenes._Acc()
    for it in items:
        acc.add(**it)
    w = oracle.World(settings)
    w.add_material(0.0, 0.6, 0.6)
    w.add_bodies(acc.bodies())
    w.refresh()
    return w


def narrow_pair(w):
    m = w.narrow(np.array([0], np.uint32), np.array([1], np.uint32))
    return m[0] if len(m) else None


def core_of(w, k):
    """World core segment of capsule collider k, with the oracle's own cos/sin of its rotation."""
    c = np.array([w.get("wX")[k], w.get("wY")[k]], np.float64)
    ax = np.array([w.get("wCos")[k], w.get("wSin")[k]], np.float64)
    h = float(w.get("p0")[k])
    return c - h * ax, c + h * ax, float(w.get("p1")[k])


def point_segment(p, a, b):
    ab = b - a
    t = np.clip((p - a) @ ab / max(ab @ ab, 1e-300), 0.0, 1.0)
    q = a + t * ab
    return float(np.linalg.norm(p - q)), q


def segments_cross(a0, a1, b0, b1):
    def cross(o, p, q):
        return (p[0] - o[0]) * (q[1] - o[1]) - (p[1] - o[1]) * (q[0] - o[0])
    d1, d2 = cross(b0, b1, a0), cross(b0, b1, a1)
    d3, d4 = cross(a0, a1, b0), cross(a0, a1, b1)
    return d1 * d2 < 0 and d3 * d4 < 0


def segment_segment(a0, a1, b0, b1):
    """Distance and a nearest pair (on a, on b) of two segments."""
    if segments_cross(a0, a1, b0, b1):
        return 0.0, None, None
    best = None
    for p, (s0, s1), on_a in ((a0, (b0, b1), True), (a1, (b0, b1), True), (b0, (a0, a1), False), (b1, (a0, a1), False)):
        d, q = point_segment(p, s0, s1)
        if best is None or d < best[0]:
            best = (d, p, q) if on_a else (d, q, p)
    return best


def outline_of(w, k):
    """World CCW outline of a box or polygon collider k."""
    c = np.array([w.get("wX")[k], w.get("wY")[k]], np.float64)
    right = np.array([w.get("wCos")[k], w.get("wSin")[k]], np.float64)
    up = np.array([-right[1], right[0]])
    if int(w.types()[k]) == 1:
        hw, hh = float(w.get("p0")[k]), float(w.get("p1")[k])
        local = np.array([[-hw, -hh], [hw, -hh], [hw, hh], [-hw, hh]])
    else:
        v, _ = w.vertices()
        st, n = int(w.get("polyStart")[k]), int(w.get("polyCount")[k])
        local = v[st:st + n].astype(np.float64)
    return c + local[:, :1] * right + local[:, 1:] * up


def inside_convex(v, p):
    e = np.roll(v, -1, axis=0) - v
    return bool(np.all(e[:, 0] * (p[1] - v[:, 1]) - e[:, 1] * (p[0] - v[:, 0]) >= 0))


def convex_segment(v, s0, s1):
    """Distance and nearest pair (on outline, on segment) of a convex outline and a segment."""
    if inside_convex(v, s0) or inside_convex(v, s1):
        return 0.0, None, None
    best = None
    for i in range(len(v)):
        d, p_out, p_seg = segment_segment(v[i], v[(i + 1) % len(v)], s0, s1)
        if d == 0.0:
            return 0.0, None, None
        if best is None or d < best[0]:
            best = (d, p_out, p_seg)
    return best


def random_polygon(rng, r):
    k = int(rng.integers(3, 9))
    ang = np.sort(rng.uniform(0.0, 2.0 * np.pi, k))
    v = np.stack([r * np.cos(ang), r * np.sin(ang)], 1)
    from scipy.spatial import ConvexHull
    h = ConvexHull(v)
    return v[h.vertices].astype(np.float32)


# ---------------------------------------------------------------- AABB and mass properties
def test_capsule_aabb_is_the_tight_box_of_its_boundary(oracle, scenes):
    rng = np.random.default_rng(11)
    items = [dict(t=3, px=float(rng.uniform(-5, 5)), py=float(rng.uniform(-5, 5)), rot=float(rng.uniform(0, 6.28)),
                  mass=1.0, p0=float(rng.uniform(0.0, 3.0)), p1=float(rng.uniform(0.05, 1.0))) for _ in range(64)]
    w = world_of(oracle, scenes, items)
    phi = np.linspace(0.0, 2.0 * np.pi, 8192, endpoint=False)
    circle = np.stack([np.cos(phi), np.sin(phi)], 1)
    for k in range(len(items)):
        a, b, r = core_of(w, k)
        pts = np.concatenate([a + r * circle, b + r * circle])
        lo, hi = pts.min(0), pts.max(0)
        box = np.array([w.get("minX")[k], w.get("minY")[k], w.get("maxX")[k], w.get("maxY")[k]], np.float64)
        assert np.all(box[:2] <= lo + 1e-5) and np.all(box[2:] >= hi - 1e-5), k
        np.testing.assert_allclose(box, np.concatenate([lo, hi]), atol=2e-5)


@pytest.mark.parametrize("length,radius", [(0.0, 0.5), (1.0, 0.25), (3.0, 0.1), (0.4, 0.6)])
def test_capsule_inertia_matches_grid_integration(oracle, scenes, length, radius):
    mass = 2.5
    w = world_of(oracle, scenes, [dict(t=3, px=0.0, py=0.0, mass=mass, p0=length, p1=radius)])
    h = 0.5 * length
    n = 3001
    xs = np.linspace(-h - radius, h + radius, n)
    ys = np.linspace(-radius, radius, n)
    X, Y = np.meshgrid(xs, ys)
    dx = np.maximum(np.abs(X) - h, 0.0)
    inside = dx * dx + Y * Y <= radius * radius
    cell = (xs[1] - xs[0]) * (ys[1] - ys[0])
    area = inside.sum() * cell
    assert area == pytest.approx(2.0 * radius * length + math.pi * radius * radius, rel=2e-3)
    polar = ((X * X + Y * Y) * inside).sum() * cell
    inertia = float(w.get("inertia")[0])
    assert inertia == pytest.approx(mass * polar / area, rel=2e-3)
    assert float(w.get("invInertia")[0]) == pytest.approx(1.0 / inertia, rel=1e-6)
    # CCD extents: the core circle is the radius, the reach the half length plus the radius
    assert float(w.get("minExtent")[0]) == np.float32(radius)
    assert float(w.get("maxExtent")[0]) == pytest.approx(h + radius, rel=1e-6)


# ---------------------------------------------------------------- narrow phase
def test_circle_capsule_nearest_point(oracle, scenes):
    rng = np.random.default_rng(12)
    checked = 0
    for _ in range(400):
        rc, rk = float(rng.uniform(0.1, 0.6)), float(rng.uniform(0.1, 0.6))
        cap = dict(t=3, px=0.0, py=0.0, rot=float(rng.uniform(0, 6.28)), mass=1.0, p0=float(rng.uniform(0.0, 2.0)), p1=rk)
        ang, dist = rng.uniform(0, 2 * np.pi), rng.uniform(0.0, 1.6)
        circ = dict(t=0, px=float(dist * np.cos(ang)), py=float(dist * np.sin(ang)), mass=1.0, p0=rc)
        w = world_of(oracle, scenes, [circ, cap])
        m = narrow_pair(w)
        c = np.array([w.get("wX")[0], w.get("wY")[0]], np.float64)
        a, b, _ = core_of(w, 1)
        d, q = point_segment(c, a, b)
        if d >= rc + rk:
            assert m is None
            continue
        if d < 1e-3:
            continue
        assert m is not None and m["count"] == 1 and m["a"] == 0 and m["b"] == 1 and m["id0"] == 0
        n = (q - c) / d
        assert m["depth"] == pytest.approx(rc + rk - d, abs=2e-5)
        np.testing.assert_allclose([m["nx"], m["ny"]], n, atol=2e-5)
        np.testing.assert_allclose([m["p0x"], m["p0y"]], c + rc * n, atol=2e-5)
        checked += 1
    assert checked > 100


def test_circle_centre_on_the_core_pushes_out_of_the_side(oracle, scenes):
    """The batched path's fallback (CircleCapsule.cpp:127-134): a centre on the core leaves along
    the capsule's side normal (sin, -cos) flipped when the centre lies on the negative side."""
    w = world_of(oracle, scenes, [dict(t=0, px=0.25, py=0.0, mass=1.0, p0=0.3), dict(t=3, px=0.0, py=0.0, mass=1.0, p0=2.0, p1=0.2)])
    assert float(w.get("wCos")[1]) == 1.0 and float(w.get("wSin")[1]) == 0.0
    m = narrow_pair(w)
    assert (m["nx"], m["ny"]) == (0.0, -1.0) and m["depth"] == np.float32(0.5)


def test_capsule_capsule_nearest_points(oracle, scenes):
    rng = np.random.default_rng(13)
    checked = 0
    for _ in range(400):
        items = []
        for _k in range(2):
            items.append(dict(t=3, px=float(rng.uniform(-1.0, 1.0)), py=float(rng.uniform(-1.0, 1.0)),
                              rot=float(rng.uniform(0, 6.28)), mass=1.0, p0=float(rng.uniform(0.0, 2.0)),
                              p1=float(rng.uniform(0.05, 0.5))))
        w = world_of(oracle, scenes, items)
        m = narrow_pair(w)
        a0, a1, ra = core_of(w, 1)   # canonical A: the higher collider index of two capsules
        b0, b1, rb = core_of(w, 0)
        d, pa, pb = segment_segment(a0, a1, b0, b1)
        ua, ub = (a1 - a0), (b1 - b0)
        cross = abs(ua[0] * ub[1] - ua[1] * ub[0]) / max(np.linalg.norm(ua) * np.linalg.norm(ub), 1e-12)
        if d >= ra + rb:
            assert m is None
            continue
        if d < 1e-3 or cross < 0.05:   # crossing cores, or near-parallel (the side-by-side rule)
            continue
        assert m is not None and m["a"] == 1 and m["b"] == 0 and m["count"] == 1 and m["id0"] == 2
        n = (pb - pa) / d
        assert m["depth"] == pytest.approx(ra + rb - d, abs=3e-5)
        np.testing.assert_allclose([m["nx"], m["ny"]], n, atol=3e-5)
        np.testing.assert_allclose([m["p0x"], m["p0y"]], pa + ra * n, atol=3e-5)
        checked += 1
    assert checked > 100


def test_capsules_side_by_side_touch_at_both_ends_of_the_overlap(oracle, scenes):
    # A = the second capsule (higher index), spanning x in [0.5, 2.5] at y = 0.5; B spans [-1, 1] at y = 0
    w = world_of(oracle, scenes, [dict(t=3, px=0.0, py=0.0, mass=1.0, p0=2.0, p1=0.25),
                                  dict(t=3, px=1.5, py=0.5, mass=1.0, p0=2.0, p1=0.375)])
    m = narrow_pair(w)
    assert m["a"] == 1 and m["count"] == 2 and (m["id0"], m["id1"]) == (0, 1)
    assert (m["nx"], m["ny"]) == (0.0, -1.0)
    assert m["depth"] == pytest.approx(0.125, abs=1e-6)
    # the overlap along A's axis is [-1, -0.5] from A's centre: x = 0.5 and 1.0, on A's surface
    np.testing.assert_allclose([m["p0x"], m["p0y"], m["p1x"], m["p1y"]], [0.5, 0.125, 1.0, 0.125], atol=1e-6)


@pytest.mark.parametrize("kind", ["box", "polygon"])
def test_hull_capsule_depth_and_normal_from_core_distance(oracle, scenes, kind):
    """Cores apart: a face contact's depth is the radius less the true distance of the core from
    the outline, along the nearest points; a vertex contact (id tagged 0x40000000 | vertex) has
    the depth and normal of its own vertex's distance from the core. The reference takes, when the
    capsule's side axis separates best, the outline's extreme vertex along that axis
    (PolygonCapsule.cpp:128-137), which need not be the nearest vertex: then only the second rule
    holds. Both together must cover almost every case."""
    rng = np.random.default_rng(14 if kind == "box" else 15)
    checked = nearest = 0
    for _ in range(1200):
        if kind == "box":
            hull = dict(t=1, px=0.0, py=0.0, rot=float(rng.uniform(0, 6.28)), mass=1.0,
                        p0=float(rng.uniform(0.3, 2.0)), p1=float(rng.uniform(0.3, 2.0)))
        else:
            hull = dict(t=2, px=0.0, py=0.0, rot=float(rng.uniform(0, 6.28)), mass=1.0,
                        verts=random_polygon(rng, float(rng.uniform(0.4, 1.2))))
        ang, dist = rng.uniform(0, 2 * np.pi), rng.uniform(0.3, 2.5)
        cap = dict(t=3, px=float(dist * np.cos(ang)), py=float(dist * np.sin(ang)), rot=float(rng.uniform(0, 6.28)),
                   mass=1.0, p0=float(rng.uniform(0.0, 2.0)), p1=float(rng.uniform(0.05, 0.5)))
        w = world_of(oracle, scenes, [hull, cap])
        m = narrow_pair(w)
        v = outline_of(w, 0)
        s0, s1, r = core_of(w, 1)
        d, p_hull, p_core = convex_segment(v, s0, s1)
        if d >= r:
            assert m is None
            continue
        if d < 1e-3:
            continue
        assert m is not None and m["a"] == 0 and m["b"] == 1 and m["count"] in (1, 2)
        ident = int(m["id0"])
        if ident & CAP_TAG:
            assert m["count"] == 1
            corner = v[ident & 0xFFFF] if kind == "polygon" else None
            if corner is None:   # box corners in the order (-,-) (+,-) (+,+) (-,+), as outline_of lists them
                corner = v[ident & 3]
            dv, q = point_segment(corner, s0, s1)
            assert m["depth"] == pytest.approx(r - dv, abs=5e-5)
            np.testing.assert_allclose([m["nx"], m["ny"]], (q - corner) / dv, atol=1e-4)
            assert dv >= d - 1e-6
        else:
            # a face wins ties with a vertex by 1e-4 of squared distance and with the side axis by
            # 1e-4 of separation (Constants.h:9): its distance lies in [d, sqrt(d^2 + 1e-4) + 1e-4]
            assert ident & TAG
            face = r - float(m["depth"])
            assert d - 5e-5 <= face <= math.sqrt(d * d + 1e-4) + 1.5e-4
            if abs(face - d) < 5e-5:
                np.testing.assert_allclose([m["nx"], m["ny"]], (p_core - p_hull) / d, atol=2e-3)
        nearest += abs(float(m["depth"]) - (r - d)) < 5e-5
        for k in range(int(m["count"])):
            p = np.array([m[f"p{k}x"], m[f"p{k}y"]], np.float64)
            dc, _ = point_segment(p, s0, s1)
            assert dc <= r + 1e-4                               # within the capsule
        checked += 1
    assert checked > 150 and nearest >= 0.98 * checked


def test_box_and_the_same_box_as_a_polygon_give_one_manifold(oracle, scenes):
    rng = np.random.default_rng(16)
    checked = 0
    for _ in range(400):
        hw, hh = float(rng.uniform(0.3, 1.5)), float(rng.uniform(0.3, 1.5))
        rot = float(rng.uniform(0, 6.28))
        ang, dist = rng.uniform(0, 2 * np.pi), rng.uniform(0.0, 2.5)
        cap = dict(t=3, px=float(dist * np.cos(ang)), py=float(dist * np.sin(ang)), rot=float(rng.uniform(0, 6.28)),
                   mass=1.0, p0=float(rng.uniform(0.0, 2.0)), p1=float(rng.uniform(0.05, 0.5)))
        square = np.array([[-hw, -hh], [hw, -hh], [hw, hh], [-hw, hh]], np.float32)
        mb = narrow_pair(world_of(oracle, scenes, [dict(t=1, px=0.0, py=0.0, rot=rot, mass=1.0, p0=2 * hw, p1=2 * hh), cap]))
        mp = narrow_pair(world_of(oracle, scenes, [dict(t=2, px=0.0, py=0.0, rot=rot, mass=1.0, verts=square), cap]))
        assert (mb is None) == (mp is None)
        if mb is None:
            continue
        assert mb["count"] == mp["count"]
        np.testing.assert_allclose([mb["nx"], mb["ny"], mb["depth"]], [mp["nx"], mp["ny"], mp["depth"]], atol=2e-5)
        pts_b = sorted((round(float(mb[f"p{k}x"]), 4), round(float(mb[f"p{k}y"]), 4)) for k in range(mb["count"]))
        pts_p = sorted((round(float(mp[f"p{k}x"]), 4), round(float(mp[f"p{k}y"]), 4)) for k in range(mp["count"]))
        np.testing.assert_allclose(pts_b, pts_p, atol=2e-4)
        checked += 1
    assert checked > 150


# ---------------------------------------------------------------- CCD
@pytest.mark.parametrize("seed", range(6))
def test_ccd_circle_into_static_capsule(oracle, scenes, seed):
    """A fast circle crossing a rotated static capsule within one step stops where its core circle
    (half its radius) first comes within the capsule's radius of the capsule's core."""
    rng = np.random.default_rng(100 + seed)
    length, radius = float(rng.uniform(0.5, 4.0)), float(rng.uniform(0.1, 0.6))
    static = dict(t=3, px=5.0, py=0.0, rot=float(rng.uniform(0.0, 6.28)), mass=0.0, p0=length, p1=radius)
    vy = float(rng.uniform(-60.0, 60.0))
    w = world_of(oracle, scenes, [static, dict(t=0, px=0.0, py=float(rng.uniform(-0.5, 0.5)), vx=600.0, vy=vy,
                                                mass=1.0, p0=0.2)], settings=dict(gy=0.0))
    w.snapshot()
    c0 = np.array([w.get("comX")[1], w.get("comY")[1]], np.float64)
    w.integrate(DT)
    w.wrap_cossin()
    w.poses()
    w.aabbs()
    c1 = np.array([w.get("comX")[1], w.get("comY")[1]], np.float64)
    a, b, r = core_of(w, 0)
    reach = r + 0.1
    ts = np.linspace(0.0, 1.0, 4001)
    dist = np.array([point_segment(c0 + t * (c1 - c0), a, b)[0] for t in ts])
    hit, t = w.ccd()
    if not np.any(dist <= reach):
        assert not hit
        return
    k = int(np.argmax(dist <= reach))
    assert k > 0
    lo, hi = ts[k - 1], ts[k]
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if point_segment(c0 + mid * (c1 - c0), a, b)[0] <= reach:
            hi = mid
        else:
            lo = mid
    assert hit
    assert t[1] == pytest.approx(hi, abs=2e-5)
