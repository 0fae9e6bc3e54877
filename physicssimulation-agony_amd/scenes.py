"""Seeded synthetic scenes for the collide-and-solve path (numpy only, no device code).

A scene is a dict: ``materials`` (list of (elasticity, static friction, dynamic friction),
created in order as material 1, 2, ...), ``bodies`` (arrays in creation order: type, px, py,
rot, vx, vy, av, mass, material, p0, p1, vstart, vcount and a float32 (V, 2) vertex pool
``verts``) and ``settings`` (overrides for the world settings). Static bodies have mass 0.
Masses are density x area, computed in float64 and rounded to float32.

Workloads (all FP32, seeded):
  pile_small    2,048 convex polygons (3-8 vertices on a jittered circle, R ~ U[0.2, 0.6]) and
                512 circles (r ~ U[0.2, 0.5]) dropped into a static 64 x 64 m box, density 1,
                friction 0.6, restitution 0, dt 1/60, 8 velocity + 3 position iterations.
  pile_large    65,536 bodies (70 % polygons, 30 % circles), radii lognormal(ln 0.3, 0.3)
                clipped to [0.1, 0.9] m, densities U[0.5, 2], packed shelf by shelf into a
                512 x 256 m box where they settle.
  scatter       bodies uniform in a square domain, lognormal sizes, random orientation, 5 %
                small bullets at 200 m/s, plus static obstacles so that CCD has work.
  batch         independent scenes with log-uniform body counts (pyramid, pile, funnel).
"""
from __future__ import annotations

import math

import numpy as np

SHAPE_CIRCLE, SHAPE_BOX, SHAPE_POLYGON, SHAPE_CAPSULE = 0, 1, 2, 3


class _Acc:
    """Accumulates bodies in creation order."""

    def __init__(self):
        self.cols = {k: [] for k in ("type", "px", "py", "rot", "vx", "vy", "av", "mass", "material", "p0", "p1",
                                     "vstart", "vcount")}
        self.verts: list[np.ndarray] = []
        self.nverts = 0

    def add(self, t, px, py, rot=0.0, vx=0.0, vy=0.0, av=0.0, mass=0.0, material=1, p0=0.0, p1=0.0, verts=None):
        c = self.cols
        c["type"].append(t); c["px"].append(px); c["py"].append(py); c["rot"].append(rot)
        c["vx"].append(vx); c["vy"].append(vy); c["av"].append(av); c["mass"].append(mass)
        c["material"].append(material); c["p0"].append(p0); c["p1"].append(p1)
        if verts is not None:
            c["vstart"].append(self.nverts); c["vcount"].append(len(verts))
            self.verts.append(np.asarray(verts, np.float32))
            self.nverts += len(verts)
        else:
            c["vstart"].append(0); c["vcount"].append(0)

    def bodies(self) -> dict:
        out = {k: np.asarray(v) for k, v in self.cols.items()}
        out["type"] = out["type"].astype(np.uint8)
        for k in ("px", "py", "rot", "vx", "vy", "av", "mass", "p0", "p1"):
            out[k] = out[k].astype(np.float32)
        for k in ("material", "vstart", "vcount"):
            out[k] = out[k].astype(np.uint32)
        out["verts"] = np.concatenate(self.verts).astype(np.float32) if self.verts else np.zeros((0, 2), np.float32)
        return out


def polygon_outline(rng: np.random.Generator, k: int, radius: float) -> np.ndarray:
    """k vertices at angles 2*pi*(i + U(-0.3, 0.3)) / k on a circle: convex, counter-clockwise."""
    ang = 2.0 * np.pi * (np.arange(k) + rng.uniform(-0.3, 0.3, k)) / k
    return np.stack([radius * np.cos(ang), radius * np.sin(ang)], axis=1).astype(np.float32)


def polygon_area(v: np.ndarray) -> float:
    x = v[:, 0].astype(np.float64)
    y = v[:, 1].astype(np.float64)
    return 0.5 * abs(float(np.dot(x, np.roll(y, -1)) - np.dot(np.roll(x, -1), y)))


def add_walls(acc: _Acc, x0: float, y0: float, x1: float, y1: float, thickness: float = 2.0, top: bool = True):
    """Static box walls around [x0, x1] x [y0, y1] (inner faces on the rectangle)."""
    w, h, t = x1 - x0, y1 - y0, thickness
    acc.add(SHAPE_BOX, (x0 + x1) / 2, y0 - t / 2, p0=w + 2 * t, p1=t)
    if top:
        acc.add(SHAPE_BOX, (x0 + x1) / 2, y1 + t / 2, p0=w + 2 * t, p1=t)
    acc.add(SHAPE_BOX, x0 - t / 2, (y0 + y1) / 2, p0=t, p1=h)
    acc.add(SHAPE_BOX, x1 + t / 2, (y0 + y1) / 2, p0=t, p1=h)


def _radii(rng, n, dist):
    kind = dist[0]
    if kind == "uniform":
        return rng.uniform(dist[1], dist[2], n)
    if kind == "lognormal":
        return np.clip(rng.lognormal(math.log(dist[1]), dist[2], n), dist[3], dist[4])
    raise ValueError(kind)


def _densities(rng, n, dist):
    return np.full(n, dist[1]) if dist[0] == "const" else rng.uniform(dist[1], dist[2], n)


def _shapes(rng, n_poly, n_circle, r_poly, r_circle, density):
    """Shape list in creation order (polygons then circles), each (type, radius, verts, mass)."""
    out = []
    rp = _radii(rng, n_poly, r_poly)
    dp = _densities(rng, n_poly, density)
    for i in range(n_poly):
        k = int(rng.integers(3, 9))
        v = polygon_outline(rng, k, float(rp[i]))
        out.append((SHAPE_POLYGON, float(rp[i]), v, float(np.float32(dp[i] * polygon_area(v)))))
    rc = _radii(rng, n_circle, r_circle)
    dc = _densities(rng, n_circle, density)
    for i in range(n_circle):
        r = float(np.float32(rc[i]))
        out.append((SHAPE_CIRCLE, r, None, float(np.float32(dc[i] * math.pi * r * r))))
    return out


def _place_uniform(rng, shapes, x0, y0, x1, y1, gap=0.02, tries=200):
    """Rejection sampling of non-overlapping bounding circles (grid hash)."""
    cell = 2.0 * max(s[1] for s in shapes) + gap
    grid: dict = {}
    pos = []
    for t, r, v, m in shapes:
        for _ in range(tries):
            x = rng.uniform(x0 + r, x1 - r)
            y = rng.uniform(y0 + r, y1 - r)
            gx, gy = int(x // cell), int(y // cell)
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for (qx, qy, qr) in grid.get((gx + dx, gy + dy), ()):
                        if (qx - x) ** 2 + (qy - y) ** 2 < (qr + r + gap) ** 2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                grid.setdefault((gx, gy), []).append((x, y, r))
                pos.append((x, y))
                break
        else:
            raise RuntimeError("could not place body without overlap; domain too small")
    return pos


def _place_shelves(shapes, x0, y0, x1, gap=0.05):
    """Shelf packing from the floor up: bodies in a row, the row height set by its largest body."""
    pos = []
    x, y, row_h = x0, y0, 0.0
    for t, r, v, m in shapes:
        if x + 2 * r + gap > x1:
            x = x0
            y += row_h + gap
            row_h = 0.0
        pos.append((x + r + gap / 2, y + r + gap / 2))
        x += 2 * r + gap
        row_h = max(row_h, 2 * r)
    return pos


def pile(n_poly: int, n_circle: int, width: float, height: float, seed: int, r_poly=("uniform", 0.2, 0.6),
         r_circle=("uniform", 0.2, 0.5), density=("const", 1.0), layout: str = "uniform", friction: float = 0.6,
         restitution: float = 0.0, shuffle: bool = False, name: str = "pile") -> dict:
    rng = np.random.default_rng(seed)
    acc = _Acc()
    add_walls(acc, -width / 2, 0.0, width / 2, height)
    shapes = _shapes(rng, n_poly, n_circle, r_poly, r_circle, density)
    if shuffle:
        order = rng.permutation(len(shapes))
        shapes = [shapes[i] for i in order]
    if layout == "uniform":
        pos = _place_uniform(rng, shapes, -width / 2, 0.0, width / 2, height)
    else:
        pos = _place_shelves(shapes, -width / 2, 0.0, width / 2)
        if pos and pos[-1][1] > height:
            raise RuntimeError("shelf packing exceeds the box height")
    rots = rng.uniform(0.0, 2.0 * np.pi, len(shapes)).astype(np.float32)
    rots = np.minimum(rots, np.float32(6.2831850))
    for (t, r, v, m), (x, y), rot in zip(shapes, pos, rots):
        if t == SHAPE_CIRCLE:
            acc.add(t, x, y, rot=float(rot), mass=m, p0=r)
        else:
            acc.add(t, x, y, rot=float(rot), mass=m, verts=v)
    return {"name": name, "materials": [(restitution, friction, friction)], "bodies": acc.bodies(),
            "settings": {"velocity_iterations": 8, "position_iterations": 3, "gravity_y": -9.81}}


def free_swarm(n_poly: int = 2048, n_circle: int = 512, size: float = 64.0, speed: float = 3.0, seed: int = 6,
               name: str = "free_swarm") -> dict:
    """The small pile's shapes without walls or gravity, started at random velocities (uniform in
    a disc of radius `speed`) and spins: every contact is between two dynamic bodies, so total
    linear momentum is conserved by the solver's equal and opposite impulses."""
    rng = np.random.default_rng(seed)
    acc = _Acc()
    shapes = _shapes(rng, n_poly, n_circle, ("uniform", 0.2, 0.6), ("uniform", 0.2, 0.5), ("const", 1.0))
    pos = _place_uniform(rng, shapes, -size / 2, -size / 2, size / 2, size / 2)
    for (t, r, v, m), (x, y) in zip(shapes, pos):
        ang, mag = rng.uniform(0.0, 2.0 * np.pi), speed * math.sqrt(rng.uniform())
        kw = dict(rot=float(rng.uniform(0.0, 6.28)), vx=float(mag * math.cos(ang)), vy=float(mag * math.sin(ang)),
                  av=float(rng.uniform(-1.0, 1.0)), mass=m)
        if t == SHAPE_CIRCLE:
            acc.add(t, x, y, p0=r, **kw)
        else:
            acc.add(t, x, y, verts=v, **kw)
    return {"name": name, "materials": [(0.0, 0.6, 0.6)], "bodies": acc.bodies(),
            "settings": {"velocity_iterations": 8, "position_iterations": 3, "gravity_y": 0.0}}


def capsule_pile(n: int = 1200, width: float = 48.0, height: float = 110.0, seed: int = 7, bullets: int = 64,
                 name: str = "capsule_pile") -> dict:
    """Every shape pair: capsules (30 %), polygons (30 %), boxes (20 %) and circles (20 %) dropped
    from random non-overlapping places into a walled box with static capsule and box ramps, plus
    `bullets` small fast circles fired down at the ramps (CCD against static capsules)."""
    rng = np.random.default_rng(seed)
    acc = _Acc()
    add_walls(acc, -width / 2, 0.0, width / 2, height)
    # static ramps: capsules and boxes at +-20 degrees, two rows
    for k, x in enumerate(np.linspace(-width / 2 + 8.0, width / 2 - 8.0, 4)):
        ang = (0.35 if k % 2 else -0.35)
        acc.add(SHAPE_CAPSULE, float(x), 14.0, rot=float(ang % (2 * np.pi)), p0=7.0, p1=0.6)
        acc.add(SHAPE_BOX, float(x) + 3.0, 28.0, rot=float((-ang) % (2 * np.pi)), p0=6.0, p1=0.8)
    shapes = []
    kinds = rng.choice(4, size=n, p=[0.3, 0.3, 0.2, 0.2])
    for kind in kinds:
        dens = float(rng.uniform(0.5, 2.0))
        if kind == 0:       # capsule: core length and radius
            r = float(np.float32(rng.uniform(0.12, 0.35)))
            length = float(np.float32(rng.uniform(0.2, 1.4)))
            m = float(np.float32(dens * (2.0 * r * length + math.pi * r * r)))
            shapes.append((SHAPE_CAPSULE, 0.5 * length + r, (length, r), m))
        elif kind == 1:
            rp = float(rng.uniform(0.2, 0.55))
            v = polygon_outline(rng, int(rng.integers(3, 9)), rp)
            shapes.append((SHAPE_POLYGON, rp, v, float(np.float32(dens * polygon_area(v)))))
        elif kind == 2:
            w_, h_ = float(np.float32(rng.uniform(0.3, 1.0))), float(np.float32(rng.uniform(0.3, 1.0)))
            shapes.append((SHAPE_BOX, 0.5 * math.hypot(w_, h_), (w_, h_), float(np.float32(dens * w_ * h_))))
        else:
            r = float(np.float32(rng.uniform(0.15, 0.45)))
            shapes.append((SHAPE_CIRCLE, r, None, float(np.float32(dens * math.pi * r * r))))
    pos = _place_uniform(rng, shapes, -width / 2, 32.0, width / 2, height)
    for (t, r, v, m), (x, y) in zip(shapes, pos):
        rot = float(min(rng.uniform(0.0, 2.0 * np.pi), 6.2831850))
        if t == SHAPE_CAPSULE:
            acc.add(t, x, y, rot=rot, mass=m, p0=v[0], p1=v[1])
        elif t == SHAPE_BOX:
            acc.add(t, x, y, rot=rot, mass=m, p0=v[0], p1=v[1])
        elif t == SHAPE_CIRCLE:
            acc.add(t, x, y, rot=rot, mass=m, p0=r)
        else:
            acc.add(t, x, y, rot=rot, mass=m, verts=v)
    # bullets between the ramps rows, fired downwards at 150 m/s (several ramp thicknesses per step)
    for k in range(bullets):
        x = float(rng.uniform(-width / 2 + 3.0, width / 2 - 3.0))
        r = float(np.float32(rng.uniform(0.05, 0.1)))
        acc.add(SHAPE_CIRCLE, x, 20.0 + float(rng.uniform(0.0, 4.0)), vx=float(rng.uniform(-20.0, 20.0)), vy=-150.0,
                mass=float(np.float32(math.pi * r * r)), p0=r)
    return {"name": name, "materials": [(0.0, 0.5, 0.5)], "bodies": acc.bodies(),
            "settings": {"velocity_iterations": 8, "position_iterations": 3, "gravity_y": -9.81}}


def pile_small(seed: int = 1) -> dict:
    """2,048 polygons + 512 circles in a 64 x 64 m box."""
    return pile(2048, 512, 64.0, 64.0, seed, name="pile_small")


def pile_large(seed: int = 2, n: int = 65536) -> dict:
    """65,536 bodies (70 % polygons) in a 512 x 256 m box, packed from the floor."""
    n_poly = int(round(0.7 * n))
    return pile(n_poly, n - n_poly, 512.0, 256.0, seed, r_poly=("lognormal", 0.3, 0.3, 0.1, 0.9),
                r_circle=("lognormal", 0.3, 0.3, 0.1, 0.9), density=("uniform", 0.5, 2.0), layout="shelves",
                shuffle=True, name="pile_large")


def scatter(n: int, size: float, seed: int, n_static: int = 0, bullet_fraction: float = 0.05,
            bullet_speed: float = 200.0, name: str = "scatter") -> dict:
    """Bodies uniform in [-size/2, size/2]^2 with random orientation, lognormal(ln 0.3, 0.3) radii
    (70 % polygons with 3-8 vertices, the rest circles), 5 % small bullets (radius x 0.25, circles)
    at 200 m/s in random directions, and static boxes and polygons so that CCD has work. No
    gravity; the world steps collide-only."""
    rng = np.random.default_rng(seed)
    # static obstacles: boxes and polygons of 3-12 m
    sx = rng.uniform(-size / 2, size / 2, n_static)
    sy = rng.uniform(-size / 2, size / 2, n_static)
    srot = np.minimum(rng.uniform(0, 2 * np.pi, n_static), 6.2831850)
    sbox = rng.random(n_static) < 0.5
    ns_poly = int((~sbox).sum())
    _, sk, sverts = _polygons(rng, ns_poly, 1.5, 4.0)
    sp0 = np.where(sbox, rng.uniform(2, 8, n_static), 0.0)
    sp1 = np.where(sbox, rng.uniform(2, 8, n_static), 0.0)
    svc = np.zeros(n_static, np.uint32)
    svc[~sbox] = sk
    # dynamic bodies
    radii = np.clip(rng.lognormal(math.log(0.3), 0.3, n), 0.1, 0.9)
    bullets = rng.random(n) < bullet_fraction
    radii = np.where(bullets, radii * 0.25, radii).astype(np.float32)
    x = rng.uniform(-size / 2, size / 2, n)
    y = rng.uniform(-size / 2, size / 2, n)
    rot = np.minimum(rng.uniform(0, 2 * np.pi, n), 6.2831850)
    ang = rng.uniform(0, 2 * np.pi, n)
    is_poly = (rng.random(n) < 0.7) & ~bullets
    dens = rng.uniform(0.5, 2.0, n)
    npoly = int(is_poly.sum())
    k = rng.integers(3, 9, npoly)
    a8 = 2.0 * np.pi * (np.arange(8)[None, :] + rng.uniform(-0.3, 0.3, (npoly, 8))) / k[:, None]
    mask = np.arange(8)[None, :] < k[:, None]
    rp = radii[is_poly].astype(np.float64)[:, None]
    dverts = np.stack([(rp * np.cos(a8))[mask], (rp * np.sin(a8))[mask]], axis=1).astype(np.float32)
    mass = np.zeros(n)
    mass[is_poly] = dens[is_poly] * _polygon_areas(dverts, k)
    mass[~is_poly] = dens[~is_poly] * np.pi * radii[~is_poly].astype(np.float64) ** 2
    vcount = np.zeros(n, np.uint32)
    vcount[is_poly] = k
    p0 = np.where(is_poly, 0.0, radii)
    vx = np.where(bullets, bullet_speed * np.cos(ang), 0.0)
    vy = np.where(bullets, bullet_speed * np.sin(ang), 0.0)
    typ = np.where(is_poly, SHAPE_POLYGON, SHAPE_CIRCLE).astype(np.uint8)
    vc_all = np.concatenate([svc, vcount])
    vstart = np.zeros(len(vc_all), np.uint32)
    has = vc_all > 0
    vstart[has] = np.concatenate([[0], np.cumsum(vc_all[has])[:-1]]).astype(np.uint32)
    body = {
        "type": np.concatenate([np.where(sbox, SHAPE_BOX, SHAPE_POLYGON).astype(np.uint8), typ]),
        "px": np.concatenate([sx, x]).astype(np.float32), "py": np.concatenate([sy, y]).astype(np.float32),
        "rot": np.concatenate([srot, rot]).astype(np.float32),
        "vx": np.concatenate([np.zeros(n_static), vx]).astype(np.float32),
        "vy": np.concatenate([np.zeros(n_static), vy]).astype(np.float32),
        "av": np.zeros(n_static + n, np.float32),
        "mass": np.concatenate([np.zeros(n_static), mass]).astype(np.float32),
        "material": np.ones(n_static + n, np.uint32),
        "p0": np.concatenate([sp0, p0]).astype(np.float32), "p1": np.concatenate([sp1, np.zeros(n)]).astype(np.float32),
        "vstart": vstart, "vcount": vc_all,
        "verts": np.concatenate([sverts, dverts]).astype(np.float32),
    }
    return {"name": name, "materials": [(0.0, 0.6, 0.6)], "bodies": body,
            "settings": {"gravity_y": 0.0, "collide_only": 1}}


def scatter_large(seed: int = 3) -> dict:
    """1,048,576 dynamic colliders in a 4096 x 4096 m domain plus 16,384 static obstacles."""
    return scatter(1 << 20, 4096.0, seed, n_static=16384, name="scatter_large")


def pyramid(rows: int, seed: int, box: float = 1.0) -> dict:
    """Stacked box pyramid on a static floor."""
    rng = np.random.default_rng(seed)
    acc = _Acc()
    acc.add(SHAPE_BOX, 0.0, -1.0, p0=rows * box * 3.0 + 10.0, p1=2.0)
    for r in range(rows):
        count = rows - r
        for c in range(count):
            x = (c - (count - 1) / 2.0) * box * 1.02 + rng.uniform(-0.01, 0.01)
            y = box / 2 + r * box * 1.001
            acc.add(SHAPE_BOX, float(x), float(y), mass=box * box, p0=box, p1=box)
    return {"name": "pyramid", "materials": [(0.0, 0.6, 0.6)], "bodies": acc.bodies(),
            "settings": {"velocity_iterations": 8, "position_iterations": 3}}


def sleepy_stacks(stacks: int = 8, height: int = 3, seed: int = 9) -> dict:
    """Stacks of boxes and polygons on a static floor that settle and fall asleep, with one ball per
    stack dropped from a height that makes it land after the stack sleeps (each at another time), and
    a static ramp the balls of the outer stacks roll off. Sleeping on (0.1 m/s, 0.5 s)."""
    rng = np.random.default_rng(seed)
    acc = _Acc()
    span = stacks * 4.0
    acc.add(SHAPE_BOX, 0.0, -1.0, p0=span + 20.0, p1=2.0)                       # floor, top at y = 0
    acc.add(SHAPE_BOX, -span / 2 - 4.0, 3.0, rot=0.5, p0=6.0, p1=0.5)            # ramp
    for k in range(stacks):
        x = (k - (stacks - 1) / 2.0) * 4.0
        for j in range(height):
            jitter = float(rng.uniform(-0.05, 0.05))
            if (j + k) % 3 == 2:
                v = polygon_outline(rng, int(rng.integers(5, 9)), 0.6)
                acc.add(SHAPE_POLYGON, x + jitter, 0.62 + 1.25 * j, mass=float(polygon_area(v)), verts=v)
            else:
                acc.add(SHAPE_BOX, x + jitter, 0.5 + 1.02 * j, mass=1.0, p0=1.0 + 0.2 * (j % 2), p1=1.0)
        acc.add(SHAPE_CIRCLE, x + float(rng.uniform(-0.3, 0.3)), 20.0 + 9.0 * k, mass=0.6, p0=0.35)
    return {"name": "sleepy_stacks", "materials": [(0.0, 0.6, 0.6)], "bodies": acc.bodies(),
            "settings": {"velocity_iterations": 8, "position_iterations": 3, "sleep_speed_threshold": 0.1,
                         "sleep_time_threshold": 0.5}}


def joint_rig(seed: int = 12, ropes: int = 3, links: int = 10) -> dict:
    """Every joint type at work beside contacts: ropes of boxes on revolute joints hanging from static
    anchors, circles paired by springs, a piston on a prismatic joint with limits, a block on a slider
    with limits, motor-driven wheels on revolute axles under a chassis, and loose polygons dropping
    onto them. Static bodies are created first, so creation order is the dense order and the joints'
    body indices are creation indices."""
    rng = np.random.default_rng(seed)
    acc = _Acc()
    statics = []

    def static(x, y, w, h, rot=0.0):
        acc.add(SHAPE_BOX, x, y, rot=rot, p0=w, p1=h)
        statics.append(len(statics))
        return len(statics) - 1

    ground = static(0.0, -1.0, 90.0, 2.0)
    anchors = [static(-30.0 + 9.0 * r, 14.0, 0.4, 0.4) for r in range(ropes)]
    rail = static(12.0, 8.0, 0.3, 0.3)
    post = static(22.0, 2.0, 0.3, 0.3)
    n_static = len(statics)
    bodies = [n_static]   # next dynamic index
    joints = []

    def dynamic(t, x, y, **kw):
        acc.add(t, x, y, **kw)
        bodies[0] += 1
        return bodies[0] - 1

    # ropes: link k spans [x0 + 0.5 k, x0 + 0.5 (k + 1)], pinned end to end
    for r in range(ropes):
        x0, y0 = -30.0 + 9.0 * r, 14.0
        prev, prev_anchor = anchors[r], (0.0, 0.0)
        for k in range(links):
            b = dynamic(SHAPE_BOX, x0 + 0.5 * k + 0.25, y0, mass=0.2, p0=0.5, p1=0.15)
            joints.append(dict(type=1, body_a=prev, body_b=b, anchor_a_x=prev_anchor[0], anchor_a_y=prev_anchor[1],
                               anchor_b_x=-0.25, anchor_b_y=0.0, stiffness=3000.0, damping=20.0))
            prev, prev_anchor = b, (0.25, 0.0)
    # spring pairs
    for k in range(6):
        x = -4.0 + 2.5 * k
        a = dynamic(SHAPE_CIRCLE, x, 6.0, vx=float(rng.uniform(-1, 1)), mass=0.5, p0=0.3)
        b = dynamic(SHAPE_CIRCLE, x + 0.4, 7.6, mass=0.5, p0=0.3)
        joints.append(dict(type=0, body_a=a, body_b=b, rest_length=1.2, stiffness=60.0, damping=0.8))
    # piston on a horizontal rail, moving at 3 m/s into its upper limit
    piston = dynamic(SHAPE_BOX, 12.0, 8.0, vx=3.0, mass=1.0, p0=0.8, p1=0.4)
    joints.append(dict(type=2, body_a=rail, body_b=piston, axis_x=1.0, axis_y=0.0, stiffness=5000.0, damping=50.0,
                       lower_limit=-2.0, upper_limit=1.5, enable_limit=1))
    # block on a tilted slider (rotation free), sliding down to its lower limit
    block = dynamic(SHAPE_BOX, 22.0, 2.0, rot=0.3, mass=0.8, p0=0.6, p1=0.6)
    joints.append(dict(type=3, body_a=post, body_b=block, axis_x=0.6, axis_y=0.8, stiffness=4000.0, damping=40.0,
                       lower_limit=-1.0, upper_limit=3.0, enable_limit=1))
    # two cars: a chassis on two motor-driven wheels
    for c in range(2):
        cx = 30.0 + 8.0 * c
        chassis = dynamic(SHAPE_BOX, cx, 1.4, mass=2.0, p0=2.4, p1=0.5)
        for side in (-1.0, 1.0):
            wheel = dynamic(SHAPE_CIRCLE, cx + side * 0.9, 0.55, mass=0.5, p0=0.5)
            joints.append(dict(type=1, body_a=chassis, body_b=wheel, anchor_a_x=side * 0.9, anchor_a_y=-0.85,
                               stiffness=8000.0, damping=60.0))
            joints.append(dict(type=4, body_a=wheel, target_speed=-6.0 if c == 0 else 4.0, max_torque=6.0))
    # loose polygons dropped onto the ropes and springs
    for k in range(16):
        v = polygon_outline(rng, int(rng.integers(3, 8)), float(rng.uniform(0.25, 0.45)))
        dynamic(SHAPE_POLYGON, float(rng.uniform(-30, 15)), float(rng.uniform(16, 24)), rot=float(rng.uniform(0, 6)),
                mass=float(polygon_area(v)), verts=v)
    return {"name": "joint_rig", "materials": [(0.0, 0.5, 0.5)], "bodies": acc.bodies(), "joints": joints,
            "settings": {"velocity_iterations": 8, "position_iterations": 3}}


# ---------------------------------------------------------------- batch of independent scenes
def _polygons(rng, n, r_lo, r_hi):
    """n convex polygons with 3-8 vertices on jittered circles: (radius, vcount, verts (V, 2))."""
    k = rng.integers(3, 9, n)
    r = rng.uniform(r_lo, r_hi, n)
    ang = 2.0 * np.pi * (np.arange(8)[None, :] + rng.uniform(-0.3, 0.3, (n, 8))) / k[:, None]
    mask = np.arange(8)[None, :] < k[:, None]
    vx = (r[:, None] * np.cos(ang))[mask]
    vy = (r[:, None] * np.sin(ang))[mask]
    return r, k, np.stack([vx, vy], axis=1).astype(np.float32)


def _polygon_areas(verts, counts):
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    x = verts[:, 0].astype(np.float64)
    y = verts[:, 1].astype(np.float64)
    nxt = np.arange(len(verts)) + 1
    ends = starts + counts
    nxt[ends - 1] = starts                      # wrap each polygon
    cross = x * y[nxt] - x[nxt] * y
    return 0.5 * np.abs(np.add.reduceat(cross, starts))


def _scene_arrays(static, dyn_type, dyn_x, dyn_y, dyn_rot, dyn_mass, dyn_p0, dyn_p1, dyn_vcount, dyn_verts):
    """Body dict from static boxes [(x, y, rot, w, h)] and dynamic body arrays."""
    ns, nd = len(static), len(dyn_type)
    st = np.asarray(static, np.float64).reshape(ns, 5)
    vstart = np.zeros(nd, np.uint32)
    vc = np.asarray(dyn_vcount, np.uint32)
    vstart[vc > 0] = np.concatenate([[0], np.cumsum(vc[vc > 0])[:-1]]).astype(np.uint32)
    return {
        "type": np.concatenate([np.full(ns, SHAPE_BOX, np.uint8), np.asarray(dyn_type, np.uint8)]),
        "px": np.concatenate([st[:, 0], dyn_x]).astype(np.float32),
        "py": np.concatenate([st[:, 1], dyn_y]).astype(np.float32),
        "rot": np.concatenate([st[:, 2], dyn_rot]).astype(np.float32),
        "vx": np.zeros(ns + nd, np.float32), "vy": np.zeros(ns + nd, np.float32), "av": np.zeros(ns + nd, np.float32),
        "mass": np.concatenate([np.zeros(ns), dyn_mass]).astype(np.float32),
        "material": np.ones(ns + nd, np.uint32),
        "p0": np.concatenate([st[:, 3], dyn_p0]).astype(np.float32),
        "p1": np.concatenate([st[:, 4], dyn_p1]).astype(np.float32),
        "vstart": np.concatenate([np.zeros(ns, np.uint32), vstart]),
        "vcount": np.concatenate([np.zeros(ns, np.uint32), vc]),
        "verts": np.asarray(dyn_verts, np.float32).reshape(-1, 2),
    }


def _grid_bodies(rng, n, x0, y0, cols, cell=1.1, r_lo=0.2, r_hi=0.5):
    """n bodies (70 % polygons, 30 % circles, density 1) on a jittered grid, row by row upwards."""
    idx = np.arange(n)
    x = x0 + cell * (idx % cols + 0.5) + rng.uniform(-0.04, 0.04, n)
    y = y0 + cell * (idx // cols + 0.5) + rng.uniform(-0.04, 0.04, n)
    is_poly = rng.random(n) < 0.7
    npoly = int(is_poly.sum())
    rp, kp, verts = _polygons(rng, npoly, r_lo, r_hi)
    rc = rng.uniform(r_lo, r_hi, n - npoly)
    typ = np.where(is_poly, SHAPE_POLYGON, SHAPE_CIRCLE).astype(np.uint8)
    mass = np.zeros(n)
    mass[is_poly] = _polygon_areas(verts, kp) if npoly else []
    mass[~is_poly] = np.pi * rc.astype(np.float32).astype(np.float64) ** 2
    p0 = np.zeros(n)
    p0[~is_poly] = rc
    vcount = np.zeros(n, np.uint32)
    vcount[is_poly] = kp
    rot = np.minimum(rng.uniform(0.0, 2.0 * np.pi, n), 6.2831850)
    return typ, x, y, rot, mass, p0, np.zeros(n), vcount, verts


def batch_scene(i: int, seed: int = 4, n_min: int = 64, n_max: int = 4096) -> dict:
    """Scene i of the batch: body count log-uniform in [n_min, n_max]; layout by i mod 3 is a box
    pyramid, a random pile in an open box, or a funnel of static boxes above a floor."""
    rng = np.random.default_rng(seed + i)
    n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
    layout = i % 3
    if layout == 0:
        rows = max(1, int((math.sqrt(8 * n + 1) - 1) / 2))
        r = np.repeat(np.arange(rows), rows - np.arange(rows))
        c = np.concatenate([np.arange(rows - k) for k in range(rows)])
        count = rows - r
        x = (c - (count - 1) / 2.0) * 1.02 + rng.uniform(-0.01, 0.01, len(r))
        y = 0.5 + r * 1.001
        m = len(r)
        static = [(0.0, -1.0, 0.0, rows * 3.0 + 10.0, 2.0)]
        body = _scene_arrays(static, np.full(m, SHAPE_BOX), x, y, np.zeros(m), np.ones(m), np.ones(m), np.ones(m),
                             np.zeros(m, np.uint32), np.zeros((0, 2)))
        name = "pyramid"
    elif layout == 1:
        cols = max(4, int(math.sqrt(n)))
        width = cols * 1.1
        height = (n // cols + 2) * 1.1 * 1.5
        static = [(0.0, -1.0, 0.0, width + 4.0, 2.0), (-width / 2 - 1.0, height / 2, 0.0, 2.0, height),
                  (width / 2 + 1.0, height / 2, 0.0, 2.0, height)]
        body = _scene_arrays(static, *_grid_bodies(rng, n, -width / 2, 0.2, cols))
        name = "pile"
    else:
        cols = max(4, int(math.sqrt(n)))
        width = cols * 1.1
        mouth = 3.0
        slope = 0.5 * (width - mouth)
        base = 6.0
        rows_h = (n // cols + 2) * 1.1
        top = base + slope + 1.0
        length = slope * math.sqrt(2.0) + 1.0
        static = [(0.0, -1.0, 0.0, width + 8.0, 2.0),
                  (-(mouth / 2 + slope / 2) - 0.5, base + slope / 2, -math.pi / 4 + 2 * math.pi, length, 1.0),
                  (mouth / 2 + slope / 2 + 0.5, base + slope / 2, math.pi / 4, length, 1.0),
                  (-width / 2 - 1.5, (top + rows_h) / 2, 0.0, 1.0, top + rows_h + 2.0),
                  (width / 2 + 1.5, (top + rows_h) / 2, 0.0, 1.0, top + rows_h + 2.0)]
        body = _scene_arrays(static, *_grid_bodies(rng, n, -width / 2, top, cols))
        name = "funnel"
    return {"name": f"batch{i}_{name}", "materials": [(0.0, 0.6, 0.6)], "bodies": body,
            "settings": {"velocity_iterations": 8, "position_iterations": 3, "gravity_y": -9.81}}


def batch(num_scenes: int = 1024, seed: int = 4, n_min: int = 64, n_max: int = 4096) -> list:
    """The scene batch: independent scenes of log-uniform body counts (one world, many scenes)."""
    return [batch_scene(i, seed, n_min, n_max) for i in range(num_scenes)]
