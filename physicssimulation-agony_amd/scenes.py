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

SHAPE_CIRCLE, SHAPE_BOX, SHAPE_POLYGON = 0, 1, 2


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
    """Bodies uniform in [-size/2, size/2]^2 with random orientation, bullets, static obstacles."""
    rng = np.random.default_rng(seed)
    acc = _Acc()
    for _ in range(n_static):
        x, y = rng.uniform(-size / 2, size / 2, 2)
        if rng.random() < 0.5:
            acc.add(SHAPE_BOX, float(x), float(y), rot=float(rng.uniform(0, 2 * np.pi)),
                    p0=float(rng.uniform(2, 8)), p1=float(rng.uniform(2, 8)))
        else:
            acc.add(SHAPE_POLYGON, float(x), float(y), rot=float(rng.uniform(0, 2 * np.pi)),
                    verts=polygon_outline(rng, int(rng.integers(3, 9)), float(rng.uniform(1.5, 4.0))))
    radii = np.clip(rng.lognormal(math.log(0.3), 0.3, n), 0.1, 0.9)
    bullets = rng.random(n) < bullet_fraction
    radii = np.where(bullets, radii * 0.25, radii)
    xs = rng.uniform(-size / 2, size / 2, n)
    ys = rng.uniform(-size / 2, size / 2, n)
    rots = np.minimum(rng.uniform(0, 2 * np.pi, n), 6.2831850)
    dirs = rng.uniform(0, 2 * np.pi, n)
    kinds = rng.random(n) < 0.7
    dens = rng.uniform(0.5, 2.0, n)
    for i in range(n):
        vx = bullet_speed * math.cos(dirs[i]) if bullets[i] else 0.0
        vy = bullet_speed * math.sin(dirs[i]) if bullets[i] else 0.0
        r = float(np.float32(radii[i]))
        if kinds[i] and not bullets[i]:
            v = polygon_outline(rng, int(rng.integers(3, 9)), r)
            acc.add(SHAPE_POLYGON, float(xs[i]), float(ys[i]), rot=float(rots[i]), vx=vx, vy=vy,
                    mass=float(np.float32(dens[i] * polygon_area(v))), verts=v)
        else:
            acc.add(SHAPE_CIRCLE, float(xs[i]), float(ys[i]), rot=float(rots[i]), vx=vx, vy=vy,
                    mass=float(np.float32(dens[i] * math.pi * r * r)), p0=r)
    return {"name": name, "materials": [(0.0, 0.6, 0.6)], "bodies": acc.bodies(),
            "settings": {"gravity_y": 0.0, "collide_only": 1}}


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
