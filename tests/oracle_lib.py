"""ctypes wrapper of oracle/liboracle.so (the CPU restatement used as the checker)."""
import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = ROOT / "oracle" / "liboracle.so"

MANIFOLD_DTYPE = np.dtype([("a", "<u4"), ("b", "<u4"), ("nx", "<f4"), ("ny", "<f4"), ("depth", "<f4"),
                           ("count", "<u4"), ("p0x", "<f4"), ("p0y", "<f4"), ("p1x", "<f4"), ("p1y", "<f4"),
                           ("id0", "<u4"), ("id1", "<u4"), ("jn0", "<f4"), ("jt0", "<f4"), ("jn1", "<f4"),
                           ("jt1", "<f4")])

EVENT_DTYPE = np.dtype([("self", "<u4"), ("other", "<u4"), ("self_body", "<u4"), ("other_body", "<u4"),
                        ("px", "<f4"), ("py", "<f4"), ("nx", "<f4"), ("ny", "<f4"), ("impulse", "<f4"),
                        ("callback", "<u4"), ("step", "<u4"), ("reserved", "<u4")])

_P = C.c_void_p
_SIGS = {
    "orc_new": (_P, []), "orc_free": (None, [_P]),
    "orc_set_settings": (None, [_P, C.c_float, C.c_float, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_float]),
    "orc_add_material": (C.c_uint32, [_P, C.c_float, C.c_float, C.c_float]),
    "orc_add_bodies": (C.c_int, [_P, C.c_uint32] + [_P] * 15),
    "orc_refresh": (None, [_P]), "orc_counts": (None, [_P, _P]),
    "orc_get_f32": (C.c_int, [_P, C.c_char_p, _P, C.c_uint32]),
    "orc_set_f32": (C.c_int, [_P, C.c_char_p, _P, C.c_uint32]),
    "orc_get_u32": (C.c_int, [_P, C.c_char_p, _P, C.c_uint32]),
    "orc_get_types": (C.c_int, [_P, _P, C.c_uint32]),
    "orc_get_verts": (C.c_int, [_P, _P, _P, C.c_uint32]),
    "orc_fast_cossin": (None, [_P, _P, _P, C.c_uint32]), "orc_wrap": (None, [_P, _P, C.c_uint32]),
    "orc_snapshot": (None, [_P]), "orc_integrate": (None, [_P, C.c_float]), "orc_wrap_cossin": (None, [_P]),
    "orc_poses": (None, [_P, C.c_int]), "orc_aabbs": (None, [_P, C.c_int]), "orc_ccd": (C.c_int, [_P, _P]),
    "orc_broad_bvh": (C.c_int64, [_P, _P, _P, C.c_uint64]), "orc_broad_brute": (C.c_int64, [_P, _P, _P, C.c_uint64]),
    "orc_narrow": (C.c_int64, [_P, _P, _P, C.c_uint64, _P, C.c_uint64]),
    "orc_set_prev_manifolds": (None, [_P, _P, C.c_uint64]), "orc_warm_start": (None, [_P, _P, C.c_uint64]),
    "orc_solve": (None, [_P, _P, C.c_uint64, _P, C.c_uint32, C.c_uint32, _P]),
    "orc_islands": (C.c_uint32, [_P, _P, C.c_uint64]),
    "orc_step": (None, [_P, C.c_float]), "orc_post_update": (None, [_P]),
    "orc_update": (None, [_P, C.c_float, C.c_uint32]),
    "orc_get_manifolds": (C.c_int64, [_P, _P, C.c_uint64]), "orc_errors": (None, [_P, _P]),
    "orc_queue_body": (None, [_P, C.c_uint32, C.c_uint32, C.c_float, C.c_float]), "orc_flush": (None, [_P]),
    "orc_pool_new": (_P, [C.c_uint32]), "orc_pool_free": (None, [_P]), "orc_pool_size": (C.c_uint32, [_P]),
    "orc_update_pool": (None, [_P, _P, C.c_float, C.c_uint32]),
    "orc_plan_waves": (C.c_int64, [_P, _P, C.c_uint64, C.c_uint32, _P, _P, _P, C.c_uint64]),
    "orc_solve_threaded": (None, [_P, _P, _P, C.c_uint64, C.c_uint32, C.c_uint32]),
    "orc_narrow_threaded": (C.c_int64, [_P, _P, _P, _P, C.c_uint64, _P, C.c_uint64]),
    "orc_set_callback": (None, [_P, C.c_uint32, C.c_uint32, C.c_float]),
    "orc_set_events_enabled": (None, [_P, C.c_int]),
    "orc_collision_events": (C.c_int64, [_P, _P, C.c_uint64, _P, C.c_uint64]),
    "orc_get_events": (C.c_int64, [_P, _P, C.c_uint64]),
    "orc_add_joints": (C.c_int, [_P, C.c_uint32, _P]),
    "orc_set_joint_iterations": (None, [_P, _P]),
    "orc_set_joint_order": (None, [_P, _P, C.c_uint32, _P, C.c_uint32, _P]),
    "orc_get_motor_impulses": (C.c_int, [_P, _P, C.c_uint32]),
    "orc_set_motor_impulses": (C.c_int, [_P, _P, C.c_uint32]),
    "orc_get_prev_manifolds": (C.c_int64, [_P, _P, C.c_uint64]),
    "orc_set_sleep": (None, [_P, C.c_float, C.c_float]),
    "orc_get_sleep": (None, [_P, _P, _P, C.c_uint32]),
    "orc_set_sleep_state": (None, [_P, _P, _P, C.c_uint32]),
    "orc_get_stash": (C.c_int64, [_P, _P, _P, C.c_uint64]),
    "orc_set_stash": (None, [_P, _P, _P, C.c_uint64]),
    "orc_wake_contacts": (C.c_int64, [_P, _P, C.c_uint64, _P, C.c_uint64]),
    "orc_update_sleeping": (C.c_int64, [_P, C.c_float, _P, C.c_uint64, _P]),
    "orc_mlp_forward": (None, [_P, _P, C.c_uint64, _P, C.c_int]),
    "orc_mlp_gradients": (C.c_double, [_P, _P, _P, C.c_uint64, _P, C.c_int]),
    "orc_mlp_loss": (C.c_double, [_P, _P, _P, _P, C.c_uint64]),
}
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C oracle`")
        h = C.CDLL(str(LIB_PATH))
        for k, (r, a) in _SIGS.items():
            f = getattr(h, k)
            f.restype = r
            f.argtypes = a
        _LIB = h
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


BODY_F32 = ["comX", "comY", "velX", "velY", "rot", "angVel", "mass", "invMass", "inertia", "invInertia", "cos",
            "sin", "localComX", "localComY", "boundRadius", "prevComX", "prevComY"]
COL_F32 = ["localOffX", "localOffY", "comOffX", "comOffY", "localRot", "wX", "wY", "wRot", "wCos", "wSin", "minX",
           "minY", "maxX", "maxY", "p0", "p1", "minExtent", "maxExtent", "prevWX", "prevWY"]


class World:
    """One oracle world (a single scene)."""

    def __init__(self, settings=None):
        self.h = lib().orc_new()
        s = dict(gx=0.0, gy=-9.81, v=6, p=3, ccd=1, fast=1.0, core=0.5)
        if settings:
            s.update(settings)
        lib().orc_set_settings(self.h, s["gx"], s["gy"], s["v"], s["p"], s["ccd"], s["fast"], s["core"])

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_free(self.h)
            self.h = None

    def add_material(self, e, us, ud):
        return lib().orc_add_material(self.h, e, us, ud)

    def add_bodies(self, b):
        n = len(b["type"])
        f = lambda k, dt=np.float32: np.ascontiguousarray(b[k], dt)
        arrs = [f("type", np.uint8), f("px"), f("py"), f("rot"), f("vx"), f("vy"), f("av"), f("mass"),
                f("material", np.uint32), f("p0"), f("p1"), f("vstart", np.uint32), f("vcount", np.uint32)]
        verts = np.ascontiguousarray(b["verts"] if len(b["verts"]) else np.zeros((1, 2)), np.float32)
        out = np.empty(n, np.uint32)
        rc = lib().orc_add_bodies(self.h, n, *[_p(a) for a in arrs], _p(verts), _p(out))
        return out, rc

    def refresh(self):
        lib().orc_refresh(self.h)

    def counts(self):
        out = np.zeros(10, np.uint32)
        lib().orc_counts(self.h, _p(out))
        keys = ["bodies", "static_bodies", "colliders", "static_colliders", "vertices", "materials", "manifolds",
                "pairs", "islands", "prev_manifolds"]
        return dict(zip(keys, (int(x) for x in out)))

    def get(self, name):
        c = self.counts()
        if name in COL_F32:
            n, dt = c["colliders"], np.float32
        elif name in BODY_F32:
            n, dt = c["bodies"], np.float32
        elif name in ("colBody", "colMaterial", "colSlot", "polyStart", "polyCount"):
            n, dt = c["colliders"], np.uint32
        else:
            n, dt = c["bodies"], np.uint32
        out = np.empty(n, dt)
        fn = lib().orc_get_f32 if dt == np.float32 else lib().orc_get_u32
        rc = fn(self.h, name.encode(), _p(out), n)
        if rc != 0:
            raise KeyError(name)
        return out

    def set(self, name, values):
        v = np.ascontiguousarray(values, np.float32)
        if lib().orc_set_f32(self.h, name.encode(), _p(v), len(v)) != 0:
            raise KeyError(name)

    def types(self):
        n = self.counts()["colliders"]
        out = np.empty(n, np.uint8)
        lib().orc_get_types(self.h, _p(out), n)
        return out

    def vertices(self):
        n = self.counts()["vertices"]
        v = np.empty((n, 2), np.float32)
        nn = np.empty((n, 2), np.float32)
        lib().orc_get_verts(self.h, _p(v), _p(nn), n)
        return v, nn

    # phases
    def snapshot(self): lib().orc_snapshot(self.h)
    def integrate(self, dt): lib().orc_integrate(self.h, dt)
    def wrap_cossin(self): lib().orc_wrap_cossin(self.h)
    def poses(self, all_=False): lib().orc_poses(self.h, int(all_))
    def aabbs(self, all_=False): lib().orc_aabbs(self.h, int(all_))
    def step(self, dt): lib().orc_step(self.h, dt)
    def post_update(self): lib().orc_post_update(self.h)
    def update(self, dt, steps=1): lib().orc_update(self.h, dt, steps)
    def queue_body(self, kind, body, a, b=0.0): lib().orc_queue_body(self.h, kind, body, a, b)
    def flush(self): lib().orc_flush(self.h)
    def update_threaded(self, pool, dt, steps=1): lib().orc_update_pool(self.h, pool.h, dt, steps)

    def plan_waves(self, m, workers):
        """The solver's wave plan: (flat order, pass starts, pass sizes), passes wave-major."""
        m = np.ascontiguousarray(m, MANIFOLD_DTYPE)
        cap = max(1024, 4 * workers)
        while True:
            flat = np.empty(max(len(m), 1), np.uint32)
            ps = np.empty(cap, np.uint32)
            sz = np.empty(cap, np.uint32)
            k = lib().orc_plan_waves(self.h, _p(m), len(m), workers, _p(flat), _p(ps), _p(sz), cap)
            if k >= 0:
                return flat[:len(m)], ps[:k], sz[:k]
            cap = -k

    def solve_threaded(self, pool, m, v_it=8, p_it=3):
        m = np.ascontiguousarray(m, MANIFOLD_DTYPE).copy()
        lib().orc_solve_threaded(self.h, pool.h, _p(m), len(m), v_it, p_it)
        return m

    def narrow_threaded(self, pool, pa, pb):
        pa = np.ascontiguousarray(pa, np.uint32)
        pb = np.ascontiguousarray(pb, np.uint32)
        cap = max(len(pa), 1)
        out = np.empty(cap, MANIFOLD_DTYPE)
        n = lib().orc_narrow_threaded(self.h, pool.h, _p(pa), _p(pb), len(pa), _p(out), cap)
        return out[:n]

    def ccd(self):
        t = np.empty(self.counts()["bodies"], np.float32)
        hit = lib().orc_ccd(self.h, _p(t))
        return bool(hit), t

    def _pairs(self, fn):
        cap = 1 << 16
        while True:
            a = np.empty(cap, np.uint32)
            b = np.empty(cap, np.uint32)
            n = fn(self.h, _p(a), _p(b), cap)
            if n >= 0:
                return a[:n], b[:n]
            cap = -n

    def broad_bvh(self): return self._pairs(lib().orc_broad_bvh)
    def broad_brute(self): return self._pairs(lib().orc_broad_brute)

    def narrow(self, pa, pb):
        pa = np.ascontiguousarray(pa, np.uint32)
        pb = np.ascontiguousarray(pb, np.uint32)
        out = np.empty(max(len(pa), 1), MANIFOLD_DTYPE)
        n = lib().orc_narrow(self.h, _p(pa), _p(pb), len(pa), _p(out), len(out))
        return out[:n]

    def set_prev_manifolds(self, m):
        m = np.ascontiguousarray(m, MANIFOLD_DTYPE)
        lib().orc_set_prev_manifolds(self.h, _p(m), len(m))

    def warm_start(self, m):
        m = np.ascontiguousarray(m, MANIFOLD_DTYPE).copy()
        lib().orc_warm_start(self.h, _p(m), len(m))
        return m

    def solve(self, m, order=None, v_it=8, p_it=3, residual=False):
        m = np.ascontiguousarray(m, MANIFOLD_DTYPE).copy()
        o = None if order is None else np.ascontiguousarray(order, np.uint32)
        r = np.zeros(max(v_it, 1), np.float32) if residual else None
        lib().orc_solve(self.h, _p(m), len(m), _p(o), v_it, p_it, _p(r))
        return (m, r) if residual else m

    def islands(self, m):
        m = np.ascontiguousarray(m, MANIFOLD_DTYPE)
        k = lib().orc_islands(self.h, _p(m), len(m))
        return self.get("islandId"), k

    def manifolds(self):
        cap = max(self.counts()["manifolds"], 1)
        out = np.empty(cap, MANIFOLD_DTYPE)
        n = lib().orc_get_manifolds(self.h, _p(out), cap)
        return out[:n]

    def set_callback(self, collider, callback_id, threshold=0.0):
        lib().orc_set_callback(self.h, collider, callback_id, threshold)

    def set_events_enabled(self, on):
        lib().orc_set_events_enabled(self.h, int(on))

    def collision_events(self, m):
        """Events of manifolds m on the current (pre-solve) state, in manifold order."""
        m = np.ascontiguousarray(m, MANIFOLD_DTYPE)
        cap = 2 * max(len(m), 1)
        out = np.empty(cap, EVENT_DTYPE)
        n = lib().orc_collision_events(self.h, _p(m), len(m), _p(out), cap)
        return out[:n]

    def events(self):
        """Events recorded since the last update call."""
        cap = 1 << 12
        while True:
            out = np.empty(cap, EVENT_DTYPE)
            n = lib().orc_get_events(self.h, _p(out), cap)
            if n >= 0:
                return out[:n]
            cap = -n

    def set_sleep(self, speed=0.1, time=0.5):
        lib().orc_set_sleep(self.h, speed, time)

    def sleep_state(self):
        n = self.counts()["bodies"]
        ids = np.empty(n, np.uint32)
        t = np.empty(n, np.float32)
        lib().orc_get_sleep(self.h, _p(ids), _p(t), n)
        return ids, t

    def set_sleep_state(self, ids, timers):
        ids = np.ascontiguousarray(ids, np.uint32)
        t = np.ascontiguousarray(timers, np.float32)
        lib().orc_set_sleep_state(self.h, _p(ids), _p(t), len(ids))

    def stash(self):
        cap = 1 << 12
        while True:
            m = np.empty(cap, MANIFOLD_DTYPE)
            isl = np.empty(cap, np.uint32)
            n = lib().orc_get_stash(self.h, _p(m), _p(isl), cap)
            if n >= 0:
                return m[:n], isl[:n]
            cap = -n

    def set_stash(self, m, island):
        m = np.ascontiguousarray(m, MANIFOLD_DTYPE)
        isl = np.ascontiguousarray(island, np.uint32)
        lib().orc_set_stash(self.h, _p(m), _p(isl), len(m))

    def wake_contacts(self, m):
        m = np.ascontiguousarray(m, MANIFOLD_DTYPE)
        cap = len(m) + len(self.stash()[0]) + 1
        out = np.empty(cap, MANIFOLD_DTYPE)
        n = lib().orc_wake_contacts(self.h, _p(m), len(m), _p(out), cap)
        return out[:n]

    def update_sleeping(self, dt, m):
        m = np.ascontiguousarray(m, MANIFOLD_DTYPE)
        out = np.empty(max(len(m), 1), MANIFOLD_DTYPE)
        n = lib().orc_update_sleeping(self.h, dt, _p(m), len(m), _p(out))
        return out[:n]

    def add_joints(self, joints):
        """Joints as a JOINT_DTYPE array (agx_joint_desc layout) or a list of dicts."""
        import importlib
        agx = importlib.import_module("agx")
        j = joints if isinstance(joints, np.ndarray) else agx.joint_array(joints)
        j = np.ascontiguousarray(j, agx.JOINT_DTYPE)
        if lib().orc_add_joints(self.h, len(j), _p(j)) != 0:
            raise ValueError("joint names a missing body")
        self.n_joints = getattr(self, "n_joints", 0) + len(j)

    def set_joint_iterations(self, it):
        it = np.ascontiguousarray(it, np.uint32)
        lib().orc_set_joint_iterations(self.h, _p(it))

    def set_joint_order(self, order, starts, types):
        order = np.ascontiguousarray(order, np.uint32)
        starts = np.ascontiguousarray(starts, np.uint32)
        types = np.ascontiguousarray(types, np.uint32)
        lib().orc_set_joint_order(self.h, _p(order), len(order), _p(starts), len(starts) - 1, _p(types))

    def motor_impulses(self):
        n = getattr(self, "n_joints", 0)
        out = np.empty(max(n, 1), np.float32)
        lib().orc_get_motor_impulses(self.h, _p(out), n)
        return out[:n]

    def set_motor_impulses(self, v):
        v = np.ascontiguousarray(v, np.float32)
        if lib().orc_set_motor_impulses(self.h, _p(v), len(v)) != 0:
            raise ValueError("joint count")

    def prev_manifolds(self):
        cap = max(self.counts()["prev_manifolds"], 1)
        out = np.empty(cap, MANIFOLD_DTYPE)
        n = lib().orc_get_prev_manifolds(self.h, _p(out), cap)
        return out[:n]

    def errors(self):
        out = np.zeros(2, np.float32)
        lib().orc_errors(self.h, _p(out))
        return out


def fast_cossin(x):
    x = np.ascontiguousarray(x, np.float32)
    c = np.empty_like(x)
    s = np.empty_like(x)
    lib().orc_fast_cossin(_p(x), _p(c), _p(s), len(x))
    return c, s


def wrap(x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().orc_wrap(_p(x), _p(y), len(x))
    return y


def world_from_scene(scene, settings=None):
    w = World(settings)
    for m in scene.get("materials", []):
        w.add_material(*m)
    idx, rc = w.add_bodies(scene["bodies"])
    if scene.get("joints"):
        w.add_joints(scene["joints"])
    w.refresh()
    return w


# ---- network of the learned warm start (oracle/mlp.cpp)
MLP_PARAMS = 19970


def mlp_forward(params, x, rounding=True):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros((x.shape[0], 2), np.float32)
    lib().orc_mlp_forward(_p(params), _p(x), x.shape[0], _p(y), 1 if rounding else 0)
    return y


def mlp_gradients(params, x, t, rounding=True):
    """(loss, gradient in host parameter layout) of the mean squared error over the rows."""
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    g = np.zeros(MLP_PARAMS, np.float32)
    loss = lib().orc_mlp_gradients(_p(params), _p(x), _p(t), x.shape[0], _p(g), 1 if rounding else 0)
    return float(loss), g


def mlp_loss(params, x, t, delta=None):
    """Exact (double, unrounded) loss with parameters params + delta."""
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    d = None if delta is None else np.ascontiguousarray(delta, np.float64)
    return float(lib().orc_mlp_loss(_p(params), _p(d), _p(x), _p(t), x.shape[0]))


class Pool:
    """Pinned worker threads plus the calling thread for the oracle's threaded step."""

    def __init__(self, participants):
        self.h = lib().orc_pool_new(int(participants))

    @property
    def size(self):
        return int(lib().orc_pool_size(self.h))

    def close(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_pool_free(self.h)
        self.h = None

    def __del__(self):
        self.close()
