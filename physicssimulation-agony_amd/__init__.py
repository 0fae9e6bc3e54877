"""MI355X collide-and-solve path of PS_AGONY2D: Python binding over the C ABI in include/agx.h.

The compute path is libagx.so (hand-written gfx950 HIP kernels); this module only moves
arrays across the C ABI. It fails loudly when libagx.so is missing: there is no CPU fallback.

Names mirror the reference's public API where one exists:
``World.update`` is ``Simulation::update`` (PS_AGONY2D/Simulation.cpp:32-74), ``World.step``
runs whole physics steps (``Simulation::physicsStep``, Simulation.cpp:255-351), and
``Builder.create_bodies`` covers ``ObjectAPI::createCircle/createBox/createPolygon``
(SimulationAPIs/ObjectAPI.h:34-37) for many bodies at once.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libagx.so"
MAX_SWEEPS = 256   # AGX_MAX_SWEEPS

SHAPE_CIRCLE, SHAPE_BOX, SHAPE_POLYGON, SHAPE_CAPSULE = 0, 1, 2, 3

(PHASE_INTEGRATE, PHASE_POSE, PHASE_CCD, PHASE_BROAD, PHASE_NARROW, PHASE_SOLVE,
 PHASE_ISLANDS, PHASE_FINISH) = range(8)

BODY_FIELDS = ["com_x", "com_y", "rotation", "vel_x", "vel_y", "ang_vel", "cos", "sin",
               "inv_mass", "inv_inertia", "mass", "inertia", "local_com_x", "local_com_y"]
COLLIDER_F32 = ["world_x", "world_y", "world_rot", "cos", "sin", "aabb_min_x", "aabb_min_y",
                "aabb_max_x", "aabb_max_y", "prev_x", "prev_y", "com_off_x", "com_off_y",
                "local_rot", "p0", "p1", "min_extent", "max_extent"]
COLLIDER_U32 = ["body", "type", "material", "slot", "poly_start", "poly_count", "scene"]

MANIFOLD_DTYPE = np.dtype([("a", "<u4"), ("b", "<u4"), ("nx", "<f4"), ("ny", "<f4"), ("depth", "<f4"),
                           ("count", "<u4"), ("p0x", "<f4"), ("p0y", "<f4"), ("p1x", "<f4"), ("p1y", "<f4"),
                           ("id0", "<u4"), ("id1", "<u4"), ("jn0", "<f4"), ("jt0", "<f4"), ("jn1", "<f4"),
                           ("jt1", "<f4")])
assert MANIFOLD_DTYPE.itemsize == 64
# agx_collision_event (Types/CollisionEvent.h:18-27 with dense indices for handles)
EVENT_DTYPE = np.dtype([("self", "<u4"), ("other", "<u4"), ("self_body", "<u4"), ("other_body", "<u4"),
                        ("px", "<f4"), ("py", "<f4"), ("nx", "<f4"), ("ny", "<f4"), ("impulse", "<f4"),
                        ("callback", "<u4"), ("step", "<u4"), ("reserved", "<u4")])
assert EVENT_DTYPE.itemsize == 48
# agx_joint_desc (ConstraintSystems/*ConstraintSystem.h CreateParams)
JOINT_DTYPE = np.dtype([("type", "<u4"), ("body_a", "<u4"), ("body_b", "<u4"), ("enable_limit", "<u4"),
                        ("anchor_a_x", "<f4"), ("anchor_a_y", "<f4"), ("anchor_b_x", "<f4"), ("anchor_b_y", "<f4"),
                        ("axis_x", "<f4"), ("axis_y", "<f4"), ("rest_length", "<f4"), ("stiffness", "<f4"),
                        ("damping", "<f4"), ("lower_limit", "<f4"), ("upper_limit", "<f4"), ("target_speed", "<f4"),
                        ("max_torque", "<f4")])
assert JOINT_DTYPE.itemsize == 68
JOINT_SPRING, JOINT_REVOLUTE, JOINT_PRISMATIC, JOINT_SLIDER, JOINT_MOTOR = range(5)


def joint_array(joints) -> np.ndarray:
    """A joint list (dicts with JOINT_DTYPE field names; missing fields 0, axis (1, 0)) as records."""
    out = np.zeros(len(joints), JOINT_DTYPE)
    out["axis_x"] = 1.0
    for i, j in enumerate(joints):
        for k, v in j.items():
            out[k][i] = v
    return out


class AgxError(RuntimeError):
    """A failed call across the C ABI (carries the library's error code and message)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"agx error {code}: {message}")
        self.code = code


class Settings(C.Structure):
    """agx_settings: Simulation::SimulationSettings (Simulation.h:50-76) and friends."""
    _fields_ = [("gravity_x", C.c_float), ("gravity_y", C.c_float),
                ("velocity_iterations", C.c_uint32), ("position_iterations", C.c_uint32),
                ("ccd_enabled", C.c_uint32), ("ccd_fast_factor", C.c_float), ("ccd_core_fraction", C.c_float),
                ("update_interval", C.c_float), ("max_delta_per_update", C.c_float), ("time_scale", C.c_float),
                ("track_errors", C.c_uint32), ("track_residual", C.c_uint32), ("residual_target", C.c_float),
                ("collide_only", C.c_uint32), ("solve_per_scene", C.c_uint32),
                ("collision_events", C.c_uint32), ("event_capacity", C.c_uint32),
                ("sleep_speed_threshold", C.c_float), ("sleep_time_threshold", C.c_float),
                ("joint_iterations", C.c_uint32 * 5)]


class Stats(C.Structure):
    _fields_ = [("pairs", C.c_uint64), ("manifolds", C.c_uint64), ("contacts", C.c_uint64),
                ("colours", C.c_uint32), ("islands", C.c_uint32), ("ccd_hits", C.c_uint32),
                ("velocity_sweeps", C.c_uint32), ("overflow", C.c_uint32), ("learned_warm_start", C.c_uint32),
                ("velocity_error", C.c_float), ("position_error", C.c_float), ("residual", C.c_float * MAX_SWEEPS),
                ("solved_steps", C.c_uint64), ("total_velocity_sweeps", C.c_uint64),
                ("total_body_sweeps", C.c_uint64), ("mean_velocity_sweeps", C.c_float),
                ("solved_per_scene", C.c_uint32), ("converged_scenes", C.c_uint32),
                ("converged_bodies", C.c_uint64), ("total_converged_bodies", C.c_uint64)]


class Gate(C.Structure):
    """agx_gate: residual gate of the learned warm start."""
    _fields_ = [("cost_cached", C.c_float), ("cost_learned", C.c_float), ("use_learned", C.c_uint32),
                ("since_probe", C.c_uint32), ("steps_cached", C.c_uint32), ("steps_learned", C.c_uint32),
                ("period", C.c_uint32), ("probing", C.c_uint32)]


class Tuning(C.Structure):
    """agx_tuning: co-resident blocks per CU of the dataflow colouring and sweep kernels, and the
    heaviest scenes whole blocks take in the per-scene solve."""
    _fields_ = [("solve_blocks_per_cu", C.c_uint32), ("colour_blocks_per_cu", C.c_uint32),
                ("ms_per_step", C.c_float), ("from_cache", C.c_uint32),
                ("max_solve_blocks_per_cu", C.c_uint32), ("max_colour_blocks_per_cu", C.c_uint32),
                ("scene_whole", C.c_uint32)]


LEARN_OFF, LEARN_PREDICT, LEARN_COLLECT, LEARN_UNGATED, LEARN_NEW_CONTACTS = 0, 1, 2, 4, 8
SET_POSITION, SET_ROTATION, SET_VELOCITY, SET_ANGULAR_VELOCITY, SET_MASS, SET_INERTIA = range(6)
MLP_INPUTS, MLP_HIDDEN, MLP_OUTPUTS, MLP_PARAMS, MLP_BATCH_MAX = 24, 128, 2, 19970, 8192


class Counts(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("scenes", "bodies", "static_bodies", "colliders", "static_colliders",
                                          "vertices", "materials", "reserved")]


_LIB = None

# Every exported function of include/agx.h with its ctypes signature.
_P = C.c_void_p
_SIGS = {
    "agx_settings_default": (None, [C.POINTER(Settings)]),
    "agx_last_error": (C.c_char_p, []),
    "agx_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "agx_builder_create": (C.c_int, [C.c_uint32, C.POINTER(_P)]),
    "agx_builder_destroy": (C.c_int, [_P]),
    "agx_builder_add_material": (C.c_int, [_P, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_uint32)]),
    "agx_builder_add_bodies": (C.c_int, [_P, C.c_uint32, C.c_uint32] + [_P] * 15),
    "agx_builder_create_body": (C.c_int, [_P, C.c_uint32] + [C.c_float] * 10 + [_P]),
    "agx_builder_add_collider": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint8] + [C.c_float] * 3 + [C.c_uint32]
                                 + [C.c_float] * 2 + [_P]),
    "agx_world_add_bodies": (C.c_int, [_P, C.c_uint32, C.c_uint32] + [_P] * 15),
    "agx_world_remove_bodies": (C.c_int, [_P, C.c_uint32, _P]),
    "agx_builder_remove_body": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
    "agx_builder_scene_counts": (C.c_int, [_P, C.c_uint32, C.POINTER(Counts)]),
    "agx_builder_read_body_f32": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_uint32]),
    "agx_builder_read_collider": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, _P, C.c_uint32]),
    "agx_builder_read_vertices": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_uint32]),
    "agx_world_create": (C.c_int, [_P, C.POINTER(Settings), C.POINTER(_P)]),
    "agx_world_destroy": (C.c_int, [_P]),
    "agx_world_last_error": (C.c_char_p, [_P]),
    "agx_world_set_settings": (C.c_int, [_P, C.POINTER(Settings)]),
    "agx_world_get_settings": (C.c_int, [_P, C.POINTER(Settings)]),
    "agx_world_counts": (C.c_int, [_P, C.POINTER(Counts)]),
    "agx_world_update": (C.c_int, [_P, C.c_float, C.POINTER(C.c_uint32)]),
    "agx_world_step": (C.c_int, [_P, C.c_float, C.c_uint32]),
    "agx_world_run_phase": (C.c_int, [_P, C.c_uint32, C.c_float]),
    "agx_world_post_update": (C.c_int, [_P]),
    "agx_world_sync": (C.c_int, [_P]),
    "agx_world_step_graph": (C.c_int, [_P, C.c_float, C.c_uint32]),
    "agx_world_read_body_f32": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32]),
    "agx_world_write_body_f32": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32]),
    "agx_world_read_collider_f32": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32]),
    "agx_world_read_collider_u32": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32]),
    "agx_world_read_pairs": (C.c_int, [_P, _P, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    "agx_world_queue_body": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_float, C.c_float]),
    "agx_world_flush": (C.c_int, [_P]),
    "agx_world_narrowphase": (C.c_int, [_P, _P, _P, C.c_uint64, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    "agx_world_solve_contacts": (C.c_int, [_P, _P, C.c_uint64, _P, _P]),
    "agx_world_read_broadphase_order": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "agx_world_read_manifolds": (C.c_int, [_P, C.c_uint32, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    "agx_world_write_manifolds": (C.c_int, [_P, C.c_uint32, _P, C.c_uint64]),
    "agx_world_read_solve_order": (C.c_int, [_P, _P, C.c_uint64, _P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "agx_world_read_islands": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "agx_world_read_stats": (C.c_int, [_P, C.POINTER(Stats)]),
    "agx_world_read_scene_residual": (C.c_int, [_P, _P, C.c_uint32]),
    "agx_world_read_scene_sweeps": (C.c_int, [_P, _P, C.c_uint32]),
    "agx_broadphase_pairs": (C.c_int, [_P, _P, _P, _P, C.c_uint32, _P, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    "agx_broadphase_pairs_noalloc": (C.c_int, [_P, _P, _P, _P, C.c_uint32, C.c_uint32, _P, _P, C.c_uint64,
                                               C.POINTER(C.c_uint64)]),
    "agx_world_set_profiling": (C.c_int, [_P, C.c_uint32]),
    "agx_world_read_profile": (C.c_int, [_P, _P, _P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "agx_profile_segment_name": (C.c_char_p, [C.c_uint32]),
    "agx_mlp_create": (C.c_int, [C.c_uint32, C.c_uint64, C.POINTER(_P)]),
    "agx_mlp_destroy": (C.c_int, [_P]),
    "agx_mlp_get_params": (C.c_int, [_P, _P, C.c_uint64]),
    "agx_mlp_set_params": (C.c_int, [_P, _P, C.c_uint64]),
    "agx_mlp_forward": (C.c_int, [_P, _P, C.c_uint64, _P]),
    "agx_mlp_gradients": (C.c_int, [_P, _P, _P, C.c_uint64, _P, C.POINTER(C.c_float)]),
    "agx_mlp_add_samples": (C.c_int, [_P, _P, _P, C.c_uint64]),
    "agx_mlp_sample_count": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "agx_mlp_read_samples": (C.c_int, [_P, _P, _P, C.c_uint64, C.c_uint64]),
    "agx_mlp_clear_samples": (C.c_int, [_P]),
    "agx_mlp_train": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, C.POINTER(C.c_float)]),
    "agx_world_set_learned": (C.c_int, [_P, _P, C.c_uint32]),
    "agx_gate_update": (C.c_int, [C.POINTER(Gate), C.c_uint32, C.c_float]),
    "agx_world_read_gate": (C.c_int, [_P, C.POINTER(Gate)]),
    "agx_world_reset_gate": (C.c_int, [_P]),
    "agx_world_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "agx_world_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "agx_world_autotune": (C.c_int, [_P, C.c_float, C.c_uint32, C.POINTER(Tuning)]),
    "agx_world_set_collider_callback": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_float]),
    "agx_world_snapshot_size": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "agx_world_save": (C.c_int, [_P, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    "agx_world_load": (C.c_int, [_P, _P, C.c_uint64]),
    "agx_world_read_sleep": (C.c_int, [_P, _P, _P, C.c_uint32]),
    "agx_world_read_stash": (C.c_int, [_P, _P, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    "agx_builder_add_joints": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P]),
    "agx_world_read_joint_order": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, _P, C.POINTER(C.c_uint32),
                                             C.POINTER(C.c_uint32)]),
    "agx_world_read_joints": (C.c_int, [_P, _P, _P, C.c_uint32]),
    "agx_world_read_events": (C.c_int, [_P, _P, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
}
EXPORTED = tuple(_SIGS)


def lib() -> C.CDLL:
    """Loads libagx.so (built in this directory by __graft_entry__.build / make)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C {PKG_DIR}` or __graft_entry__.build()")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


# Live worlds and networks, released newest first when the interpreter exits (before the HIP
# runtime's own exit handlers), so no device object is freed from a finaliser after teardown began.
_LIVE: "weakref.WeakValueDictionary[int, object]" = weakref.WeakValueDictionary()
_NEXT_ID = [0]


def _track(obj) -> None:
    _NEXT_ID[0] += 1
    _LIVE[_NEXT_ID[0]] = obj


@atexit.register
def _close_all() -> None:
    for key in sorted(list(_LIVE.keys()), reverse=True):
        obj = _LIVE.get(key)
        if obj is not None:
            obj.close()


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def default_settings(**overrides) -> Settings:
    s = Settings()
    lib().agx_settings_default(C.byref(s))
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


def _check(rc: int, world=None):
    if rc != 0:
        msg = (lib().agx_world_last_error(world) if world else lib().agx_last_error()).decode()
        raise AgxError(rc, msg)


class Builder:
    """Scene builder (ObjectAPI creation + Simulation::createMaterial), S independent scenes."""

    def __init__(self, num_scenes: int = 1):
        self._h = C.c_void_p()
        _check(lib().agx_builder_create(num_scenes, C.byref(self._h)))
        self.num_scenes = num_scenes

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.agx_builder_destroy(self._h)
            self._h = None

    def create_material(self, elasticity: float, static_friction: float, dynamic_friction: float) -> int:
        out = C.c_uint32()
        _check(lib().agx_builder_add_material(self._h, elasticity, static_friction, dynamic_friction, C.byref(out)))
        return out.value

    def create_bodies(self, scene: int, bodies: dict) -> np.ndarray:
        """Creates bodies from the arrays of a scene description (see scenes.py); returns dense indices."""
        b = normalise_bodies(bodies)
        n = len(b["type"])
        out = np.empty(n, np.uint32)
        rc = lib().agx_builder_add_bodies(
            self._h, scene, n, _ptr(b["type"]), _ptr(b["px"]), _ptr(b["py"]), _ptr(b["rot"]), _ptr(b["vx"]),
            _ptr(b["vy"]), _ptr(b["av"]), _ptr(b["mass"]), _ptr(b["material"]), _ptr(b["p0"]), _ptr(b["p1"]),
            _ptr(b["vstart"]), _ptr(b["vcount"]), _ptr(b["verts"]), _ptr(out))
        if rc < 0:
            _check(rc)
        self.rejected = rc
        return out

    def create_body(self, scene: int, x: float, y: float, mass: float, inertia: float, rot: float = 0.0,
                    vel=(0.0, 0.0), ang_vel: float = 0.0, com=(0.0, 0.0)) -> int:
        """A body without colliders (ObjectAPI::createBody): mass properties as given."""
        out = C.c_uint32()
        _check(lib().agx_builder_create_body(self._h, scene, x, y, rot, vel[0], vel[1], ang_vel, mass, inertia, com[0],
                                             com[1], C.byref(out)))
        return out.value

    def add_collider(self, scene: int, body: int, shape: int, p0: float, p1: float = 0.0, offset=(0.0, 0.0),
                     local_rot: float = 0.0, material: int = 0) -> int:
        """A circle, box or capsule collider on a body (ObjectAPI::create*Collider); its dense index."""
        out = C.c_uint32()
        _check(lib().agx_builder_add_collider(self._h, scene, body, shape, offset[0], offset[1], local_rot, material, p0, p1,
                                              C.byref(out)))
        return out.value

    def add_joints(self, scene: int, joints) -> np.ndarray:
        """Joints (a JOINT_DTYPE array or a list of dicts) between bodies of a scene."""
        j = joints if isinstance(joints, np.ndarray) else joint_array(joints)
        j = np.ascontiguousarray(j, JOINT_DTYPE)
        out = np.empty(max(len(j), 1), np.uint32)
        _check(lib().agx_builder_add_joints(self._h, scene, len(j), _ptr(j), _ptr(out)))
        return out[:len(j)]

    def counts(self, scene: int = 0) -> Counts:
        c = Counts()
        _check(lib().agx_builder_scene_counts(self._h, scene, C.byref(c)))
        return c

    def body(self, scene: int, field: str) -> np.ndarray:
        n = self.counts(scene).bodies
        out = np.empty(n, np.float32)
        _check(lib().agx_builder_read_body_f32(self._h, scene, BODY_FIELDS.index(field), _ptr(out), n))
        return out

    def collider(self, scene: int, field: str) -> np.ndarray:
        n = self.counts(scene).colliders
        as_u32 = field in COLLIDER_U32
        out = np.empty(n, np.uint32 if as_u32 else np.float32)
        idx = COLLIDER_U32.index(field) if as_u32 else COLLIDER_F32.index(field)
        _check(lib().agx_builder_read_collider(self._h, scene, idx, int(as_u32), _ptr(out), n))
        return out

    def vertices(self, scene: int = 0):
        n = self.counts(scene).vertices
        v = np.empty((n, 2), np.float32)
        nn = np.empty((n, 2), np.float32)
        _check(lib().agx_builder_read_vertices(self._h, scene, _ptr(v), _ptr(nn), n))
        return v, nn


def normalise_bodies(bodies: dict) -> dict:
    """Contiguous typed arrays for the body-creation call."""
    n = len(bodies["type"])
    out = {"type": np.ascontiguousarray(bodies["type"], np.uint8)}
    for k in ("px", "py", "rot", "vx", "vy", "av", "mass", "p0", "p1"):
        out[k] = np.ascontiguousarray(bodies.get(k, np.zeros(n)), np.float32)
    out["material"] = np.ascontiguousarray(bodies.get("material", np.zeros(n)), np.uint32)
    out["vstart"] = np.ascontiguousarray(bodies.get("vstart", np.zeros(n)), np.uint32)
    out["vcount"] = np.ascontiguousarray(bodies.get("vcount", np.zeros(n)), np.uint32)
    verts = bodies.get("verts")
    out["verts"] = np.ascontiguousarray(verts if verts is not None and len(verts) else np.zeros((1, 2)), np.float32)
    return out


class World:
    """A device world of one or more scenes (Simulation, Simulation.h:36-255)."""

    def __init__(self, builder: Builder, settings: Settings | None = None):
        self._h = C.c_void_p()
        s = settings if settings is not None else default_settings()
        _check(lib().agx_world_create(builder._h, C.byref(s), C.byref(self._h)))
        self.counts = self._read_counts()
        _track(self)

    def _read_counts(self) -> Counts:
        c = Counts()
        _check(lib().agx_world_counts(self._h, C.byref(c)), self._h)
        return c

    def close(self):
        if getattr(self, "_h", None):
            lib().agx_world_destroy(self._h)
            self._h = None

    def __del__(self):
        if _LIB is not None:
            self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- settings / stepping
    def settings(self) -> Settings:
        s = Settings()
        _check(lib().agx_world_get_settings(self._h, C.byref(s)), self._h)
        return s

    def set_settings(self, s: Settings):
        _check(lib().agx_world_set_settings(self._h, C.byref(s)), self._h)

    def update(self, dt: float) -> int:
        n = C.c_uint32()
        _check(lib().agx_world_update(self._h, dt, C.byref(n)), self._h)
        return n.value

    def step(self, dt: float, n_steps: int = 1):
        _check(lib().agx_world_step(self._h, dt, n_steps), self._h)

    def step_graph(self, dt: float, n_steps: int = 1):
        """As step(), with every step after the first replayed from a captured HIP graph."""
        _check(lib().agx_world_step_graph(self._h, dt, n_steps), self._h)

    def run_phase(self, phase: int, dt: float):
        _check(lib().agx_world_run_phase(self._h, phase, dt), self._h)

    def post_update(self):
        _check(lib().agx_world_post_update(self._h), self._h)

    def sync(self):
        _check(lib().agx_world_sync(self._h), self._h)

    # -- read-outs
    def body(self, field: str) -> np.ndarray:
        out = np.empty(self.counts.bodies, np.float32)
        _check(lib().agx_world_read_body_f32(self._h, BODY_FIELDS.index(field), _ptr(out), len(out)), self._h)
        return out

    def queue_body(self, kind: int, body: int, a: float, b: float = 0.0):
        """Deferred modification (SET_*), applied at the next step or flush."""
        _check(lib().agx_world_queue_body(self._h, kind, body, a, b), self._h)

    def flush(self):
        _check(lib().agx_world_flush(self._h), self._h)

    def narrowphase(self, a, b) -> np.ndarray:
        """Manifolds of the given collider pairs at the current poses (warm-started)."""
        a = np.ascontiguousarray(a, np.uint32)
        b = np.ascontiguousarray(b, np.uint32)
        out = np.empty(max(len(a), 1), MANIFOLD_DTYPE)
        cnt = C.c_uint64()
        _check(lib().agx_world_narrowphase(self._h, _ptr(a), _ptr(b), len(a), _ptr(out), len(out), C.byref(cnt)), self._h)
        return out[:cnt.value]

    def create_bodies(self, scene: int, bodies: dict) -> np.ndarray:
        """Creates bodies after the upload (the world is rebuilt around them); returns their world
        dense indices at creation (UINT32_MAX for a rejected body)."""
        b = normalise_bodies(bodies)
        n = len(b["type"])
        out = np.empty(max(n, 1), np.uint32)
        rc = lib().agx_world_add_bodies(
            self._h, scene, n, _ptr(b["type"]), _ptr(b["px"]), _ptr(b["py"]), _ptr(b["rot"]), _ptr(b["vx"]),
            _ptr(b["vy"]), _ptr(b["av"]), _ptr(b["mass"]), _ptr(b["material"]), _ptr(b["p0"]), _ptr(b["p1"]),
            _ptr(b["vstart"]), _ptr(b["vcount"]), _ptr(b["verts"]), _ptr(out))
        if rc < 0:
            _check(rc, self._h)
        self.counts = self._read_counts()
        return out[:n]

    def remove_bodies(self, bodies):
        """Destroys bodies (world dense indices) with their colliders and joints; the world is rebuilt."""
        b = np.ascontiguousarray(bodies, np.uint32)
        _check(lib().agx_world_remove_bodies(self._h, len(b), _ptr(b)), self._h)
        self.counts = self._read_counts()

    def solve_contacts(self, manifolds, order=None) -> np.ndarray:
        """Solves the given manifolds against the current bodies (which it updates) in the given
        sequential order (None: the world's own colouring); returns the solved manifolds."""
        m = np.ascontiguousarray(manifolds, MANIFOLD_DTYPE)
        o = None if order is None else np.ascontiguousarray(order, np.uint32)
        out = np.empty(max(len(m), 1), MANIFOLD_DTYPE)
        _check(lib().agx_world_solve_contacts(self._h, _ptr(m), len(m), _ptr(o), _ptr(out)), self._h)
        return out[:len(m)]

    def set_body(self, field: str, values):
        v = np.ascontiguousarray(values, np.float32)
        _check(lib().agx_world_write_body_f32(self._h, BODY_FIELDS.index(field), _ptr(v), len(v)), self._h)

    def collider(self, field: str) -> np.ndarray:
        n = self.counts.colliders
        if field in COLLIDER_U32:
            out = np.empty(n, np.uint32)
            _check(lib().agx_world_read_collider_u32(self._h, COLLIDER_U32.index(field), _ptr(out), n), self._h)
        else:
            out = np.empty(n, np.float32)
            _check(lib().agx_world_read_collider_f32(self._h, COLLIDER_F32.index(field), _ptr(out), n), self._h)
        return out

    def pairs(self):
        cnt = C.c_uint64()
        rc = lib().agx_world_read_pairs(self._h, None, None, 0, C.byref(cnt))
        n = cnt.value
        a = np.empty(max(n, 1), np.uint32)
        b = np.empty(max(n, 1), np.uint32)
        _check(lib().agx_world_read_pairs(self._h, _ptr(a), _ptr(b), n, C.byref(cnt)), self._h)
        return a[:n], b[:n]

    def broadphase_order(self, which: int = 0):
        """(keys, colliders, morton_bits) of the last build of the dynamic (0) or static (1) tree."""
        c = self.counts
        n = (c.colliders - c.static_colliders) if which == 0 else c.static_colliders
        keys = np.empty(max(n, 1), np.uint32)
        cols = np.empty(max(n, 1), np.uint32)
        bits = C.c_uint32()
        _check(lib().agx_world_read_broadphase_order(self._h, which, _ptr(keys), _ptr(cols), n, C.byref(bits)), self._h)
        return keys[:n], cols[:n], bits.value

    def manifolds(self, which: int = 0) -> np.ndarray:
        cnt = C.c_uint64()
        lib().agx_world_read_manifolds(self._h, which, None, 0, C.byref(cnt))
        out = np.empty(max(cnt.value, 1), MANIFOLD_DTYPE)
        _check(lib().agx_world_read_manifolds(self._h, which, _ptr(out), cnt.value, C.byref(cnt)), self._h)
        return out[:cnt.value]

    def set_collider_callback(self, collider: int, callback_id: int, threshold: float = 0.0):
        """Simulation::setColliderCollisionCallback (Simulation.cpp:152-172)."""
        _check(lib().agx_world_set_collider_callback(self._h, collider, callback_id, threshold), self._h)

    def events(self):
        """Collision events of the last update/step call and how many were dropped beyond the capacity."""
        cnt, dropped = C.c_uint64(), C.c_uint64()
        rc = lib().agx_world_read_events(self._h, None, 0, C.byref(cnt), C.byref(dropped))
        out = np.empty(max(cnt.value, 1), EVENT_DTYPE)
        if cnt.value:
            _check(lib().agx_world_read_events(self._h, _ptr(out), cnt.value, C.byref(cnt), C.byref(dropped)), self._h)
        elif rc:
            _check(rc, self._h)
        return out[:cnt.value], dropped.value

    def joint_order(self):
        """(order, colour starts, first colour of each type) of the joints' velocity passes."""
        n, k = C.c_uint32(), C.c_uint32()
        _check(lib().agx_world_read_joint_order(self._h, None, 0, None, 0, None, C.byref(n), C.byref(k)), self._h)
        order = np.empty(max(n.value, 1), np.uint32)
        starts = np.empty(k.value + 1, np.uint32)
        types = np.empty(6, np.uint32)
        _check(lib().agx_world_read_joint_order(self._h, _ptr(order), n.value, _ptr(starts), k.value + 1, _ptr(types),
                                                C.byref(n), C.byref(k)), self._h)
        return order[:n.value], starts, types

    def joints(self):
        """(joint records, the motors' accumulated impulses)."""
        n = C.c_uint32()
        _check(lib().agx_world_read_joint_order(self._h, None, 0, None, 0, None, C.byref(n), None), self._h)
        out = np.empty(max(n.value, 1), JOINT_DTYPE)
        imp = np.empty(max(n.value, 1), np.float32)
        _check(lib().agx_world_read_joints(self._h, _ptr(out), _ptr(imp), n.value), self._h)
        return out[:n.value], imp[:n.value]

    def sleep_state(self):
        """(sleep island id per body, 0 when awake; sleep timer per body)."""
        n = self.counts.bodies
        ids = np.empty(n, np.uint32)
        t = np.empty(n, np.float32)
        _check(lib().agx_world_read_sleep(self._h, _ptr(ids), _ptr(t), n), self._h)
        return ids, t

    def stash(self):
        """Manifolds of sleeping islands and their sleep island ids."""
        cnt = C.c_uint64()
        _check(lib().agx_world_read_stash(self._h, None, None, 0, C.byref(cnt)), self._h)
        m = np.empty(max(cnt.value, 1), MANIFOLD_DTYPE)
        isl = np.empty(max(cnt.value, 1), np.uint32)
        if cnt.value:
            _check(lib().agx_world_read_stash(self._h, _ptr(m), _ptr(isl), cnt.value, C.byref(cnt)), self._h)
        return m[:cnt.value], isl[:cnt.value]

    def save(self) -> bytes:
        """The world's state as a snapshot (agx_world_save)."""
        n = C.c_uint64()
        _check(lib().agx_world_snapshot_size(self._h, C.byref(n)), self._h)
        buf = C.create_string_buffer(n.value)
        _check(lib().agx_world_save(self._h, buf, n.value, C.byref(n)), self._h)
        return buf.raw[:n.value]

    def load(self, snapshot: bytes):
        """Restores a snapshot of a world built from the same scenes (agx_world_load)."""
        buf = C.create_string_buffer(bytes(snapshot), len(snapshot))
        _check(lib().agx_world_load(self._h, buf, len(snapshot)), self._h)

    def set_warm_start_source(self, manifolds: np.ndarray):
        m = np.ascontiguousarray(manifolds, MANIFOLD_DTYPE)
        _check(lib().agx_world_write_manifolds(self._h, 1, _ptr(m), len(m)), self._h)

    def solve_order(self):
        k = C.c_uint32()
        starts = np.empty(2048, np.uint32)
        lib().agx_world_read_solve_order(self._h, None, 0, _ptr(starts), len(starts), C.byref(k))
        total = int(starts[k.value]) if k.value else 0
        order = np.empty(max(total, 1), np.uint32)
        _check(lib().agx_world_read_solve_order(self._h, _ptr(order), total, _ptr(starts), len(starts), C.byref(k)),
               self._h)
        return order[:total], starts[:k.value + 1].copy()

    def islands(self):
        ids = np.empty(self.counts.bodies, np.uint32)
        cnt = C.c_uint32()
        _check(lib().agx_world_read_islands(self._h, _ptr(ids), len(ids), C.byref(cnt)), self._h)
        return ids, cnt.value

    def set_profiling(self, enabled: bool):
        _check(lib().agx_world_set_profiling(self._h, int(enabled)), self._h)

    def profile(self) -> dict:
        """{segment: (total ms, launches)} since the last call (HIP events on the world's stream)."""
        ms = np.zeros(32, np.float64)
        n = np.zeros(32, np.uint32)
        k = C.c_uint32()
        _check(lib().agx_world_read_profile(self._h, _ptr(ms), _ptr(n), 32, C.byref(k)), self._h)
        return {lib().agx_profile_segment_name(i).decode(): (float(ms[i]), int(n[i])) for i in range(k.value)}

    def stats(self) -> Stats:
        s = Stats()
        _check(lib().agx_world_read_stats(self._h, C.byref(s)), self._h)
        return s

    def scene_residual(self) -> np.ndarray:
        """Relative residual of each scene after its last velocity sweep of the last per-scene solve."""
        out = np.empty(self.counts.scenes, np.float32)
        _check(lib().agx_world_read_scene_residual(self._h, _ptr(out), len(out)), self._h)
        return out

    def scene_sweeps(self) -> np.ndarray:
        """Velocity sweeps each scene ran in the last step solved scene by scene."""
        out = np.empty(self.counts.scenes, np.uint32)
        _check(lib().agx_world_read_scene_sweeps(self._h, _ptr(out), len(out)), self._h)
        return out

    # -- learned warm start
    def set_learned(self, mlp: "Mlp | None", mode: int):
        """AGX_LEARN_* flags; the world keeps a reference to the network."""
        self._mlp = mlp
        _check(lib().agx_world_set_learned(self._h, mlp._h if mlp is not None else None, mode), self._h)

    def gate(self) -> Gate:
        g = Gate()
        _check(lib().agx_world_read_gate(self._h, C.byref(g)), self._h)
        return g

    def reset_gate(self):
        _check(lib().agx_world_reset_gate(self._h), self._h)

    # -- launch shapes
    def tuning(self) -> Tuning:
        t = Tuning()
        _check(lib().agx_world_get_tuning(self._h, C.byref(t)), self._h)
        return t

    def set_tuning(self, solve_blocks_per_cu: int, colour_blocks_per_cu: int, scene_whole: int = 0):
        t = Tuning(solve_blocks_per_cu=solve_blocks_per_cu, colour_blocks_per_cu=colour_blocks_per_cu,
                   scene_whole=scene_whole)
        _check(lib().agx_world_set_tuning(self._h, C.byref(t)), self._h)

    def autotune(self, dt: float, steps_per_trial: int = 2) -> Tuning:
        """Times graph-replayed steps per launch shape and keeps the fastest (advances the world)."""
        t = Tuning()
        _check(lib().agx_world_autotune(self._h, dt, steps_per_trial, C.byref(t)), self._h)
        return t


class Mlp:
    """The learned warm start's network (24 -> 128 -> 128 -> 2) with its training rows, on the GPU."""

    def __init__(self, seed: int = 1, sample_capacity: int = 1 << 20):
        self._h = C.c_void_p()
        _check(lib().agx_mlp_create(seed, sample_capacity, C.byref(self._h)))
        _track(self)

    def close(self):
        if getattr(self, "_h", None):
            lib().agx_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        if _LIB is not None:
            self.close()

    def params(self) -> np.ndarray:
        out = np.empty(MLP_PARAMS, np.float32)
        _check(lib().agx_mlp_get_params(self._h, _ptr(out), MLP_PARAMS))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().agx_mlp_set_params(self._h, _ptr(p), len(p)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        y = np.empty((len(x), MLP_OUTPUTS), np.float32)
        _check(lib().agx_mlp_forward(self._h, _ptr(x), len(x), _ptr(y)))
        return y

    def gradients(self, x, t):
        """(loss, gradient in host parameter layout) of one batch."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, MLP_OUTPUTS)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = C.c_float()
        _check(lib().agx_mlp_gradients(self._h, _ptr(x), _ptr(t), len(x), _ptr(g), C.byref(loss)))
        return float(loss.value), g

    def add_samples(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, MLP_OUTPUTS)
        _check(lib().agx_mlp_add_samples(self._h, _ptr(x), _ptr(t), len(x)))

    def sample_count(self) -> int:
        n = C.c_uint64()
        _check(lib().agx_mlp_sample_count(self._h, C.byref(n)))
        return n.value

    def samples(self, offset: int = 0, n: int | None = None):
        n = self.sample_count() - offset if n is None else n
        x = np.empty((n, MLP_INPUTS), np.float32)
        t = np.empty((n, MLP_OUTPUTS), np.float32)
        _check(lib().agx_mlp_read_samples(self._h, _ptr(x), _ptr(t), offset, n))
        return x, t

    def clear_samples(self):
        _check(lib().agx_mlp_clear_samples(self._h))

    def train(self, iterations: int, batch: int = 4096, learning_rate: float = 1e-3, seed: int = 1) -> float:
        loss = C.c_float()
        _check(lib().agx_mlp_train(self._h, iterations, batch, learning_rate, seed, C.byref(loss)))
        return float(loss.value)


def gate_update(g: Gate, used_learned: bool, cost: float) -> Gate:
    """The residual gate's update on the host (the same code the device runs after each step)."""
    _check(lib().agx_gate_update(C.byref(g), int(bool(used_learned)), cost))
    return g


def broadphase_pairs(aabb: np.ndarray, owner: np.ndarray, is_dynamic: np.ndarray, scene=None, cap=None):
    """BroadPhaseCollisionDetector::findCollisions on injected AABBs (n x 4: minX, minY, maxX, maxY)."""
    aabb = np.ascontiguousarray(aabb, np.float32)
    owner = np.ascontiguousarray(owner, np.uint32)
    dyn = np.ascontiguousarray(is_dynamic, np.uint8)
    sc = None if scene is None else np.ascontiguousarray(scene, np.uint32)
    n = len(owner)
    cap = cap if cap is not None else max(1 << 16, 64 * n)
    a = np.empty(cap, np.uint32)
    b = np.empty(cap, np.uint32)
    cnt = C.c_uint64()
    _check(lib().agx_broadphase_pairs(_ptr(aabb), _ptr(owner), _ptr(dyn), _ptr(sc), n, _ptr(a), _ptr(b), cap,
                                      C.byref(cnt)))
    return a[:cnt.value], b[:cnt.value]


def device_count() -> int:
    n = C.c_int()
    lib().agx_device_count(C.byref(n))
    return n.value


def build_scene(builder: Builder, scene: dict, scene_index: int = 0) -> np.ndarray:
    """Creates a scene description's materials (first scene only), bodies and joints."""
    if scene_index == 0 or not getattr(builder, "_materials_done", False):
        for m in scene.get("materials", []):
            builder.create_material(*m)
        builder._materials_done = True
    idx = builder.create_bodies(scene_index, scene["bodies"])
    if scene.get("joints"):
        builder.add_joints(scene_index, scene["joints"])
    return idx
