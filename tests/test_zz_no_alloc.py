"""agx_broadphase_pairs_noalloc against agx_broadphase_pairs, through the C ABI.

The allocation-free variant works on device buffers (allocated here through the HIP runtime the
library runs on) and keeps every intermediate array in the library's static device array. Its pairs must be
the existing entry point's, bit for bit and in the same order, on two calls in a row (the second
smaller, with stale contents of the first left in the array), and it must refuse inputs beyond its
limits with an error instead of touching memory.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(n, scenes_n, seed):
    rng = np.random.default_rng(seed)
    size = np.sqrt(n) * 1.2
    c = rng.uniform(0, size, (n, 2)).astype(np.float32)
    e = rng.uniform(0.1, 1.2, (n, 2)).astype(np.float32)
    aabb = np.ascontiguousarray(np.concatenate([c - e, c + e], axis=1), np.float32)
    owner = np.arange(n, dtype=np.uint32)
    owner[1::17] = owner[0::17][: len(owner[1::17])]
    dyn = (rng.random(n) > 0.1).astype(np.uint8)
    scene = rng.integers(0, scenes_n, n).astype(np.uint32) if scenes_n > 1 else None
    return aabb, owner, dyn, scene


class _Device:
    """Device buffers from the HIP runtime the library itself runs on (test plumbing only)."""

    H2D, D2H = 1, 2

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.held = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(arr.nbytes, 16)) == 0
        self.held.append(p)
        if arr.nbytes:
            assert self.hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, self.H2D) == 0
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype)
        if out.nbytes:
            assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        for p in self.held:
            self.hip.hipFree(p)
        self.held = []


def _noalloc(agx, aabb, owner, dyn, scene, num_scenes, cap):
    dev = _Device()
    try:
        d_aabb, d_owner, d_dyn = dev.put(aabb), dev.put(owner), dev.put(dyn)
        d_scene = None if scene is None else dev.put(scene)
        d_a = dev.put(np.full(max(cap, 1), 0xFFFFFFFF, np.uint32))
        d_b = dev.put(np.full(max(cap, 1), 0xFFFFFFFF, np.uint32))
        cnt = C.c_uint64()
        rc = agx.lib().agx_broadphase_pairs_noalloc(d_aabb, d_owner, d_dyn, d_scene, len(owner), num_scenes, d_a, d_b,
                                                     cap, C.byref(cnt))
        n = cnt.value if rc == 0 else 0
        return rc, dev.get(d_a, n, np.uint32), dev.get(d_b, n, np.uint32)
    finally:
        dev.free()


def test_noalloc_broadphase_matches_the_allocating_entry_point(agx):
    for n, scenes_n, seed in ((20000, 7, 1), (3000, 1, 2)):   # the second call is smaller
        aabb, owner, dyn, scene = _inputs(n, scenes_n, seed)
        ref_a, ref_b = agx.broadphase_pairs(aabb, owner, dyn, scene)
        cap = 64 * n
        rc, a, b = _noalloc(agx, aabb, owner, dyn, scene, scenes_n, cap)
        assert rc == 0, agx.lib().agx_last_error()
        assert len(ref_a) > n // 10
        assert np.array_equal(a, ref_a) and np.array_equal(b, ref_b), (n, scenes_n)
    # a collider count beyond the maximum, a scene index beyond num_scenes, a small cap
    big = agx.lib().agx_broadphase_pairs_noalloc
    cnt = C.c_uint64()
    rc = big(None, None, None, None, (1 << 16) + 1, 1, None, None, 0, C.byref(cnt))
    assert rc == -2 and b"maximum" in agx.lib().agx_last_error()
    aabb, owner, dyn, scene = _inputs(500, 3, 3)
    rc, _, _ = _noalloc(agx, aabb, owner, dyn, scene, 2, 1 << 14)
    assert rc == -1 and b"num_scenes" in agx.lib().agx_last_error()
    rc, _, _ = _noalloc(agx, aabb, owner, dyn, scene, 3, 4)
    assert rc == -2
    # and a last full call after the refused ones
    ref_a, ref_b = agx.broadphase_pairs(aabb, owner, dyn, scene)
    rc, a, b = _noalloc(agx, aabb, owner, dyn, scene, 3, 1 << 14)
    assert rc == 0 and np.array_equal(a, ref_a) and np.array_equal(b, ref_b)
