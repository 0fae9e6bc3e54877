"""GPU parity at the full sizes of the benchmark workloads, and of the broad phase's sorted keys.

- The 65,536-body pile (70 % polygons, lognormal radii, 512 x 256 m box), settled through step
  graphs and checked one step phase by phase against the oracle (parity_steps.check_step): pairs
  and every manifold field bit for bit, impulses and bodies within 1e-4 in the exported colour
  order (the device-wide dataflow solve), islands bit for bit, constraint errors within 1e-4.
- The 1,048,576-collider scatter with 16,384 statics and 5 % bullets at 200 m/s: CCD against the
  oracle (times of impact within 3e-5 of the sweep, the same bodies hit), the pair set bit for bit
  against the oracle's exact sweep and prune on the same AABBs, manifolds bit for bit.
- Morton keys: each key is the closed-form interleave of its collider's quantised AABB centre
  (Broad/BvhTree.cpp:27-40, :191-280) with the scene above it, keys ascend and equal keys keep
  ascending collider order (the stable sort of BvhTree.cpp:282-354).
"""
import numpy as np
import pytest

from parity_steps import check_step, oracle_from_gpu, pair_keys, assert_manifolds_equal

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def make_world(agx, scenes_list, **settings):
    b = agx.Builder(len(scenes_list))
    for s, sc in enumerate(scenes_list):
        agx.build_scene(b, sc, s)
    s = agx.default_settings(**{**scenes_list[0].get("settings", {}), **settings})
    return agx.World(b, s)


def test_pile_65536_step_matches_oracle_phase_by_phase(agx, oracle, scenes):
    scene = scenes.pile_large(seed=2)
    w = make_world(agx, [scene], track_errors=1)
    assert w.counts.bodies - w.counts.static_bodies == 65536
    w.step_graph(DT, 150)
    print(f"pile_large settled: {w.stats().manifolds} manifolds", flush=True)
    out = check_step(agx, oracle, scene, w, DT, 8, 3, min_manifolds=30000)
    assert not w.stats().solved_per_scene          # one large scene: the device-wide dataflow solve
    assert out["two_point"] > 1000 and out["colours"] >= 8
    print(out, flush=True)


def _quantised_keys(aabb, scene_of, scene_lo, scene_hi, morton_bits):
    """Closed form of the sort keys in float32 arithmetic: centre (min + max) * (q / (2 extent))
    + (-lo * q / extent), clamped to [0, q] and truncated, bits of x at even and of y at odd
    positions, the scene index above them."""
    f = np.float32
    axis_bits = morton_bits // 2
    q = f((1 << axis_bits) - 1)
    lo = scene_lo[scene_of]
    hi = np.maximum(scene_hi[scene_of], lo + f(1e-5))
    scale = (q / (hi - lo)).astype(np.float32)
    centre2 = (aabb[:, :2] + aabb[:, 2:]).astype(np.float32)
    t = centre2 * (scale * f(0.5)) + (-lo * scale)
    t = np.minimum(np.maximum(t.astype(np.float32), f(0)), q).astype(np.uint32)
    code = np.zeros(len(t), np.uint64)
    for bit in range(axis_bits):
        code |= ((t[:, 0].astype(np.uint64) >> bit) & 1) << np.uint64(2 * bit)
        code |= ((t[:, 1].astype(np.uint64) >> bit) & 1) << np.uint64(2 * bit + 1)
    return ((scene_of.astype(np.uint64) << np.uint64(morton_bits)) | code).astype(np.uint64)


@pytest.mark.parametrize("layout", ["single", "batch"])
def test_morton_keys_closed_form_and_stable_sort(agx, scenes, layout):
    if layout == "single":
        parts = [scenes.pile_small(seed=7)]
    else:
        parts = [scenes.batch_scene(i, seed=4, n_min=64, n_max=600) for i in range(11)]
    w = make_world(agx, parts, velocity_iterations=8)
    w.step(DT, 30)
    for p in range(agx.PHASE_BROAD + 1):
        w.run_phase(p, DT)
    aabb = np.stack([w.collider(f) for f in ("aabb_min_x", "aabb_min_y", "aabb_max_x", "aabb_max_y")], 1)
    scene_of = w.collider("scene")
    ns = w.counts.static_colliders
    is_dyn = w.body("inv_mass")[w.collider("body")] != 0
    for which, sel in ((0, np.flatnonzero(is_dyn)), (1, np.flatnonzero(~is_dyn))):
        keys, cols, bits = w.broadphase_order(which)
        assert len(keys) == len(sel) > 0
        assert np.array_equal(np.sort(cols), sel)              # a permutation of the tree's colliders
        nsc = len(parts)
        lo = np.full((nsc, 2), np.inf, np.float32)
        hi = np.full((nsc, 2), -np.inf, np.float32)
        for s in range(nsc):
            m = sel[scene_of[sel] == s]
            if len(m):
                lo[s] = aabb[m, :2].min(0)
                hi[s] = aabb[m, 2:].max(0)
        expect = _quantised_keys(aabb[sel], scene_of[sel], lo, hi, bits)
        by_col = dict(zip(sel.tolist(), expect.tolist()))
        assert [by_col[c] for c in cols.tolist()] == keys.astype(np.uint64).tolist()
        k64 = keys.astype(np.uint64)
        assert np.all(k64[1:] >= k64[:-1])                      # ascending
        same = k64[1:] == k64[:-1]
        assert np.all(cols[1:][same] > cols[:-1][same])          # stable: equal keys in collider order
        assert len(np.unique(k64)) > 0.9 * len(k64)
    assert ns == int(np.count_nonzero(~is_dyn))


def test_scatter_1m_ccd_pairs_and_manifolds_match_oracle(agx, oracle, scenes):
    scene = scenes.scatter_large(seed=3)
    w = make_world(agx, [scene])
    n_static = w.counts.static_colliders
    assert w.counts.colliders - n_static == 1 << 20 and n_static == 16384
    w.step(DT, 2)
    print("scatter_large built and stepped", flush=True)
    ow = oracle_from_gpu(oracle, scene, w, dict(gy=0.0))
    ow.snapshot()
    w.run_phase(agx.PHASE_INTEGRATE, DT)
    w.run_phase(agx.PHASE_POSE, DT)
    ow.integrate(DT)
    ow.wrap_cossin()
    ow.poses()
    ow.aabbs()
    com0 = np.stack([w.body("com_x"), w.body("com_y")], 1)
    prev = np.stack([ow.get("prevComX"), ow.get("prevComY")], 1)
    assert np.array_equal(com0[:, 0], ow.get("comX")) and np.array_equal(w.collider("aabb_max_x"), ow.get("maxX"))
    # CCD: the same bodies move back, to the same time of impact
    w.run_phase(agx.PHASE_CCD, DT)
    hit, t_orc = ow.ccd()
    assert hit
    com1 = np.stack([w.body("com_x"), w.body("com_y")], 1)
    moved = np.any(com1 != com0, axis=1)
    hit_orc = t_orc < 1.0
    print(f"ccd: {int(moved.sum())} bodies moved back on the GPU, {int(hit_orc.sum())} in the oracle", flush=True)
    assert int(hit_orc.sum()) > 1000
    assert np.array_equal(moved, hit_orc)
    # the GPU's time of impact from the move back: com1 = com0 + (com0 - prev) (t - 1)
    disp = (com0 - prev)[hit_orc].astype(np.float64)
    back = (com1 - com0)[hit_orc].astype(np.float64)
    t_gpu = 1.0 + np.sum(back * disp, 1) / np.sum(disp * disp, 1)
    np.testing.assert_allclose(t_gpu, t_orc[hit_orc], rtol=0, atol=3e-5)
    # broad phase on the GPU's AABBs: the oracle's exact sweep and prune
    w.run_phase(agx.PHASE_BROAD, DT)
    ga, gb = w.pairs()
    for f, o in (("aabb_min_x", "minX"), ("aabb_min_y", "minY"), ("aabb_max_x", "maxX"), ("aabb_max_y", "maxY"),
                 ("world_x", "wX"), ("world_y", "wY")):
        ow.set(o, w.collider(f))
    oa, ob = ow.broad_brute()
    kg = pair_keys(ga, gb, n_static)
    assert len(kg) > 10000 and len(np.unique(kg)) == len(kg)
    assert np.array_equal(kg, pair_keys(oa, ob, n_static))
    print(f"pairs: {len(kg)}", flush=True)
    # narrow phase on the GPU's pairs and poses
    w.run_phase(agx.PHASE_NARROW, DT)
    mg = w.manifolds(0)
    mo = ow.warm_start(ow.narrow(ga, gb))
    assert len(mg) > 1000
    assert_manifolds_equal(mg, mo, "scatter narrow")
    for p in range(agx.PHASE_SOLVE, agx.PHASE_FINISH + 1):
        w.run_phase(p, DT)
    st = w.stats()
    assert st.overflow == 0 and st.ccd_hits == int(hit_orc.sum())
