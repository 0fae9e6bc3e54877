"""The oracle's threaded step (the CPU baseline's structure, oracle/threaded.cpp) against its
sequential paths:
- the wave plan (SolvingPlanner.cpp:141-310) covers every manifold once, and within a wave no two
  workers' passes touch one cache-line group of dynamic bodies (bodies >> 4), so passes may run at
  once;
- the threaded solve (waves behind barriers, BodyCollisionSolver.cpp:982-1031) equals the
  sequential solve in the plan's flat order bit for bit;
- the threaded narrow phase (1,024-pair chunks merged in order) equals the sequential one;
- a threaded run keeps the pairs, manifold count and islands of a sequential run of the same state
  for one step.
"""
import numpy as np
import pytest

DT = 1.0 / 60.0


@pytest.fixture(scope="module")
def settled(oracle, scenes):
    # 10,000 bodies packed in shelves in a narrow box, fallen into contact for 3 s
    scene = scenes.pile(7000, 3000, 60.0, 400.0, 5, r_poly=("lognormal", 0.3, 0.3, 0.1, 0.9),
                        r_circle=("lognormal", 0.3, 0.3, 0.1, 0.9), layout="shelves", name="threaded")
    w = oracle.world_from_scene(scene, dict(v=8, p=3))
    w.update(DT, 180)
    return scene, w


@pytest.fixture(scope="module")
def pool(oracle):
    p = oracle.Pool(4)
    yield p
    p.close()


def test_wave_plan_is_a_partition_with_disjoint_worker_groups(oracle, settled):
    _, w = settled
    m = w.manifolds()
    assert len(m) > 6640
    col_body = w.get("colBody")
    n_static = w.counts()["static_bodies"]
    flat, starts, sizes = w.plan_waves(m, 4)
    assert np.array_equal(np.sort(flat), np.arange(len(m)))
    assert len(starts) % 4 == 0 and np.array_equal(starts[1:], starts[:-1] + sizes[:-1]) and starts[0] == 0
    waves = len(starts) // 4
    assert 1 <= waves < 200
    for wave in range(waves):
        owner = {}
        for k in range(4):
            p = wave * 4 + k
            for c in flat[starts[p]:starts[p] + sizes[p]]:
                for b in (col_body[m["a"][c]], col_body[m["b"][c]]):
                    if b < n_static:
                        continue
                    g = int(b) >> 4
                    assert owner.setdefault(g, k) == k, (wave, g)
        assert sum(sizes[wave * 4:(wave + 1) * 4]) > 0


def test_threaded_solve_equals_sequential_solve_in_plan_order(oracle, settled, pool):
    _, w = settled
    m = w.manifolds()
    workers = min(pool.size, len(m) // 3320)
    assert workers >= 2
    flat, _, _ = w.plan_waves(m, workers)
    state = {f: w.get(f).copy() for f in ("velX", "velY", "angVel", "comX", "comY")}
    seq = w.solve(m, order=flat, v_it=8, p_it=3)
    after_seq = {f: w.get(f).copy() for f in state}
    for f, v in state.items():
        w.set(f, v)
    thr = w.solve_threaded(pool, m, v_it=8, p_it=3)
    for f in state:
        assert np.array_equal(w.get(f), after_seq[f]), f
    assert seq.tobytes() == thr.tobytes()
    for f, v in state.items():
        w.set(f, v)


def test_threaded_narrow_phase_equals_sequential(oracle, settled, pool):
    _, w = settled
    pa, pb = w.broad_bvh()
    assert len(pa) > 6000
    assert w.narrow(pa, pb).tobytes() == w.narrow_threaded(pool, pa, pb).tobytes()


def test_threaded_step_keeps_pairs_and_islands(oracle, settled, pool):
    scene, w = settled
    a = oracle.world_from_scene(scene, dict(v=8, p=3))
    b = oracle.world_from_scene(scene, dict(v=8, p=3))
    for f in ("comX", "comY", "rot", "velX", "velY", "angVel", "cos", "sin", "wX", "wY"):
        a.set(f, w.get(f))
        b.set(f, w.get(f))
    a.set_prev_manifolds(w.manifolds())
    b.set_prev_manifolds(w.manifolds())
    a.update(DT, 1)
    b.update_threaded(pool, DT, 1)
    ma, mb = a.manifolds(), b.manifolds()
    assert len(ma) == len(mb) > 6640
    for f in ("a", "b", "nx", "ny", "depth", "count"):
        assert np.array_equal(ma[f], mb[f]), f
    assert np.array_equal(a.get("islandId"), b.get("islandId"))
    # the solves ran in different orders: eight sweeps leave different impulses, but the same
    # total support of the pile to within a few per cent
    assert np.sum(mb["jn0"] + mb["jn1"]) == pytest.approx(np.sum(ma["jn0"] + ma["jn1"]), rel=0.05)
