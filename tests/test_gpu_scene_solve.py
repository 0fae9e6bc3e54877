"""The per-scene solve (each scene of a batch solved by one workgroup with its bodies in LDS)
against the dataflow solve over the whole batch: the same sequential colour-major order, so fixed
sweep counts give identical bits (and the per-scene islands the same ids as the device-wide
union-find); with a residual target every scene stops where the oracle, replaying that scene
alone in the GPU's order, reaches the target."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("com_x", "com_y", "rotation", "vel_x", "vel_y", "ang_vel")
DT = 1.0 / 60.0


def _world(agx, parts, **settings):
    b = agx.Builder(len(parts))
    for i, sc in enumerate(parts):
        agx.build_scene(b, sc, i)
    return agx.World(b, agx.default_settings(**settings))


def _same(a, b):
    for f in FIELDS:
        np.testing.assert_array_equal(a.body(f), b.body(f), err_msg=f)
    np.testing.assert_array_equal(a.manifolds(), b.manifolds())
    # islands: device-wide union-find (dataflow solve) and per-scene union-find in LDS agree exactly
    ia, ka = a.islands()
    ib, kb = b.islands()
    assert ka == kb
    np.testing.assert_array_equal(ia, ib)


def test_per_scene_solve_matches_dataflow_bit_for_bit(agx, scenes):
    parts = scenes.batch(24)
    base = dict(velocity_iterations=8, position_iterations=3, track_residual=1, track_errors=1)
    flow = _world(agx, parts, solve_per_scene=0, **base)
    per = _world(agx, parts, solve_per_scene=1, **base)
    print("bodies", flow.counts.bodies, flush=True)
    for n in (1, 10, 29):
        flow.step(DT, n)
        per.step_graph(DT, n)
        _same(flow, per)
        # the per-scene colouring (rounds in LDS) gives the dataflow colouring's colours exactly
        of, cf = flow.solve_order()
        op, cp = per.solve_order()
        np.testing.assert_array_equal(cf, cp)
        np.testing.assert_array_equal(of, op)
    sf, sp = flow.stats(), per.stats()
    assert sf.solved_per_scene == 0 and sp.solved_per_scene == 1
    assert sf.overflow == 0 and sp.overflow == 0 and sp.manifolds > 0
    assert sp.velocity_sweeps == 8 and sp.mean_velocity_sweeps == pytest.approx(8.0)
    assert sp.total_body_sweeps == sf.total_body_sweeps
    np.testing.assert_allclose(sp.residual[:8], sf.residual[:8], rtol=1e-4)
    assert sp.velocity_error == pytest.approx(sf.velocity_error, rel=1e-4)


@pytest.mark.parametrize("whole_cap", ["0", "5"])
def test_lane_groups_match_dataflow_bit_for_bit(agx, scenes, monkeypatch, whole_cap):
    """Scenes past the first `whole_cap` (largest first) are coloured and solved by lane groups of a
    block that meet at LDS arrival counters instead of block barriers: still the dataflow's colours
    and bits."""
    parts = scenes.batch(40, n_max=1500)
    base = dict(velocity_iterations=8, position_iterations=3, track_residual=1)
    flow = _world(agx, parts, solve_per_scene=0, **base)
    monkeypatch.setenv("AGX_SCENE_WHOLE_CAP", whole_cap)
    per = _world(agx, parts, solve_per_scene=1, **base)
    print("bodies", flow.counts.bodies, "whole", whole_cap, flush=True)
    for n in (1, 12):
        flow.step(DT, n)
        per.step(DT, n)
        _same(flow, per)
        # the colouring by lane groups gives the dataflow colouring's colours too
        of, cf = flow.solve_order()
        op, cp = per.solve_order()
        np.testing.assert_array_equal(cf, cp)
        np.testing.assert_array_equal(of, op)
    sf, sp = flow.stats(), per.stats()
    assert sp.solved_per_scene == 1 and sp.overflow == 0 and sp.manifolds > 0
    assert sp.total_body_sweeps == sf.total_body_sweeps
    np.testing.assert_allclose(sp.residual[:8], sf.residual[:8], rtol=1e-4)


def test_many_small_scenes_keep_size_order(agx, scenes):
    """More scenes than the per-step work order sorts (4,096): the scenes go in size order, most of
    them to lane groups; still the dataflow's colours and bits."""
    parts = [scenes.pile(3, 2, 6.0, 6.0, seed=100 + i) for i in range(4200)]
    base = dict(velocity_iterations=8, position_iterations=3)
    flow = _world(agx, parts, solve_per_scene=0, **base)
    per = _world(agx, parts, solve_per_scene=1, **base)
    print("scenes", len(parts), "bodies", flow.counts.bodies, flush=True)
    for n in (1, 20):
        flow.step(DT, n)
        per.step(DT, n)
        _same(flow, per)
        np.testing.assert_array_equal(flow.solve_order()[1], per.solve_order()[1])
    assert per.stats().solved_per_scene == 1 and per.stats().overflow == 0 and per.stats().manifolds > 0


def test_per_scene_colouring_from_global_scratch(agx, scenes, monkeypatch):
    """Scenes with more manifolds than the colouring keeps in LDS colour from global scratch
    (forced here with a tiny LDS cap): still the dataflow colouring's colours and bits."""
    parts = scenes.batch(16)
    base = dict(velocity_iterations=6, position_iterations=2)
    flow = _world(agx, parts, solve_per_scene=0, **base)
    monkeypatch.setenv("AGX_COLOUR_LDS_CAP", "64")
    per = _world(agx, parts, solve_per_scene=1, **base)
    for n in (1, 6):
        flow.step(DT, n)
        per.step(DT, n)
        _same(flow, per)
        np.testing.assert_array_equal(flow.solve_order()[0], per.solve_order()[0])
    assert per.stats().solved_per_scene == 1


@pytest.mark.parametrize("whole_cap", [None, "0"])
def test_per_scene_target_matches_oracle_scene_by_scene(agx, oracle, scenes, monkeypatch, whole_cap):
    """One step with a 1e-3 residual target, solved scene by scene (by whole blocks, or with
    whole_cap 0 by lane groups): every scene ran the sweeps the oracle needs to reach the target on
    that scene alone (replayed in the GPU's order), and ended in the oracle's state after exactly
    those sweeps."""
    if whole_cap is not None:
        monkeypatch.setenv("AGX_SCENE_WHOLE_CAP", whole_cap)
    parts = [scenes.batch_scene(i, seed=4, n_min=64, n_max=600) for i in range(9)]
    target = 1e-3
    w = _world(agx, parts, velocity_iterations=64, residual_target=target, position_iterations=3, track_residual=1)
    w.step(DT, 45)
    cscene = w.collider("scene")
    bscene = np.empty(w.counts.bodies, np.uint32)
    bscene[w.collider("body")] = cscene
    for p in range(agx.PHASE_SOLVE):
        w.run_phase(p, DT)
    mg = w.manifolds(0)
    state = {f: w.body(f) for f in ("com_x", "com_y", "rotation", "vel_x", "vel_y", "ang_vel", "cos", "sin")}
    w.run_phase(agx.PHASE_SOLVE, DT)
    st = w.stats()
    assert st.solved_per_scene == 1 and st.overflow == 0
    sweeps = w.scene_sweeps()
    gres = w.scene_residual()
    order, _ = w.solve_order()
    after = {f: w.body(f) for f in ("vel_x", "vel_y", "ang_vel", "com_x", "com_y")}
    pos_of = np.empty(len(order), np.int64)
    pos_of[order] = np.arange(len(order))
    renames = {"com_x": "comX", "com_y": "comY", "rotation": "rot", "vel_x": "velX", "vel_y": "velY",
               "ang_vel": "angVel", "cos": "cos", "sin": "sin"}
    checked = 0
    for s_, sc in enumerate(parts):
        bs = np.flatnonzero(bscene == s_)
        cs = np.flatnonzero(cscene == s_)
        c_off = int(cs[0])
        msel = np.flatnonzero((mg["a"] >= c_off) & (mg["a"] < c_off + len(cs)))
        ms = mg[msel].copy()
        ms["a"] -= c_off
        ms["b"] -= c_off
        local = np.argsort(pos_of[msel], kind="stable").astype(np.uint32)

        def fresh():
            ow = oracle.world_from_scene(sc, dict(v=64, p=3))
            for f, o in renames.items():
                ow.set(o, state[f][bs])
            return ow
        _, res = fresh().solve(ms, order=local, v_it=64, p_it=3, residual=True)
        below = np.flatnonzero(res <= target)
        k = int(below[0]) + 1 if len(below) else 64
        # a residual within rounding of the target may fall either side on the GPU
        near = abs(res[k - 1] - target) < 1e-4 * target or (k > 1 and abs(res[k - 2] - target) < 1e-4 * target)
        if near:
            continue
        assert sweeps[s_] == k, (s_, int(sweeps[s_]), k, res[:k + 1])
        assert gres[s_] == pytest.approx(res[k - 1], rel=1e-3), (s_, float(gres[s_]), float(res[k - 1]))
        ow = fresh()
        ow.solve(ms, order=local, v_it=k, p_it=3)
        for f in ("vel_x", "vel_y", "ang_vel", "com_x", "com_y"):
            np.testing.assert_allclose(after[f][bs], ow.get(renames[f]), rtol=1e-4, atol=1e-4, err_msg=f"{s_} {f}")
        checked += 1
    assert checked >= 7
    # the convergence counts: every scene that stopped before the cap met the target; at the cap,
    # those whose last residual is at most the target
    dyn = np.bincount(bscene[w.body("inv_mass") != 0], minlength=len(parts))
    early = sweeps < 64
    assert np.all(gres[early] <= target * (1 + 1e-4))
    at_cap_met = (sweeps == 64) & (gres <= target * (1 + 1e-4))
    assert int(early.sum()) <= st.converged_scenes <= int(early.sum() + at_cap_met.sum())
    assert int(dyn[early].sum()) <= st.converged_bodies <= int(dyn[early | at_cap_met].sum())
    assert st.converged_bodies > 0
    assert st.velocity_sweeps == sweeps.max()
    assert len(set(sweeps.tolist())) > 1          # scenes stopped apart
    assert st.mean_velocity_sweeps < st.velocity_sweeps


def test_scene_too_large_for_lds_uses_the_dataflow(agx, scenes):
    small = _world(agx, [scenes.pile_small(seed=1)], velocity_iterations=8)
    small.step(DT, 3)
    assert small.stats().solved_per_scene == 1
    big = _world(agx, [scenes.pile(9000, 3000, 140.0, 140.0, seed=3)], velocity_iterations=8)
    assert big.counts.bodies > 11000
    big.step(DT, 3)
    assert big.stats().solved_per_scene == 0 and big.stats().overflow == 0


def test_dataflow_solve_counts_convergence(agx, scenes):
    """One scene solved by the device-wide dataflow with a residual target: the step's sweeps stop
    at the target for all its bodies together, so either every dynamic body counts as converged or
    none does, and the counts add up over the steps."""
    w = _world(agx, [scenes.pile_small(seed=1)], velocity_iterations=agx.MAX_SWEEPS, residual_target=1e-3,
               track_residual=1, solve_per_scene=0)
    ndyn = w.counts.bodies - w.counts.static_bodies
    total0 = w.stats().total_converged_bodies
    steps = 0
    for _ in range(6):
        w.step(DT, 5)
        steps += 5
        st = w.stats()
        assert st.solved_per_scene == 0 and st.overflow == 0
        if st.velocity_sweeps < agx.MAX_SWEEPS:
            assert st.converged_scenes == 1 and st.converged_bodies == ndyn
        assert st.converged_bodies in (0, ndyn)
    assert 0 < st.total_converged_bodies - total0 <= steps * ndyn


def test_scene_launch_shapes_step_alike_and_autotune(agx, scenes):
    """How many of the heaviest scenes whole blocks take (the rest go to lane groups) changes no
    scene's result; the tuner of a world solved scene by scene picks among such shares (not among
    shapes of the dataflow kernels, which its steps never run) and reuses its pick from the cache."""
    parts = scenes.batch(40, n_max=1500)
    base = dict(velocity_iterations=64, residual_target=1e-3, position_iterations=3, track_residual=1)
    a, b = _world(agx, parts, **base), _world(agx, parts, **base)
    t = a.tuning()
    assert t.scene_whole >= 1
    b.set_tuning(t.solve_blocks_per_cu, t.colour_blocks_per_cu, 1)
    assert b.tuning().scene_whole == 1
    a.step(DT, 12)
    b.step_graph(DT, 12)
    _same(a, b)
    np.testing.assert_array_equal(a.scene_sweeps(), b.scene_sweeps())
    r = a.autotune(DT, 2)
    assert r.from_cache == 0 and r.ms_per_step > 0 and 1 <= r.scene_whole <= len(parts)
    assert a.tuning().scene_whole == r.scene_whole
    again = a.autotune(DT, 2)
    assert again.from_cache == 1 and again.scene_whole == r.scene_whole
