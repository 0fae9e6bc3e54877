"""Steps replayed from captured HIP graphs against the same steps as plain launches: bit-identical
state, manifolds and islands (fixed sweep counts), across parity switches, settings changes and
the learned warm start."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("com_x", "com_y", "rotation", "vel_x", "vel_y", "ang_vel")


def _world(agx, scenes, **settings):
    scene = scenes.pile_small(seed=1)
    b = agx.Builder()
    agx.build_scene(b, scene)
    return agx.World(b, agx.default_settings(**{**scene["settings"], "track_residual": 1, **settings}))


def _same(a, b):
    for f in FIELDS:
        np.testing.assert_array_equal(a.body(f), b.body(f), err_msg=f)
    ma, mb = a.manifolds(), b.manifolds()
    assert len(ma) == len(mb)
    np.testing.assert_array_equal(ma, mb)
    ia, ca = a.islands()
    ib, cb = b.islands()
    assert ca == cb
    np.testing.assert_array_equal(ia, ib)


def test_graph_steps_match_launched_steps(agx, scenes):
    a, b = _world(agx, scenes), _world(agx, scenes)
    dt = 1.0 / 60.0
    for n in (1, 7, 24, 3, 2):          # odd counts switch the ping-pong parity between calls
        a.step(dt, n)
        b.step_graph(dt, n)
        _same(a, b)                     # host reads between replays
    s = a.settings()
    s.velocity_iterations = 4
    a.set_settings(s)
    b.set_settings(s)                   # the graphs are captured again
    a.step(dt, 5)
    b.step_graph(dt, 5)
    _same(a, b)
    st = b.stats()
    assert st.velocity_sweeps == 4 and st.solved_steps == 42 and st.manifolds > 0


def test_graph_steps_with_learned_warm_start(agx, scenes):
    a, b = _world(agx, scenes), _world(agx, scenes)
    dt = 1.0 / 60.0
    a.step(dt, 30)
    b.step(dt, 30)
    m = agx.Mlp(seed=2, sample_capacity=1 << 16)
    for w in (a, b):
        w.set_learned(m, agx.LEARN_PREDICT | agx.LEARN_UNGATED)
    a.step(dt, 9)
    b.step_graph(dt, 9)
    _same(a, b)
    assert b.stats().learned_warm_start == 1


def test_launch_shapes_give_identical_steps_and_autotune(agx, scenes):
    """The colouring and the solve order inside a colour do not depend on how many blocks run the
    dataflow kernels, so every launch shape steps to the same bits; the tuner picks one of them and
    reuses it from its cache."""
    a, b = _world(agx, scenes, solve_per_scene=0), _world(agx, scenes, solve_per_scene=0)
    t = a.tuning()
    assert t.solve_blocks_per_cu >= 1 and t.max_solve_blocks_per_cu >= t.solve_blocks_per_cu
    b.set_tuning(1, 1)
    dt = 1.0 / 60.0
    a.step(dt, 20)
    b.step_graph(dt, 20)
    _same(a, b)
    with pytest.raises(agx.AgxError):
        b.set_tuning(t.max_solve_blocks_per_cu + 1, 1)
    r = a.autotune(dt, 2)
    assert r.from_cache == 0 and r.ms_per_step > 0
    assert 1 <= r.solve_blocks_per_cu <= r.max_solve_blocks_per_cu
    assert 1 <= r.colour_blocks_per_cu <= r.max_colour_blocks_per_cu
    now = a.tuning()
    assert (now.solve_blocks_per_cu, now.colour_blocks_per_cu) == (r.solve_blocks_per_cu, r.colour_blocks_per_cu)
    again = a.autotune(dt, 2)
    assert again.from_cache == 1
    assert (again.solve_blocks_per_cu, again.colour_blocks_per_cu) == (r.solve_blocks_per_cu, r.colour_blocks_per_cu)
