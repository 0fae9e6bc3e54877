"""Steps replayed from captured HIP graphs against the same steps as plain launches: bit-identical
state, manifolds and islands (fixed sweep counts), across parity switches, settings changes and
the learned warm start."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("com_x", "com_y", "rotation", "vel_x", "vel_y", "ang_vel")


def _world(agx, scenes):
    scene = scenes.pile_small(seed=1)
    b = agx.Builder()
    agx.build_scene(b, scene)
    return agx.World(b, agx.default_settings(**scene["settings"], track_residual=1))


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
