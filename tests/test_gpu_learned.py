"""GPU checks of the learned warm start: the MFMA forward and backward kernels against the oracle
network, training on the GPU, the world's feature/collect/predict hooks and the residual gate."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

IN, H, OUT = 24, 128, 2
TENSORS = {"W1": (0, H * IN), "b1": (H * IN, H * IN + H), "W2": (H * IN + H, H * IN + H + H * H),
           "b2": (H * IN + H + H * H, H * IN + 2 * H + H * H), "W3": (H * IN + 2 * H + H * H, H * IN + 2 * H + H * H + 2 * H),
           "b3": (H * IN + 2 * H + H * H + 2 * H, 19970)}


def features(rng, n):
    x = rng.normal(0, 1, (n, IN)).astype(np.float32)
    x[:, 4:6] = rng.integers(0, 2, (n, 2))          # flags
    x[:, 0] = np.abs(rng.normal(0, 8, n))            # cached impulses span a wide range
    return x


def test_forward_matches_oracle(agx, oracle):
    m = agx.Mlp(seed=3, sample_capacity=16)
    x = features(np.random.default_rng(1), 1000)
    y = m.forward(x)
    yo = oracle.mlp_forward(m.params(), x, rounding=True)
    scale = max(1.0, float(np.abs(yo).max()))
    err = np.abs(y - yo)
    # same bf16 rounding points; only fp32 summation order differs (it can flip a bf16 rounding)
    assert err.max() <= 2e-2 * scale
    assert np.sqrt((err ** 2).mean()) <= 2e-3 * scale


def test_gradients_match_oracle(agx, oracle):
    m = agx.Mlp(seed=5, sample_capacity=16)
    rng = np.random.default_rng(2)
    x = features(rng, 512)
    t = rng.normal(0, 2, (512, OUT)).astype(np.float32)
    loss, g = m.gradients(x, t)
    lo, go = oracle.mlp_gradients(m.params(), x, t, rounding=True)
    assert loss == pytest.approx(lo, rel=2e-3)
    for name, (a, b) in TENSORS.items():
        rel = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-12)
        assert rel < 2e-2, (name, rel)
    # a batch that is not a multiple of 16 is rejected
    with pytest.raises(agx.AgxError):
        m.gradients(x[:100], t[:100])


def test_training_on_gpu_fits_a_smooth_target(agx, oracle):
    rng = np.random.default_rng(3)
    x = features(rng, 1 << 16)
    t = np.stack([0.5 * x[:, 0] + np.tanh(x[:, 6]), x[:, 7] * x[:, 10] - 0.3 * x[:, 2]], 1).astype(np.float32)
    m = agx.Mlp(seed=7, sample_capacity=1 << 17)
    m.add_samples(x, t)
    assert m.sample_count() == len(x)
    xr, tr = m.samples(100, 10)
    np.testing.assert_array_equal(tr, t[100:110])
    y0 = m.forward(x[:4096])
    before = float(((y0 - t[:4096]) ** 2).sum(1).mean())
    m.train(1500, 2048, 1e-3, seed=1)
    m.train(500, 2048, 3e-4, seed=2)
    y1 = m.forward(x[:4096])
    after = float(((y1 - t[:4096]) ** 2).sum(1).mean())
    assert after < before / 20, (before, after)
    # the trained parameters still agree with the oracle's forward pass
    yo = oracle.mlp_forward(m.params(), x[:2048], rounding=True)
    assert np.abs(yo - y1[:2048]).max() <= 2e-2 * max(1.0, float(np.abs(yo).max()))


def _world(agx, scenes, **settings):
    scene = scenes.pile_small(seed=1)
    b = agx.Builder()
    agx.build_scene(b, scene)
    s = agx.default_settings(**scene["settings"])
    for k, v in settings.items():
        setattr(s, k, v)
    return agx.World(b, s)


def _state(w):
    return np.stack([w.body(f) for f in ("com_x", "com_y", "rotation", "vel_x", "vel_y", "ang_vel")])


def test_zero_network_reproduces_cached_warm_start(agx, scenes):
    """Corrections of an all-zero network are exactly zero, so predict mode is bit-identical to the
    cached warm start (fixed sweep counts keep both runs on the same schedule)."""
    a = _world(agx, scenes)
    b = _world(agx, scenes)
    m = agx.Mlp(seed=1, sample_capacity=16)
    m.set_params(np.zeros(agx.MLP_PARAMS, np.float32))
    dt = 1.0 / 60.0
    a.step(dt, 30)
    b.step(dt, 30)
    np.testing.assert_array_equal(_state(a), _state(b))
    b.set_learned(m, agx.LEARN_PREDICT | agx.LEARN_UNGATED)
    a.step(dt, 20)
    b.step(dt, 20)
    assert b.stats().learned_warm_start == 1 and a.stats().learned_warm_start == 0
    np.testing.assert_array_equal(_state(a), _state(b))
    # a nonzero correction does change the step (the prediction is applied)
    p = np.zeros(agx.MLP_PARAMS, np.float32)
    p[-2] = 0.5     # b3[0]: +0.5 S on every normal impulse
    m.set_params(p)
    a.step(dt, 1)
    b.step(dt, 1)
    assert not np.array_equal(_state(a), _state(b))


def test_collect_rows_are_the_steps_contact_points(agx, scenes):
    w = _world(agx, scenes)
    dt = 1.0 / 60.0
    w.step(dt, 40)
    m = agx.Mlp(seed=1, sample_capacity=1 << 16)
    w.set_learned(m, agx.LEARN_COLLECT)
    w.step(dt, 1)
    st = w.stats()
    assert st.learned_warm_start == 0
    assert m.sample_count() == st.contacts > 0
    x, t = m.samples()
    assert np.isfinite(x).all() and np.isfinite(t).all()
    assert set(np.unique(x[:, 4])) <= {0.0, 1.0} and set(np.unique(x[:, 5])) <= {0.0, 1.0}
    # the final normal impulse is never negative: target >= -(cached normal impulse) in units of S
    assert (t[:, 0] + x[:, 0] >= -1e-2 * np.maximum(1.0, np.abs(x[:, 0]))).all()
    # a settled pile mostly reuses cached impulses
    assert x[:, 4].mean() > 0.5
    w.step(dt, 1)
    assert m.sample_count() == st.contacts + w.stats().contacts
    w.set_learned(None, agx.LEARN_OFF)
    n = m.sample_count()
    w.step(dt, 1)
    assert m.sample_count() == n


def _batch_world(agx, scenes, n_scenes, **settings):
    batch = scenes.batch(n_scenes)
    b = agx.Builder(len(batch))
    for i, sc in enumerate(batch):
        agx.build_scene(b, sc, i)
    return agx.World(b, agx.default_settings(**settings))


def test_gate_rejects_a_harmful_network(agx, scenes):
    """A network that cancels every cached normal impulse (a cold start) leaves a larger residual
    after the early sweeps: the gate hands the warm start back to the cached impulses and probes the
    network ever more rarely."""
    w = _batch_world(agx, scenes, 64, velocity_iterations=16, position_iterations=3, track_residual=1)
    dt = 1.0 / 60.0
    w.step(dt, 240)
    m = agx.Mlp(seed=1, sample_capacity=16)
    p = np.zeros(agx.MLP_PARAMS, np.float32)
    p[-2] = -1e4     # b3[0]: jn = max(0, jn_cached - 64 S) = 0 for every point in range
    m.set_params(p)
    w.set_learned(m, agx.LEARN_PREDICT)
    used = []
    for _ in range(48):
        w.step(dt, 1)
        used.append(int(w.stats().learned_warm_start))
    g = w.gate()
    info = (used, g.cost_learned, g.cost_cached)
    assert g.cost_learned > g.cost_cached, info
    assert g.use_learned == 0, info
    assert used[:2] == [1, 0] and sum(used[2:]) <= 2, info     # probes after 16 and 32 more steps
    assert g.steps_cached + g.steps_learned == 48


def test_trained_network_needs_fewer_sweeps_than_cached_warm_start(agx, scenes):
    """What the trained network gains: two copies of a 256-scene batch settled at the 1e-3 residual
    target (each scene stops at its own sweep); the network is trained on the GPU on one copy's own
    solves (rows of contact points without a cached impulse), then both step on, one warm-started
    by the network at those points (gated), and the sweeps per dynamic body are compared."""
    dt = 1.0 / 60.0
    base = dict(velocity_iterations=64, residual_target=1e-3, position_iterations=3, track_residual=1)
    worlds = [_batch_world(agx, scenes, 256, **base) for _ in range(2)]
    for w in worlds:
        w.step_graph(dt, 150)
    plain, learned = worlds
    m = agx.Mlp(seed=1, sample_capacity=4 << 20)
    learned.set_learned(m, agx.LEARN_COLLECT | agx.LEARN_NEW_CONTACTS)
    for _ in range(12):
        plain.step(dt, 1)
        learned.step(dt, 1)
    assert m.sample_count() > 10000
    m.train(3000, 4096, 1e-3, seed=1)
    m.train(1000, 4096, 2e-4, seed=2)
    learned.set_learned(m, agx.LEARN_PREDICT | agx.LEARN_NEW_CONTACTS)
    s0 = [w.stats().total_body_sweeps for w in worlds]
    steps = 40
    for _ in range(steps):
        plain.step(dt, 1)
        learned.step(dt, 1)
    s1 = [w.stats().total_body_sweeps for w in worlds]
    dyn = plain.counts.bodies - plain.counts.static_bodies
    sweeps_plain, sweeps_learned = (s1[0] - s0[0]) / (steps * dyn), (s1[1] - s0[1]) / (steps * dyn)
    g = learned.gate()
    print(f"sweeps per body to 1e-3: cached {sweeps_plain:.2f}, learned {sweeps_learned:.2f}; gate steps "
          f"{g.steps_cached} cached / {g.steps_learned} learned")
    assert g.steps_learned > 0
    assert sweeps_learned < sweeps_plain


def test_new_contacts_mode_rows_are_the_uncached_points(agx, scenes):
    """In new-contacts mode a world collects one row per contact point without a cached impulse:
    exactly the rows with feature 4 (has a cached impulse) equal to 0 that full collection gives on
    an identical world, except that the mode keeps no accumulator history (features 8, 17 and 18
    read "none"); a network whose correction is +0.5 S then changes the step."""
    dt = 1.0 / 60.0
    full, new = _world(agx, scenes), _world(agx, scenes)
    for w in (full, new):
        w.step(dt, 12)      # still settling: new contacts every step
    mf = agx.Mlp(seed=1, sample_capacity=1 << 16)
    mn = agx.Mlp(seed=1, sample_capacity=1 << 16)
    full.set_learned(mf, agx.LEARN_COLLECT)
    new.set_learned(mn, agx.LEARN_COLLECT | agx.LEARN_NEW_CONTACTS)
    full.step(dt, 1)
    new.step(dt, 1)
    np.testing.assert_array_equal(_state(full), _state(new))
    xf, tf = mf.samples()
    xn, tn = mn.samples()
    assert 0 < len(xn) < len(xf)
    assert (xn[:, 4] == 0).all() and (xn[:, 0] == 0).all() and (xn[:, 1] == 0).all()
    hist = [8, 17, 18]
    assert (xn[:, hist] == 0).all()
    sel = xf[:, 4] == 0
    xf = xf.copy()
    xf[:, hist] = 0
    key = lambda x, t: np.lexsort(np.concatenate([x, t], 1).T[::-1])
    a, b = np.concatenate([xf[sel], tf[sel]], 1), np.concatenate([xn, tn], 1)
    np.testing.assert_array_equal(a[key(xf[sel], tf[sel])], b[key(xn, tn)])
    # prediction in new-contacts mode: a constant +0.5 S normal correction reaches only new points
    p = np.zeros(agx.MLP_PARAMS, np.float32)
    p[-2] = 0.5
    mp = agx.Mlp(seed=1, sample_capacity=1 << 16)
    mp.set_params(p)
    full.set_learned(None, agx.LEARN_OFF)
    new.set_learned(mp, agx.LEARN_PREDICT | agx.LEARN_UNGATED | agx.LEARN_NEW_CONTACTS)
    full.step(dt, 1)
    new.step(dt, 1)
    assert new.stats().learned_warm_start == 1
    assert not np.array_equal(_state(full), _state(new))
