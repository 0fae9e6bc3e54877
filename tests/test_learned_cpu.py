"""CPU checks of the learned warm start: the oracle network pinned by finite differences and closed
forms, and the residual gate's state machine (the host copy of the device update)."""
import numpy as np
import pytest

IN, H, OUT = 24, 128, 2
OFF_W1, OFF_B1 = 0, H * IN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + OUT * H
PARAMS = OFF_B3 + OUT


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, np.float32)
    p[OFF_W1:OFF_B1] = rng.normal(0, np.sqrt(2 / IN), H * IN) * scale
    p[OFF_B1:OFF_W2] = rng.normal(0, 0.1, H)
    p[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2 / H), H * H) * scale
    p[OFF_B2:OFF_W3] = rng.normal(0, 0.1, H)
    p[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(2 / H), OUT * H) * scale
    p[OFF_B3:] = rng.normal(0, 0.1, OUT)
    return p


def test_parameter_count_matches_header(oracle, agx):
    assert PARAMS == oracle.MLP_PARAMS == agx.MLP_PARAMS == 19970


def test_oracle_gradient_matches_finite_differences(oracle):
    """Analytic backward (exact mode) against central differences of the loss in double."""
    p = random_params(1)
    rng = np.random.default_rng(2)
    x = rng.normal(0, 1, (12, IN)).astype(np.float32)
    t = rng.normal(0, 1, (12, OUT)).astype(np.float32)
    loss, g = oracle.mlp_gradients(p, x, t, rounding=False)
    assert loss == pytest.approx(oracle.mlp_loss(p, x, t), rel=1e-12)
    # probe entries of every tensor, including ones the ReLU masks (their gradient must be 0)
    picks = []
    for lo, hi in ((OFF_W1, OFF_B1), (OFF_B1, OFF_W2), (OFF_W2, OFF_B2), (OFF_B2, OFF_W3), (OFF_W3, OFF_B3),
                   (OFF_B3, PARAMS)):
        picks += list(rng.choice(np.arange(lo, hi), size=min(12, hi - lo), replace=False))
    h = 1e-6
    for k in picks:
        d = np.zeros(PARAMS)
        d[k] = h
        fd = (oracle.mlp_loss(p, x, t, d) - oracle.mlp_loss(p, x, t, -d)) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-4, abs=1e-7), k


def test_oracle_identity_network_closed_form(oracle):
    """W1, W2 = identity on the first units, W3 picks units 0 and 1: y = (relu(x0), relu(x1)); with
    rounding, the bf16 value of x (activations and inputs are rounded the same way)."""
    p = np.zeros(PARAMS, np.float32)
    for k in range(IN):
        p[OFF_W1 + k * IN + k] = 1.0
    for k in range(H):
        p[OFF_W2 + k * H + k] = 1.0
    p[OFF_W3 + 0 * H + 0] = 1.0
    p[OFF_W3 + 1 * H + 1] = 1.0
    p[OFF_B3:] = (0.5, -0.25)
    x = np.random.default_rng(3).normal(0, 3, (64, IN)).astype(np.float32)
    y = oracle.mlp_forward(p, x, rounding=False)
    np.testing.assert_array_equal(y, np.maximum(x[:, :2], 0) + np.float32([0.5, -0.25]))
    yr = oracle.mlp_forward(p, x, rounding=True)
    xb = (x.view(np.uint32) + 0x7FFF + ((x.view(np.uint32) >> 16) & 1)) & 0xFFFF0000
    xb = xb.view(np.float32)
    np.testing.assert_array_equal(yr, np.maximum(xb[:, :2], 0) + np.float32([0.5, -0.25]))


def test_oracle_loss_is_mean_squared_error_of_forward(oracle):
    p = random_params(4)
    rng = np.random.default_rng(5)
    x = rng.normal(0, 1, (32, IN)).astype(np.float32)
    t = rng.normal(0, 1, (32, OUT)).astype(np.float32)
    y = oracle.mlp_forward(p, x, rounding=False).astype(np.float64)
    assert oracle.mlp_loss(p, x, t) == pytest.approx(((y - t) ** 2).sum() / len(x), rel=1e-6)
    # a target equal to the output has zero loss and zero gradient
    yt = oracle.mlp_forward(p, x, rounding=False)
    loss, g = oracle.mlp_gradients(p, x, yt, rounding=False)
    assert loss < 1e-10 and np.abs(g).max() < 1e-5


def test_rounded_oracle_stays_close_to_exact(oracle):
    """The GPU's rounding points perturb the exact gradient by bf16-sized amounts only."""
    p = random_params(6)
    rng = np.random.default_rng(7)
    x = rng.normal(0, 1, (64, IN)).astype(np.float32)
    t = rng.normal(0, 1, (64, OUT)).astype(np.float32)
    le, ge = oracle.mlp_gradients(p, x, t, rounding=False)
    lr, gr = oracle.mlp_gradients(p, x, t, rounding=True)
    assert lr == pytest.approx(le, rel=2e-2)
    cos = float(ge @ gr / (np.linalg.norm(ge) * np.linalg.norm(gr)))
    assert cos > 0.999
    # and differs from it (the rounding points are really taken)
    assert np.abs(ge - gr).max() > 0


# ---- residual gate (host copy of the device update, agx_gate_update)

def run_gate(agx, costs, steps=64):
    g = agx.Gate()
    g.use_learned = 1
    used = []
    for _ in range(steps):
        u = int(g.use_learned)
        used.append(u)
        agx.gate_update(g, u, costs[u])
    return g, used


def probe_steps(used, method):
    return [k for k, u in enumerate(used) if k >= 2 and u == method]


def test_gate_prefers_learned_when_cheaper(agx):
    g, used = run_gate(agx, {0: 20.0, 1: 12.0}, steps=200)
    assert used[:2] == [1, 0]             # first step tries the network, then the cached warm start once
    assert g.cost_learned == pytest.approx(12.0) and g.cost_cached == pytest.approx(20.0)
    # afterwards the cached warm start only runs as a probe; each losing probe doubles the period
    probes = probe_steps(used, 0)
    gaps = np.diff([1] + probes)
    assert list(gaps) == [16, 32, 64], probes


def test_gate_falls_back_when_prediction_raises_residual(agx):
    g, used = run_gate(agx, {0: 12.0, 1: 20.0}, steps=200)
    assert g.use_learned == 0
    probes = probe_steps(used, 1)          # the network only runs as a probe, ever more rarely
    assert list(np.diff([1] + probes)) == [16, 32, 64], probes
    assert g.period == 128


def test_gate_winning_probe_resets_period(agx):
    g = agx.Gate()
    g.use_learned = 1
    cost = {0: 12.0, 1: 20.0}
    used = []
    for k in range(120):
        u = int(g.use_learned)
        used.append(u)
        agx.gate_update(g, u, cost[u])
        if k == 60:
            cost[1] = 4.0                   # the network becomes much better
    assert g.use_learned == 1 and g.period == 16


def test_gate_switches_when_learned_degrades(agx):
    g = agx.Gate()
    g.use_learned = 1
    for _ in range(20):
        agx.gate_update(g, int(g.use_learned), 10.0 if g.use_learned else 15.0)
    assert g.use_learned == 1
    # the network starts to cost more than the cached warm start
    switched = None
    for k in range(20):
        agx.gate_update(g, int(g.use_learned), 30.0 if g.use_learned else 15.0)
        if g.use_learned == 0:
            switched = k
            break
    assert switched == 0          # the first learned step costlier than the cached average hands over


def test_gate_noisy_learned_step_triggers_one_cached_step(agx):
    g = agx.Gate()
    g.use_learned = 1
    for _ in range(10):
        agx.gate_update(g, int(g.use_learned), 10.0 if g.use_learned else 14.0)
    while g.use_learned == 0:
        agx.gate_update(g, 0, 14.0)
    agx.gate_update(g, 1, 15.0)        # one bad learned step (its average stays below the cached one)
    assert g.use_learned == 0 and g.cost_learned < g.cost_cached
    agx.gate_update(g, 0, 14.0)
    assert g.use_learned == 1


def test_gate_ties_keep_cached(agx):
    g, used = run_gate(agx, {0: 8.0, 1: 8.0})
    assert g.use_learned == 0
