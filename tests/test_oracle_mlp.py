"""Pins the numpy oracle of the learned-safety network (oracle/mlp.py):
closed-form features, BF16 rounding on known values, and the analytic
gradient against central finite differences in exact float64 arithmetic."""
import numpy as np
import pytest

from oracle import mlp


def test_parameter_count():
    assert mlp.NPARAMS == 51 * 128 + 128 + 3 * (128 * 128 + 128) + 128 + 1 == 56321


def test_features_closed_form():
    f = mlp.features(np.array([[0.0, 0.5, -0.25]]))[0]
    assert f.shape == (51,)
    assert np.array_equal(f[:3], [0.0, 0.5, -0.25])
    # k = 0: sin(pi u), cos(pi u); k = 1: sin(2 pi u) ...
    assert f[3] == 0.0 and f[6] == 1.0                       # axis 0, u = 0
    assert abs(f[4] - 1.0) < 1e-15 and abs(f[7]) < 1e-15     # axis 1, sin(pi/2), cos(pi/2)
    assert abs(f[5] + np.sqrt(0.5)) < 1e-15                  # axis 2, sin(-pi/4)
    assert abs(f[3 + 6 + 1]) < 1e-15 and abs(f[3 + 6 + 4] + 1.0) < 1e-15  # k=1, axis 1: sin(pi), cos(pi)
    assert abs(f[3 + 42 + 2] - np.sin(-0.25 * 128 * np.pi)) < 1e-12       # k=7, axis 2


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 1.0 + 2.0 ** -7, 3.0e38])
    r = mlp.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                    # half-way, ties to even
    assert r[2] == 1.0 + 2.0 ** -6        # half-way, ties to even (up)
    assert r[3] == -2.5
    assert r[4] == 1.0 + 2.0 ** -7        # representable
    assert np.isfinite(r[5])


def test_zero_network():
    p = np.zeros(mlp.NPARAMS)
    u = np.random.default_rng(1).uniform(-1, 1, (32, 3))
    t = np.random.default_rng(2).uniform(0.1, 1.0, 32)
    y, _ = mlp.forward(p, u, emulate=False)
    assert np.all(y == 0.0)
    loss, g = mlp.loss_and_grad(p, u, t, over_weight=4.0, emulate=False)
    assert abs(loss - np.mean(t * t)) < 1e-15          # e = -t < 0: weight 1
    assert abs(g[-1] - np.mean(-2.0 * t)) < 1e-15       # d/db5
    assert np.all(g[:-1 - 128] == 0.0)                  # zero activations' paths: SiLU(0) = 0


@pytest.mark.parametrize("seed", [3, 4])
def test_gradient_finite_differences(seed):
    rng = np.random.default_rng(seed)
    p = mlp.random_params(seed, scale=1.5)
    u = rng.uniform(-1, 1, (12, 3))
    t = rng.uniform(-0.3, 0.3, 12)
    _, g = mlp.loss_and_grad(p, u, t, over_weight=3.0, emulate=False)
    # indices in every block: W1, b1, W2, b2, W3, b3, W4, b4, w5, b5
    offs = np.cumsum([0, 128 * 51, 128, 16384, 128, 16384, 128, 16384, 128, 128, 1])
    idx = []
    for a, b in zip(offs[:-1], offs[1:]):
        idx += list(rng.integers(a, b, size=min(4, b - a)))
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = mlp.loss_and_grad(pp, u, t, 3.0, emulate=False)
        lm, _ = mlp.loss_and_grad(pm, u, t, 3.0, emulate=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_asymmetric_weight():
    """An overestimate of the same size costs over_weight times more."""
    p = np.zeros(mlp.NPARAMS)
    p[-1] = 0.5  # y = b5 = 0.5 everywhere
    u = np.zeros((2, 3))
    l_over, _ = mlp.loss_and_grad(p, u, np.array([0.25, 0.25]), over_weight=4.0, emulate=False)
    l_under, _ = mlp.loss_and_grad(p, u, np.array([0.75, 0.75]), over_weight=4.0, emulate=False)
    assert abs(l_over - 4.0 * l_under) < 1e-15


def test_emulated_close_to_exact():
    p = mlp.random_params(7)
    u = np.random.default_rng(8).uniform(-1, 1, (64, 3))
    y0, _ = mlp.forward(p, u, emulate=False)
    y1, _ = mlp.forward(p, u, emulate=True)
    assert np.max(np.abs(y0 - y1)) < 0.05 * (1 + np.max(np.abs(y0)))
