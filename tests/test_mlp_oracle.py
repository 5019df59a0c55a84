"""Pin the oracle's FP64 surrogate network: its back-propagated gradient
against central finite differences of its own loss, and Adam's first steps
against the closed form of Kingma & Ba (2015), Algorithm 1."""
import numpy as np

from oracle import pyoracle as po


def _data(seed=0, n=40):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.4, 1249)
    X = rng.uniform(0, 1.5, (n, 4))
    Y = rng.uniform(0, 1, n)
    return w, X, Y


def test_mlp_gradient_matches_finite_differences():
    w, X, Y = _data()
    g, _ = po.mlp_loss_grad(w, X, Y)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(128, 6, replace=False), 128 + rng.choice(32, 3),
                          160 + rng.choice(1024, 10, replace=False), 1184 + rng.choice(32, 3),
                          1216 + rng.choice(32, 4), [1248]])
    eps = 1e-6
    for p in idx:
        wp, wm = w.copy(), w.copy()
        wp[p] += eps
        wm[p] -= eps
        fd = (po.mlp_loss_grad(wp, X, Y)[1] - po.mlp_loss_grad(wm, X, Y)[1]) / (2 * eps)
        assert abs(fd - g[p]) <= 1e-6 * max(1.0, abs(g[p])), p


def test_mlp_forward_is_the_network():
    w, X, _ = _data(3, 5)
    W1 = w[:128].reshape(4, 32)
    b1 = w[128:160]
    W2 = w[160:1184].reshape(32, 32)
    b2 = w[1184:1216]
    W3 = w[1216:1248]
    b3 = w[1248]
    for x in X:
        h1 = np.maximum(x @ W1 + b1, 0)
        h2 = np.maximum(h1 @ W2 + b2, 0)
        assert abs(po.mlp_forward(w, x) - (h2 @ W3 + b3)) < 1e-12


def adam_reference(w, grads, lr):
    """Kingma & Ba (2015) Algorithm 1 with beta1 0.9, beta2 0.999, eps 1e-8."""
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    for t, g in enumerate(grads, start=1):
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        w = w - lr * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
    return w


def test_adam_first_step_is_lr_times_sign():
    # after one step mhat = g and vhat = g^2, so the update is lr * g/(|g| + eps)
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    w1 = adam_reference(w, [g], 0.01)
    assert np.allclose(w - w1, 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)
