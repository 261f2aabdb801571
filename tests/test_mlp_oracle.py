"""CPU: the policy-MLP oracle's gradients are pinned by central finite
differences of its own loss (the reference has no learned component, so the
mathematics is the only anchor), and the loss has its closed form at zero
weights."""
import numpy as np

import oracle_lib


def _batch(rng, n=23):
    X = rng.normal(size=(n, 14)).astype(np.float32)
    Y = (rng.random((n, 2)) < 0.5).astype(np.float32)
    M = (rng.random((n, 2)) < 0.7).astype(np.float32)
    return X, Y, M


def _params(rng):
    P = np.zeros(5250)
    P[:896] = rng.uniform(-0.65, 0.65, 896)             # W1 (fan-in 14)
    P[896:960] = rng.uniform(-0.1, 0.1, 64)
    P[960:5056] = rng.uniform(-0.3, 0.3, 4096)          # W2
    P[5056:5120] = rng.uniform(-0.1, 0.1, 64)
    P[5120:5248] = rng.uniform(-0.3, 0.3, 128)          # W3
    P[5248:] = rng.uniform(-0.1, 0.1, 2)
    return P


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    X, Y, M = _batch(rng)
    P = _params(rng)
    _, _, g = oracle_lib.mlp_loss_grad(P, X, Y, M)
    eps = 1e-6
    idx = list(rng.choice(5250, size=60, replace=False)) + [0, 895, 896, 959, 960, 5055, 5120, 5248, 5249]
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += eps
        Pm[i] -= eps
        fd = (oracle_lib.mlp_loss_grad(Pp, X, Y, M, False)[0] - oracle_lib.mlp_loss_grad(Pm, X, Y, M, False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_oracle_loss_closed_form_at_zero():
    rng = np.random.default_rng(1)
    X, Y, M = _batch(rng)
    loss, logits, g = oracle_lib.mlp_loss_grad(np.zeros(5250), X, Y, M)
    assert np.allclose(logits, 0.0)
    assert abs(loss - np.log(2.0)) < 1e-12          # every masked output contributes log 2, averaged
    # only the output biases move at zero weights: dL/db3[o] = mean over masked of (0.5 - y)
    exp_b3 = [((0.5 - Y[:, o]) * M[:, o]).sum() / M.sum() for o in range(2)]
    assert np.allclose(g[5248:], exp_b3)
    assert np.allclose(g[:5248], 0.0)
