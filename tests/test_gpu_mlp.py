"""GPU: the policy MLP's hand-written HIP kernels against the CPU oracle.

Tolerances: fp32 device arithmetic vs the fp64 oracle, |a-b| <= 1e-4 + 1e-3|b|
for gradients and 1e-4 + 1e-4|b| for fp32 logits; the bf16 inference path is
compared with the oracle evaluated on the bf16-rounded weights, 2e-2 + 2e-2|b|
(bf16 keeps 8 significant bits)."""
import numpy as np
import pytest

import oracle_lib
import tensorfim_amd as tf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = tf.Context(0)
    yield c
    c.close()


def _batch(rng, n):
    X = rng.normal(size=(n, 14)).astype(np.float32)
    Y = (rng.random((n, 2)) < 0.5).astype(np.float32)
    M = (rng.random((n, 2)) < 0.7).astype(np.float32)
    return X, Y, M


def _bf16_round(p):
    u = p.astype(np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("n", [1, 5, 257, 4000])
def test_gradients_match_oracle(ctx, n):
    rng = np.random.default_rng(n)
    mlp = tf.Mlp(ctx, seed=7)
    P = mlp.params()
    X, Y, M = _batch(rng, n)
    g, loss = mlp.gradients(X, Y, M)
    l_ref, _, g_ref = oracle_lib.mlp_loss_grad(P.astype(np.float64), X, Y, M)
    assert abs(loss - l_ref) <= 1e-4 + 1e-4 * abs(l_ref)
    np.testing.assert_allclose(g, g_ref, rtol=1e-3, atol=1e-4)
    mlp.close()


def test_forward_fp32_and_bf16(ctx):
    rng = np.random.default_rng(3)
    mlp = tf.Mlp(ctx, seed=11)
    P = mlp.params()
    X, Y, M = _batch(rng, 999)
    _, ref, _ = oracle_lib.mlp_loss_grad(P.astype(np.float64), X, Y, M, False)
    np.testing.assert_allclose(mlp.forward(X), ref, rtol=1e-4, atol=1e-4)
    _, ref_b, _ = oracle_lib.mlp_loss_grad(_bf16_round(P).astype(np.float64), X, Y, M, False)
    np.testing.assert_allclose(mlp.forward(X, bf16=True), ref_b, rtol=2e-2, atol=2e-2)
    mlp.close()


def test_training_learns_a_separable_rule(ctx):
    rng = np.random.default_rng(5)
    X = rng.normal(size=(2048, 14)).astype(np.float32)
    Y = np.stack([(X[:, 0] + 0.5 * X[:, 3] > 0), (X[:, 7] < 0.2)], axis=1).astype(np.float32)
    M = np.ones_like(Y)
    mlp = tf.Mlp(ctx, seed=1)
    first = mlp.train_step(X, Y, M, lr=3e-3)
    for _ in range(300):
        last = mlp.train_step(X, Y, M, lr=3e-3)
    assert last < 0.25 * first
    acc = ((mlp.forward(X, bf16=True) > 0) == (Y > 0.5)).mean()
    assert acc > 0.95
    mlp.close()


def test_forward_submit_wait_orders_with_training(ctx):
    """An inference queued before a training step sees the weights before the
    step; one queued after sees the updated ones; both match the synchronous
    forward at those weights."""
    rng = np.random.default_rng(3)
    mlp = tf.Mlp(ctx, seed=11)
    X = rng.normal(size=(300, 14)).astype(np.float32)
    before = mlp.forward(X)
    mlp.forward_submit(X)
    Xt, Yt, Mt = _batch(rng, 2048)
    mlp.train_step(Xt, Yt, Mt, lr=1e-2)
    got_before = mlp.forward_wait()
    np.testing.assert_array_equal(got_before, before)
    mlp.forward_submit(X[:17])
    got_after = mlp.forward_wait()
    np.testing.assert_array_equal(got_after, mlp.forward(X[:17]))
    assert np.abs(got_after - before[:17]).max() > 0   # the step changed the weights
    with pytest.raises(tf.TfimError):
        mlp.forward_wait()                            # nothing outstanding
    mlp.close()


def test_multi_step_training_equals_single_steps(ctx):
    """tfim_mlp_train(steps=4), which stages the batch once, leaves the same
    parameters as four tfim_mlp_train_step calls on that batch (same kernels
    in the same order), and reports the loss the last step saw."""
    rng = np.random.default_rng(21)
    X, Y, M = _batch(rng, 3000)
    a, b = tf.Mlp(ctx, seed=4), tf.Mlp(ctx, seed=4)
    for _ in range(4):
        la = a.train_step(X, Y, M, lr=5e-3)
    lb = b.train(X, Y, M, lr=5e-3, steps=4)
    np.testing.assert_array_equal(a.params(), b.params())
    assert la == lb
    a.close()
    b.close()
