"""Candidate scorer: the CPU reference pinned by finite differences, and the
device's MFMA training kernels checked against it."""
import numpy as np
import pytest

import mlp_reference as ref


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = (rng.random((n, ref.IN)) < 0.1).astype(np.float32) + rng.normal(0, 0.1, (n, ref.IN)).astype(np.float32)
    y = (X[:, :8].sum(1) * 0.3 + 0.1 * rng.normal(size=n)).astype(np.float32)
    return X, y


def test_reference_gradient_matches_finite_differences():
    # FP64 central differences of the FP32 reference's loss at random entries
    # of every parameter block.
    X, y = _data(32, 1)
    p = ref.init_params(2).astype(np.float64)
    _, g = ref.loss_and_grads(p.astype(np.float32), X, y, bf16=False)

    def loss64(q):
        W1 = q[ref.OFF_W1:ref.OFF_B1].reshape(ref.IN_PAD, ref.HID)
        b1 = q[ref.OFF_B1:ref.OFF_W2]
        W2 = q[ref.OFF_W2:ref.OFF_B2].reshape(ref.HID, ref.HID)
        b2 = q[ref.OFF_B2:ref.OFF_W3]
        w3 = q[ref.OFF_W3:ref.OFF_B3]
        b3 = q[ref.OFF_B3]
        Xp = np.zeros((X.shape[0], ref.IN_PAD))
        Xp[:, :ref.IN] = X
        H1 = np.maximum(Xp @ W1 + b1, 0)
        H2 = np.maximum(H1 @ W2 + b2, 0)
        return np.mean((H2 @ w3 + b3 - y) ** 2)

    rng = np.random.default_rng(3)
    blocks = [(ref.OFF_W1, ref.IN * ref.HID), (ref.OFF_B1, ref.HID), (ref.OFF_W2, ref.HID * ref.HID),
              (ref.OFF_B2, ref.HID), (ref.OFF_W3, ref.HID), (ref.OFF_B3, 1)]
    h = 1e-5
    for off, size in blocks:
        for i in rng.integers(0, size, 6):
            q = p.copy()
            q[off + i] += h
            lp = loss64(q)
            q[off + i] -= 2 * h
            lm = loss64(q)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[off + i]) <= 2e-3 * max(1.0, abs(fd)), (off, i, fd, g[off + i])


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.0e-3], dtype=np.float32)
    # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1.0)
    out = ref.round_bf16(vals)
    assert out[0] == 1.0 and out[1] == 1.0 and out[2] == np.float32(1.0078125) and out[3] == np.float32(1.0078125)
    assert out[4] == -2.5
    assert abs(out[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


@pytest.mark.gpu
def test_device_scores_match_reference(gpu_context):
    X, _ = _data(1000, 4)
    p = ref.init_params(5)
    s = gpu_context.mlp_score(p, X)
    s_ref, _ = ref.forward(p, X, bf16=True)
    # same BF16 operands; FP32 sums in a different order
    assert np.allclose(s, s_ref, rtol=2e-3, atol=2e-3 * np.abs(s_ref).max())


@pytest.mark.gpu
def test_device_gradients_match_reference(gpu_context):
    X, y = _data(777, 6)   # not a multiple of 64: padded rows must not count
    p = ref.init_params(7)
    loss, g = gpu_context.mlp_gradients(p, X, y)
    loss_ref, g_ref = ref.loss_and_grads(p, X, y, bf16=True)
    assert abs(loss - loss_ref) <= 1e-3 * abs(loss_ref)
    blocks = {"W1": (ref.OFF_W1, ref.OFF_B1), "b1": (ref.OFF_B1, ref.OFF_W2), "W2": (ref.OFF_W2, ref.OFF_B2),
              "b2": (ref.OFF_B2, ref.OFF_W3), "w3": (ref.OFF_W3, ref.OFF_B3), "b3": (ref.OFF_B3, ref.PARAMS)}
    for name, (a, b) in blocks.items():
        d, r = g[a:b], g_ref[a:b]
        # BF16 activations may round the other way where the FP32 sums differ
        # in their last bits: 1% of the block's norm
        assert np.linalg.norm(d - r) <= 1e-2 * np.linalg.norm(r) + 1e-7, name
    # and close to the pure FP32 gradient: 5% of the norm
    _, g32 = ref.loss_and_grads(p, X, y, bf16=False)
    assert np.linalg.norm(g - g32) <= 5e-2 * np.linalg.norm(g32)


@pytest.mark.gpu
def test_device_training_reduces_loss(gpu_context):
    X, y = _data(8192, 8)
    p0 = ref.init_params(9)
    loss0, _ = ref.loss_and_grads(p0, X, y, bf16=True)
    p, losses = gpu_context.mlp_train(X, y, p0, epochs=6, batch=512, lr=1e-3, seed=3)
    loss1, _ = ref.loss_and_grads(p, X, y, bf16=True)
    assert losses[-1] < losses[0]
    assert loss1 < 0.2 * loss0
    assert np.all(p[ref.OFF_W1 + ref.IN * ref.HID:ref.OFF_B1] == 0)   # padded rows of W1 stay zero
