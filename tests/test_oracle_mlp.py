"""Pins of the MLP oracle (CPU): BF16 rounding known answers and the backward formulas
against central finite differences of the loss (FP64, no rounding)."""
import numpy as np
import pytest

from mlp_oracle import MLPRef, bf16, pad16


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 3.0e38], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins; 1 + 3*2^-8 ties up
    expect = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, 2.9975145e38], dtype=np.float32)
    np.testing.assert_array_equal(bf16(x)[:5], expect[:5])
    assert abs(bf16(x)[5] - expect[5]) / expect[5] < 2 ** -8


def _random_net(widths, seed):
    rng = np.random.default_rng(seed)
    K = [pad16(w) for w in widths[:-1]]
    N = [pad16(w) for w in widths[1:]]
    p = []
    for l in range(len(widths) - 1):
        W = np.zeros((N[l], K[l]))
        W[:widths[l + 1], :widths[l]] = rng.standard_normal((widths[l + 1], widths[l])) / np.sqrt(widths[l])
        b = np.zeros(N[l])
        b[:widths[l + 1]] = 0.1 * rng.standard_normal(widths[l + 1])
        p += [W.ravel(), b]
    return np.concatenate(p)


@pytest.mark.parametrize("widths", [(5, 16, 3), (34, 24, 20, 12)])
def test_backward_is_fd_of_loss(widths):
    rng = np.random.default_rng(1)
    p = _random_net(widths, 2)
    net = MLPRef(widths, p)
    X = rng.standard_normal((7, widths[0]))
    T = rng.standard_normal((7, widths[-1]))
    loss, g = net.loss_and_grads(X, T, quantize=False)
    live = np.flatnonzero(p != 0.0)  # padded entries are structurally zero
    pick = rng.choice(live, size=min(40, live.size), replace=False)
    h = 1e-6
    for i in pick:
        q = p.copy()
        q[i] += h
        lp, _ = MLPRef(widths, q).loss_and_grads(X, T, quantize=False)
        q[i] -= 2 * h
        lm, _ = MLPRef(widths, q).loss_and_grads(X, T, quantize=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_forward_closed_form_single_layer():
    """One linear layer: Y = X W^T + b exactly (inputs and weights representable in BF16)."""
    widths = (3, 2)
    W = np.zeros((16, 16))
    W[0, :3] = [1.0, -2.0, 0.5]
    W[1, :3] = [0.25, 0.0, 4.0]
    b = np.zeros(16)
    b[:2] = [1.0, -1.0]
    net = MLPRef(widths, np.concatenate([W.ravel(), b]))
    X = np.array([[1.0, 2.0, 3.0], [-1.0, 0.5, 0.0]])
    Y = net.forward(X)
    np.testing.assert_array_equal(Y, [[1.0 - 4.0 + 1.5 + 1.0, 0.25 + 12.0 - 1.0], [-1.0 - 1.0 + 0 + 1.0, -0.25 - 1.0]])
