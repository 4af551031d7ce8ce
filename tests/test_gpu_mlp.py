"""GPU checks of the matrix-core MLP against the CPU oracle (oracle/mlp_oracle.py), which
rounds to BF16 at the same points: the MFMA fragment layout (exact), forward outputs and
backward gradients (FP32 accumulation order only), and that training reduces the loss.

Tolerances: the exact-layout case is compared bit for bit; elsewhere 5e-3 relative in the
2-norm, the size of one BF16 rounding of an activation that the FP32 accumulation order
can flip."""
import numpy as np
import pytest
import torch

import pgfem3d_amd as P
from mlp_oracle import MLPRef, pad16

pytestmark = pytest.mark.gpu
WIDTHS = (34, 128, 128, 128, 128, 12)


def _cuda(x, device):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_mfma_layout_exact(device):
    """One 48 -> 32 linear layer with a permutation-and-scale weight: every output is one
    input times a power of two, exact in BF16, so any fragment mix-up shows as a mismatch."""
    widths = (40, 30)
    net = P.MLP(widths, seed=0)
    K, N = pad16(40), pad16(30)
    W = np.zeros((N, K), dtype=np.float32)
    perm = np.random.default_rng(0).permutation(40)[:30]
    for n in range(30):
        W[n, perm[n]] = 2.0 ** (n % 5 - 2)
    b = np.zeros(N, dtype=np.float32)
    b[:30] = np.arange(30)
    net.set_params(np.concatenate([W.ravel(), b]))
    X = np.random.default_rng(1).integers(-64, 64, size=(77, 40)).astype(np.float32)
    Y = torch.empty((77, 30), dtype=torch.float32, device=device)
    net.forward(_cuda(X, device), Y)
    expect = X[:, perm] * (2.0 ** (np.arange(30) % 5 - 2)) + np.arange(30)
    np.testing.assert_array_equal(Y.cpu().numpy(), expect.astype(np.float32))
    net.close()


def test_forward_and_gradients_match_oracle(device):
    rng = np.random.default_rng(3)
    net = P.MLP(WIDTHS, seed=11)
    p = net.get_params()
    ref = MLPRef(WIDTHS, p)
    B = 320
    X = rng.standard_normal((B, 34)).astype(np.float32)
    T = rng.standard_normal((B, 12)).astype(np.float32)
    Y = torch.empty((B, 12), dtype=torch.float32, device=device)
    net.forward(_cuda(X, device), Y)
    assert _rel(Y.cpu().numpy(), ref.forward(X)) < 5e-3
    loss = net.backward(_cuda(X, device), _cuda(T, device))
    g = net.get_grads()
    loss_ref, g_ref = ref.loss_and_grads(X, T)
    assert abs(loss - loss_ref) < 5e-3 * loss_ref
    off = 0
    for l in range(len(WIDTHS) - 1):  # every layer's weights and bias
        n = pad16(WIDTHS[l + 1]) * pad16(WIDTHS[l]) + pad16(WIDTHS[l + 1])
        assert _rel(g[off:off + n], g_ref[off:off + n]) < 5e-3, l
        off += n
    # an Adam step consumes the gradient and moves the parameters by about lr
    net.adam_step(1e-3)
    assert np.abs(net.get_grads()).max() == 0.0
    d = net.get_params() - p
    assert 0.5e-3 < np.abs(d).max() <= 1.01e-3
    net.close()


def test_training_reduces_loss(device):
    """A smooth 34 -> 12 map is learned by the library's own backward and Adam kernels."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((34, 12)) / 6
    X = rng.standard_normal((16384, 34)).astype(np.float32)
    T = np.tanh(X @ A).astype(np.float32)
    net = P.MLP(WIDTHS, seed=1)
    Xd, Td = _cuda(X, device), _cuda(T, device)
    first = net.backward(Xd[:4096], Td[:4096])
    net.adam_step(1e-3)
    for ep in range(60):
        for s in range(0, 16384, 4096):
            loss = net.backward(Xd[s:s + 4096], Td[s:s + 4096])
            net.adam_step(1e-3)
    assert loss < 0.05 * first
    net.close()


def test_fused_inference_layout_exact(device):
    """The fused inference kernel (activations in registers, the transposed product) on the
    permutation-and-scale layer: exact, as the layered forward is; and on a two-layer net with
    an odd batch size its outputs equal the layered forward's."""
    widths = (40, 30)
    net = P.MLP(widths, seed=0)
    K, N = pad16(40), pad16(30)
    W = np.zeros((N, K), dtype=np.float32)
    perm = np.random.default_rng(0).permutation(40)[:30]
    for n in range(30):
        W[n, perm[n]] = 2.0 ** (n % 5 - 2)
    b = np.zeros(N, dtype=np.float32)
    b[:30] = np.arange(30)
    net.set_params(np.concatenate([W.ravel(), b]))
    X = np.random.default_rng(1).integers(-64, 64, size=(300, 40)).astype(np.float32)
    Y = torch.empty((300, 30), dtype=torch.float32, device=device)
    net.infer(_cuda(X, device), Y)
    expect = X[:, perm] * (2.0 ** (np.arange(30) % 5 - 2)) + np.arange(30)
    np.testing.assert_array_equal(Y.cpu().numpy(), expect.astype(np.float32))
    net.close()


def test_fused_inference_matches_layered_forward(device):
    rng = np.random.default_rng(5)
    net = P.MLP(WIDTHS, seed=7)
    ref = MLPRef(WIDTHS, net.get_params())
    for B in (1, 37, 1000):
        X = rng.standard_normal((B, 34)).astype(np.float32)
        Y1 = torch.empty((B, 12), dtype=torch.float32, device=device)
        Y2 = torch.empty((B, 12), dtype=torch.float32, device=device)
        net.forward(_cuda(X, device), Y1)
        net.infer(_cuda(X, device), Y2)
        assert _rel(Y2.cpu().numpy(), ref.forward(X)) < 5e-3
        # same products in the same k order: equal up to the MFMA's own summation
        assert _rel(Y2.cpu().numpy(), Y1.cpu().numpy().astype(np.float64)) < 1e-6, B
    net.close()
