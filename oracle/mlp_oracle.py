"""CPU oracle of the warm-start MLP -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the library's MLP (pgfem-3d_amd/csrc/pgf_mlp.hip): layers with
zero-padded widths (multiples of 16), exact GELU after every layer but the last, BF16
operands with wide accumulation, mean-squared-error loss, and its backward pass.  With
``quantize=True`` it rounds at the same points as the kernels (inputs, weights, hidden
activations, loss gradient and back-propagated pre-activation gradients to BF16, round to
nearest even; pre-activations kept in FP32), so the GPU results agree up to the order of
FP32 accumulation.  With ``quantize=False`` everything is FP64, which is what the
finite-difference pin of the backward formulas uses (tests/test_oracle_mlp.py).
"""
import math

import numpy as np


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


_erf = np.vectorize(math.erf)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def pad16(n):
    return ((n + 15) // 16) * 16


class MLPRef:
    def __init__(self, widths, params):
        self.widths = list(widths)
        self.L = len(widths) - 1
        self.K = [pad16(w) for w in widths[:-1]]
        self.N = [pad16(w) for w in widths[1:]]
        self.W, self.b = [], []
        off = 0
        for l in range(self.L):
            n = self.N[l] * self.K[l]
            self.W.append(np.asarray(params[off:off + n], dtype=np.float64).reshape(self.N[l], self.K[l]))
            off += n
            self.b.append(np.asarray(params[off:off + self.N[l]], dtype=np.float64))
            off += self.N[l]
        self.n_param = off

    def _q(self, x, quantize):
        return bf16(x).astype(np.float64) if quantize else np.asarray(x, dtype=np.float64)

    def forward(self, X, quantize=True):
        B = X.shape[0]
        H = np.zeros((B, self.K[0]))
        H[:, :X.shape[1]] = X
        Hs, Zs = [self._q(H, quantize)], [None]
        for l in range(self.L):
            Z = Hs[-1] @ self._q(self.W[l], quantize).T
            if quantize:
                Z = Z.astype(np.float32).astype(np.float64)
            Z = Z + self.b[l]
            if quantize:
                Z = Z.astype(np.float32).astype(np.float64)
            if l + 1 < self.L:
                Zs.append(Z)
                Hs.append(self._q(gelu(Z), quantize))
            else:
                Y = Z[:, :self.widths[-1]]
        self._cache = (Hs, Zs)
        return Y

    def loss_and_grads(self, X, T, quantize=True):
        """Mean squared error over all outputs and its gradient, in the library's layout."""
        Y = self.forward(X, quantize)
        Hs, Zs = self._cache
        B, no = T.shape
        e = Y - T
        loss = float(np.sum(e * e) / (B * no))
        dZ = np.zeros((B, self.N[-1]))
        dZ[:, :no] = 2.0 * e / (B * no)
        dZ = self._q(dZ, quantize)
        gW, gb = [None] * self.L, [None] * self.L
        for l in range(self.L - 1, -1, -1):
            gW[l] = dZ.T @ Hs[l]
            gb[l] = dZ.sum(axis=0)
            if l == 0:
                break
            dH = dZ @ self._q(self.W[l], quantize)
            dZ = self._q(dH * gelu_d(Zs[l]), quantize)
        g = np.concatenate([np.concatenate([gW[l].ravel(), gb[l]]) for l in range(self.L)])
        return loss, g

    def set_flat(self, p):
        self.__init__(self.widths, p)
