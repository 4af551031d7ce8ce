"""Learned warm start of the crystal-plasticity local Newton: data logging, normalisation and
training with the library's own MLP kernels (pgf_mlp_*), and attachment to a handle.

The network maps the 34 features of a QP -- F(n) pF(n)^-1, F(n+1) - F(n), the 12 resolved
shear stresses over the slip resistance tau_a(n)/g(n) (the model has one slip resistance per
QP, whose ratio to each system's stress is what the flow rule reads) and the grain's
orientation quaternion -- to its 12 slip increments
dt * gamma_dot(n+1).  The library turns the prediction into a start (M, lambda, g) of the FP64
local Newton, which still converges to tol_M and falls back to the plain start when the
predicted start does not converge.  torch is used only to hold the data on the GPU and to
shuffle and slice it; every forward, backward and optimiser step runs in the library.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

import pgfem3d_amd as P

N_FEAT, N_OUT = 34, 12
WIDTHS = (N_FEAT, 128, 128, 128, 128, N_OUT)


@dataclass
class Normalisation:
    shift: np.ndarray      # [34]
    scale: np.ndarray      # [34]
    out_scale: np.ndarray  # [12]


def log_step(h: P.Handle, X_all, T_all, row0: int) -> int:
    """Append the features/targets of the handle's last integration at row row0 of the
    preallocated float32 GPU tensors; returns the next free row."""
    import torch

    X = torch.empty((h.n_qp, N_FEAT), dtype=torch.float32, device=X_all.device)
    T = torch.empty((h.n_qp, N_OUT), dtype=torch.float32, device=X_all.device)
    h.log_samples(X, T)
    X_all[row0:row0 + h.n_qp] = X
    T_all[row0:row0 + h.n_qp] = T
    return row0 + h.n_qp


def normalisation(X, T) -> Normalisation:
    """Per-feature shift/scale to zero mean and unit spread; constant features keep scale 1.
    Targets are scaled by one common factor (the RMS of all slip increments), so that the
    prediction of every system keeps its physical proportion."""
    mu = X.mean(dim=0)
    sd = X.std(dim=0)
    sd = sd.where(sd > 1e-12 * (mu.abs() + 1.0), sd.new_ones(()))
    rms = float(T.pow(2).mean().sqrt())
    out = np.full(N_OUT, 1.0 / rms if rms > 0 else 1.0, dtype=np.float32)
    return Normalisation(mu.cpu().numpy().astype(np.float32), (1.0 / sd).cpu().numpy().astype(np.float32), out)


def train(X, T, epochs=6, batch=8192, lr=2e-3, seed=0, split=0.9, widths=WIDTHS, verbose=False):
    """Train on raw logged samples (float32 GPU tensors) with a seeded train/validation split.
    Returns (mlp, normalisation, history) with history = list of (epoch, train, val loss)."""
    import torch

    nrm = normalisation(X, T)
    dev = X.device
    sh = torch.from_numpy(nrm.shift).to(dev)
    sc = torch.from_numpy(nrm.scale).to(dev)
    Xn = ((X - sh) * sc).contiguous()
    Tn = (T * float(nrm.out_scale[0])).contiguous()
    g = torch.Generator(device="cpu").manual_seed(seed)
    perm = torch.randperm(X.shape[0], generator=g).to(dev)
    n_tr = (int(split * X.shape[0]) // batch) * batch
    tr, va = perm[:n_tr], perm[n_tr:]
    n_va = (va.numel() // 16) * 16
    va = va[:n_va]
    mlp = P.MLP(widths, seed=seed)
    hist = []
    steps_per_epoch = n_tr // batch
    total = epochs * steps_per_epoch
    k = 0
    for ep in range(epochs):
        t0 = time.perf_counter()
        order = tr[torch.randperm(n_tr, generator=g).to(dev)]
        run = 0.0
        for s in range(steps_per_epoch):
            idx = order[s * batch:(s + 1) * batch]
            xb = Xn.index_select(0, idx).contiguous()
            tb = Tn.index_select(0, idx).contiguous()
            run += mlp.backward(xb, tb)
            # cosine decay of the learning rate over the whole run
            mlp.adam_step(lr * 0.5 * (1.0 + np.cos(np.pi * k / max(1, total))))
            k += 1
        val = validation_loss(mlp, Xn, Tn, va)
        hist.append((ep, run / max(1, steps_per_epoch), val))
        if verbose:
            print(f"[surrogate] epoch {ep}: train {hist[-1][1]:.4e} val {val:.4e} "
                  f"({time.perf_counter() - t0:.1f}s)", flush=True)
    return mlp, nrm, hist


def validation_loss(mlp: P.MLP, Xn, Tn, idx, chunk=65536) -> float:
    import torch

    if idx.numel() == 0:
        return float("nan")
    tot = 0.0
    for s in range(0, idx.numel(), chunk):
        j = idx[s:s + chunk]
        xb = Xn.index_select(0, j).contiguous()
        yb = torch.empty((j.numel(), N_OUT), dtype=torch.float32, device=Xn.device)
        mlp.forward(xb, yb)
        tot += float((yb - Tn.index_select(0, j)).pow(2).sum())
    return tot / (idx.numel() * N_OUT)


def attach(h: P.Handle, mlp: P.MLP, nrm: Normalisation):
    h.set_surrogate(mlp, nrm.shift, nrm.scale, nrm.out_scale)
