"""Pins of the oracle's cohesive interface element (exponential traction-separation law)."""
import numpy as np
import pytest

import oracle as O


class Props:
    def __init__(self, sig_c=500.0, chi_c=1e-3, beta=0.5, alpha=1.0):
        self.sig_c, self.chi_c, self.beta, self.alpha = sig_c, chi_c, beta, alpha


def _unit(v):
    return v / np.linalg.norm(v)


def test_pure_opening_envelope_peaks_at_sig_c():
    """t(X n) = sig_c (X/chi_c) exp(1 - X/chi_c) n on the loading envelope; max sig_c at chi_c."""
    p = Props()
    n = _unit(np.array([0.3, -0.2, 0.9]))
    for X in (0.1e-3, 0.5e-3, 1e-3, 2.5e-3):
        t, rc = O.coh_traction(p, X * n, n)
        expect = p.sig_c * X / p.chi_c * np.exp(1 - X / p.chi_c)
        np.testing.assert_allclose(t, expect * n, rtol=1e-14, atol=0)
    t, _ = O.coh_traction(p, p.chi_c * n, n)
    assert abs(np.linalg.norm(t) - p.sig_c) < 1e-12 * p.sig_c


def test_unloading_is_secant_and_damaged_element_is_killed():
    p = Props()
    n = np.array([0.0, 0.0, 1.0])
    Xmax = 2.5e-3
    tmax = p.sig_c * Xmax / p.chi_c * np.exp(1 - Xmax / p.chi_c)
    t, rc = O.coh_traction(p, 1e-3 * n, n, (Xmax, tmax))
    np.testing.assert_allclose(t, tmax / Xmax * 1e-3 * n, rtol=1e-14)
    # far on the softening branch t_hat = |t|/X drops below sig_c/100 (per unit opening, as
    # the reference compares it): reported as killed
    t, rc = O.coh_traction(p, 15e-3 * n, n)
    assert rc == 100
    t, rc = O.coh_traction(p, 8e-3 * n, n)
    assert rc == 0


@pytest.mark.parametrize("case", ["opening", "unloading", "contact"])
def test_tangents_are_fd_of_the_traction(case):
    rng = np.random.default_rng(3)
    p = Props(beta=0.7, alpha=2.0)
    n = _unit(rng.standard_normal(3))
    shear = rng.standard_normal(3)
    shear -= shear.dot(n) * n
    jn = {"opening": 0.8e-3, "unloading": 0.6e-3, "contact": -0.3e-3}[case]
    jump = jn * n + 0.4e-3 * _unit(shear)
    vars_n = (0.0, 0.0) if case != "unloading" else (2e-3, 300.0)
    if case == "contact":
        vars_n = (0.4e-3, 0.0)
    mt, gt = O.coh_tangents(p, jump, n, vars_n)
    h = 1e-10
    fd = np.zeros((3, 3))
    for j in range(3):
        d = np.zeros(3)
        d[j] = h
        fd[:, j] = (O.coh_traction(p, jump + d, n, vars_n)[0] - O.coh_traction(p, jump - d, n, vars_n)[0]) / (2 * h)
    extra_m = extra_g = np.zeros((3, 3))
    if case == "unloading":
        # on the secant branch t_hat = t_max/X_max is constant, but the reference's tangent
        # (Needleman_compute_dt_hat) uses d t_hat/d X_eff = t_max/X_max: the only difference
        # from the exact derivative is that term, restated as the reference has it
        b2 = p.beta ** 2
        Xn = jump @ n
        Cs = jump - Xn * n
        Xeff = np.sqrt(b2 * Cs @ Cs + Xn * Xn)
        dth = vars_n[1] / vars_n[0]
        dCdX = (b2 * jump + (1 - b2) * Xn * n) / Xeff
        dCdN = (1 - b2) * Xn / Xeff * jump
        extra_m = dth * np.outer(b2 * Cs + Xn * n, dCdX)
        extra_g = dth * np.outer(b2 * Cs + Xn * n, dCdN)
    assert np.abs(mt - extra_m - fd).max() < 1e-6 * np.abs(mt).max()
    # geometric tangent: derivative wrt the normal along directions tangent to the unit sphere
    for v in (_unit(np.cross(n, [1.0, 0, 0])), _unit(np.cross(n, [0, 1.0, 0]))):
        hn = 1e-7
        tp = O.coh_traction(p, jump, _unit(n + hn * v), vars_n)[0]
        tm = O.coh_traction(p, jump, _unit(n - hn * v), vars_n)[0]
        d_fd = (tp - tm) / (2 * hn)
        assert np.abs((gt - extra_g) @ v - d_fd).max() < 1e-6 * np.abs(gt).max()


def _pair(nn):
    """A flat interface at z = 0 whose current edge vectors are unit when the two sides move
    antisymmetrically: the unit-leg triangle or the side-2 square (Gauss points map with unit
    tangent vectors)."""
    if nn == 3:
        base = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
    else:
        base = np.array([[0.0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0]])
    return np.vstack([base, base])


@pytest.mark.parametrize("nn", [3, 4])
def test_uniform_opening_total_force_is_traction_times_area(nn):
    p = Props(beta=0.5)
    x0 = _pair(nn)
    area = 0.5 if nn == 3 else 4.0
    X = 0.7e-3
    u = np.zeros((2 * nn, 3))
    u[nn:, 2] = X  # + side lifted
    fe, K, v1, err = O.coh_element(p, x0, u)
    assert err == 0
    tz = p.sig_c * X / p.chi_c * np.exp(1 - X / p.chi_c)
    fe = fe.reshape(2 * nn, 3)
    np.testing.assert_allclose(fe[nn:].sum(0), [0, 0, tz * area], rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(fe[:nn].sum(0), [0, 0, -tz * area], rtol=1e-13, atol=1e-12)
    # committed state: (|Xn|, |t|) at every QP
    np.testing.assert_allclose(v1[:, 0], X, rtol=1e-14)
    np.testing.assert_allclose(v1[:, 1], tz, rtol=1e-13)
    # a closed interface carries nothing
    fe0, _, v0, _ = O.coh_element(p, x0, np.zeros((2 * nn, 3)))
    assert np.abs(fe0).max() == 0.0 and np.abs(v0).max() == 0.0


@pytest.mark.parametrize("nn", [3, 4])
@pytest.mark.parametrize("normal_open", [0.6e-3, -0.2e-3])
def test_element_tangent_is_fd_of_residual(nn, normal_open):
    """K = d fe / d u where the reference's dn/dxb is exact (unit current edge vectors):
    antisymmetric side displacements keep the mean surface at the reference one.  Loading
    (opening) and contact branches."""
    rng = np.random.default_rng(nn)
    p = Props(beta=0.6, alpha=2.0)
    x0 = _pair(nn)
    d = np.zeros((nn, 3))
    d[:, 2] = normal_open * (1 + 0.3 * rng.random(nn))
    d[:, :2] = 0.3e-3 * rng.standard_normal((nn, 2))
    u = np.vstack([-0.5 * d, 0.5 * d]).reshape(-1)
    vars_n = np.zeros((nn, 2))
    if normal_open < 0:
        vars_n[:, 0] = 0.3e-3
    fe, K, _, err = O.coh_element(p, x0, u, vars_n)
    assert err == 0
    h = 1e-10
    Kfd = np.zeros_like(K)
    for j in range(6 * nn):
        e = np.zeros(6 * nn)
        e[j] = h
        Kfd[:, j] = (O.coh_element(p, x0, u + e, vars_n)[0] - O.coh_element(p, x0, u - e, vars_n)[0]) / (2 * h)
    assert np.abs(K - Kfd).max() < 1e-6 * np.abs(K).max()
