"""Pins of the oracle's constitutive updates (CPU only): closed forms, invariants and central
finite differences, chosen so that a dropped term, a wrong sign or a transposed index fails."""
import numpy as np
import pytest

import oracle as O
import pgfem3d_amd as P

HE = P.neo_hookean(200e3, 0.3)
J2 = P.j2_linear(70e3, 0.33, 250.0, 1000.0)


def _L_dot(L, dE):
    return np.einsum("ijkl,kl->ij", L, dE)


def _fd_tangent_err(stress, L, F, h=1e-6):
    """max over 9 directions of |dS_fd - L:dE| / max|dS| with dE = sym(F^T dF)."""
    err = scale = 0.0
    for k in range(9):
        dF = np.zeros(9)
        dF[k] = h
        dF = dF.reshape(3, 3)
        dS = (stress(F + dF) - stress(F - dF)) / (2 * h)
        dE = 0.5 * (F.T @ dF + dF.T @ F) / h
        err = max(err, np.abs(dS - _L_dot(L, dE)).max())
        scale = max(scale, np.abs(dS).max())
    return err / scale


def _rand_F(rng, amp):
    F = np.eye(3) + amp * rng.standard_normal((3, 3))
    if np.linalg.det(F) <= 0.2:
        F = np.eye(3)
    return F


def test_he_stress_free_reference():
    S, L, rc = O.he_update(HE, np.eye(3))
    assert rc == 0 and np.abs(S).max() == 0.0


def test_he_uniaxial_closed_form():
    """F = diag(l,1,1): neo-Hookean S = G J^-2/3 (I - trC/3 C^-1) + kappa J U'(J) C^-1,
    written out per component (Doll-Schweizerhof 7 volumetric potential)."""
    E, nu = 200e3, 0.3
    G, kappa = E / (2 * (1 + nu)), E / (3 * (1 - 2 * nu))
    for lam in (0.9, 1.01, 1.3):
        S, _, rc = O.he_update(HE, np.diag([lam, 1.0, 1.0]))
        J = lam
        dU = 0.5 * (np.exp(J - 1) - 1 / J)
        trC = lam**2 + 2
        S11 = G * J ** (-2 / 3) * (1 - trC / 3 / lam**2) + kappa * J * dU / lam**2
        S22 = G * J ** (-2 / 3) * (1 - trC / 3) + kappa * J * dU
        assert rc == 0
        np.testing.assert_allclose([S[0, 0], S[1, 1], S[2, 2]], [S11, S22, S22], rtol=1e-12, atol=1e-9)
        assert np.abs(S - np.diag(np.diag(S))).max() == 0.0


def test_he_tangent_matches_fd():
    rng = np.random.default_rng(0)
    for _ in range(5):
        F = _rand_F(rng, 0.15)
        S, L, rc = O.he_update(HE, F)
        assert rc == 0
        assert _fd_tangent_err(lambda G: O.he_update(HE, G, False)[0], L, F) < 1e-7


def test_he_frame_invariance():
    rng = np.random.default_rng(1)
    F = _rand_F(rng, 0.2)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    S1 = O.he_update(HE, F, False)[0]
    S2 = O.he_update(HE, Q @ F, False)[0]
    np.testing.assert_allclose(S2, S1, rtol=0, atol=1e-9 * np.abs(S1).max())


@pytest.mark.parametrize("vol", [2, 3, 99])
def test_he_volumetric_potentials(vol):
    """The three volumetric potentials (src/new_potentials.cc:575-616): pure dilatation
    F = a I gives S = kappa J U'(J) C^-1 with zero deviatoric part."""
    m = P.neo_hookean(200e3, 0.3, vol_pot=vol)
    kappa = 200e3 / (3 * (1 - 0.6))
    a = 1.02
    J = a**3
    dU = {2: 0.5 * (np.exp(J - 1) - 1 / J), 3: (J - 1) / (2 * J) + 0.5 * np.log(J), 99: J - 1}[vol]
    S = O.he_update(m, a * np.eye(3), False)[0]
    np.testing.assert_allclose(np.diag(S), kappa * J * dU / a**2 * np.ones(3), rtol=1e-12)


def test_j2_elastic_range_equals_neo_hookean():
    rng = np.random.default_rng(2)
    he = P.neo_hookean(70e3, 0.33)
    F = np.eye(3) + 4e-4 * rng.standard_normal((3, 3))
    r = O.j2_update(J2, F, [1, 1, 1, 0, 0, 0], 0.0)
    S, L, _ = O.he_update(he, F)
    assert r["gam"] == 0.0 and r["ep"] == 0.0
    np.testing.assert_allclose(r["S"], S, rtol=0, atol=1e-12 * np.abs(S).max())
    dE = rng.standard_normal((3, 3))
    dE = dE + dE.T  # the two tangents agree on symmetric strain increments
    np.testing.assert_allclose(_L_dot(r["L"], dE), _L_dot(L, dE), rtol=1e-10, atol=1e-6)


def test_j2_consistency_after_return():
    """After a plastic step ||dev tau|| = sqrt(2/3)(k0 + hp ep) and ep = sqrt(2/3) gam."""
    rng = np.random.default_rng(3)
    for amp in (0.01, 0.05):
        F = np.eye(3) + np.diag([amp, -amp / 3, -amp / 3]) + 0.2 * amp * rng.standard_normal((3, 3))
        r = O.j2_update(J2, F, [1, 1, 1, 0, 0, 0], 0.0)
        assert r["gam"] > 0
        tau = F @ r["S"] @ F.T
        s = tau - np.trace(tau) / 3 * np.eye(3)
        np.testing.assert_allclose(np.sqrt((s * s).sum()), np.sqrt(2 / 3) * (250.0 + 1000.0 * r["ep"]), rtol=1e-12)
        np.testing.assert_allclose(r["ep"], np.sqrt(2 / 3) * r["gam"], rtol=1e-14)


def test_j2_small_strain_uniaxial_closed_form():
    """Small-strain limit of uniaxial stress (lateral stretch chosen so sigma_22 = 0): the
    stress follows sigma_y + E_t (eps - eps_y), E_t = E H/(E+H) up to O(eps) corrections."""
    E, nu, sy, H = 70e3, 0.33, 250.0, 1000.0
    eps = 0.006
    lo, hi = -0.01, 0.0
    for _ in range(80):  # bisection on the lateral strain for sigma_22 = 0
        mid = 0.5 * (lo + hi)
        F = np.diag([1 + eps, 1 + mid, 1 + mid])
        r = O.j2_update(J2, F, [1, 1, 1, 0, 0, 0], 0.0)
        sig = F @ r["S"] @ F.T / np.linalg.det(F)
        if sig[1, 1] > 0:
            hi = mid
        else:
            lo = mid
    Et = E * H / (E + H)
    expect = sy + Et * (eps - sy / E)
    assert abs(sig[0, 0] - expect) / expect < 2e-2


def test_j2_tangent_matches_fd_plastic():
    rng = np.random.default_rng(4)
    for k in range(4):
        F = np.eye(3) + np.diag([0.02, -0.007, -0.007]) + 0.004 * rng.standard_normal((3, 3))
        cn = np.array([1, 1, 1, 0, 0, 0.0]) + np.r_[0.003 * rng.standard_normal(3), 0.001 * rng.standard_normal(3)]
        ep = 0.01 * k
        r = O.j2_update(J2, F, cn, ep)
        assert r["gam"] > 0
        stress = lambda G: O.j2_update(J2, G, cn, ep, False)["S"]  # noqa: E731
        assert _fd_tangent_err(stress, r["L"], F) < 1e-6


def test_j2_tangent_symmetrised_version_fails_fd():
    """Guard on the check itself: dropping the unsymmetric n (x) dev(n^2) half (as a
    symmetrised tangent would) is detectable by the FD test above."""
    F = np.eye(3) + np.diag([0.03, -0.01, -0.012]) + np.array([[0, 0.01, 0], [0, 0, 0.004], [0.002, 0, 0]])
    cn = np.array([1, 1, 1, 0, 0, 0.0])
    r = O.j2_update(J2, F, cn, 0.0)
    L = r["L"]
    Lsym = 0.5 * (L + L.transpose(2, 3, 0, 1))
    stress = lambda G: O.j2_update(J2, G, cn, 0.0, False)["S"]  # noqa: E731
    assert _fd_tangent_err(stress, L, F) < 1e-6
    assert _fd_tangent_err(stress, Lsym, F) > 1e-5


# ------------------------------------------------------------------ J2 + continuum damage
J2D = P.j2_damage(70e3, 0.33, 250.0, 1000.0, mu=2.0, p1=2.0, p2=1.5, Yin=5.0)


def _stretch(a):
    lat = -0.35 * a
    return np.diag([1.0 + a, 1.0 + lat, 1.0 + lat])


def test_j2_damage_off_is_plain_j2():
    off = P.j2_damage(70e3, 0.33, 250.0, 1000.0, mu=0.0, p1=2.0, p2=1.5, Yin=5.0)
    F = _stretch(0.03) + 0.003 * np.random.default_rng(1).standard_normal((3, 3))
    cn = np.array([1, 1, 1, 0, 0, 0.0])
    a = O.j2_update(J2, F, cn, 0.0)
    b = O.j2d_update(off, F, cn, 0.0, 0.0, 0.0, 1.0)
    assert b["w"] == 0.0 and np.array_equal(a["S"], b["S"]) and np.array_equal(a["L"], b["L"])


def test_weibull_closed_forms():
    """G(Yin) = 0, G(Yin (1 + P1)) = 0.9999 (1 - 1/e), and H = dG/dY (central differences)."""
    g, h = O.weibull(J2D, J2D.Yin)
    assert g == 0.0 and h == 0.0
    g, _ = O.weibull(J2D, J2D.Yin * (1.0 + J2D.p1))
    assert abs(g - 0.9999 * (1.0 - np.exp(-1.0))) < 1e-15
    for Y in (6.0, 9.0, 20.0, 40.0):
        d = 1e-5 * Y
        fd = (O.weibull(J2D, Y + d)[0] - O.weibull(J2D, Y - d)[0]) / (2 * d)
        assert abs(O.weibull(J2D, Y)[1] - fd) < 1e-8 * max(abs(fd), 1e-3)


def test_damage_driving_energy_closed_form():
    """Y0 = G/2 (J^-2/3 (l1^2 + l2^2 + l3^2) - 3) + kappa U(J) for principal stretches, with the
    Doll-Schweizerhof 7 potential U = (e^(J-1) - ln J - 1)/2."""
    lam = np.array([1.04, 0.985, 0.99])
    J = lam.prod()
    G, kappa = J2D.G, J2D.E / (3 * (1 - 2 * J2D.nu))
    U = 0.5 * (np.exp(J - 1) - np.log(J) - 1)
    ref = 0.5 * G * (J ** (-2 / 3) * (lam ** 2).sum() - 3) + kappa * U
    assert abs(O.j2d_Y0(J2D, np.diag(lam)) - ref) < 1e-12 * ref
    # the damage threshold of this material is crossed by a 3% stretch, not by 0.5%
    assert O.j2d_Y0(J2D, _stretch(0.005)) < J2D.Yin < O.j2d_Y0(J2D, _stretch(0.03))


def test_viscous_damage_under_constant_load_closed_form():
    """Held at one F from rest, backward-Euler viscous damage relaxes towards G(Y):
    w_k = X_k = G (1 - (1 + dt mu)^-k)."""
    F = _stretch(0.03)
    Y = O.j2d_Y0(J2D, F)
    G, _ = O.weibull(J2D, Y)
    dt = 0.25
    w = X = 0.0
    cn = np.array([1, 1, 1, 0, 0, 0.0])
    ep = 0.0
    for k in range(1, 9):
        r = O.j2d_update(J2D, F, cn, ep, w, X, dt, want_L=False)
        cn, ep, w, X = r["Cpinv"], r["ep"], r["w"], r["X"]
        exact = G * (1.0 - (1.0 + dt * J2D.mu) ** (-k))
        assert r["damaged"] == 1 and abs(w - exact) < 1e-14 and abs(X - exact) < 1e-14, k
    # S = (1 - w) Sbar with Sbar the undamaged J2 stress at the same plastic state
    r = O.j2d_update(J2D, F, cn, ep, w, X, dt)
    assert r["damaged"] == 1  # X_k < G: still relaxing
    bare = O.j2_update(J2, F, cn, ep)
    np.testing.assert_allclose(r["S"], (1.0 - r["w"]) * bare["S"], rtol=1e-14, atol=1e-12)


def test_damage_capped_at_w_max():
    capped = P.j2_damage(70e3, 0.33, 250.0, 1000.0, mu=50.0, p1=0.5, p2=1.5, Yin=5.0, w_max=0.2)
    r = O.j2d_update(capped, _stretch(0.04), np.array([1, 1, 1, 0, 0, 0.0]), 0.0, 0.0, 0.0, 1.0, want_L=False)
    assert r["w"] == 0.2 and r["X"] > 0.2


def test_j2_damage_tangent_matches_fd_elastic_range():
    """Damaging but elastic (no yield): Sbar = S0, so the reference's symmetrised damage term
    is the exact derivative and L = FD of S (w depends on F through Y0)."""
    el = P.j2_damage(70e3, 0.33, 1e9, 1000.0, mu=2.0, p1=2.0, p2=1.5, Yin=5.0)
    rng = np.random.default_rng(8)
    cn = np.array([1, 1, 1, 0, 0, 0.0])
    for k in range(3):
        F = _stretch(0.025 + 0.01 * k) + 0.003 * rng.standard_normal((3, 3))
        w_n, X_n = 0.01 * k, 0.02 * k
        r = O.j2d_update(el, F, cn, 0.0, w_n, X_n, 0.5)
        assert r["damaged"] == 1 and r["gam"] == 0.0 and r["w"] > w_n
        stress = lambda G: O.j2d_update(el, G, cn, 0.0, w_n, X_n, 0.5, False)["S"]  # noqa: E731
        assert _fd_tangent_err(stress, r["L"], F) < 1e-6


def test_j2_damage_tangent_plastic_differs_only_by_the_symmetrisation():
    """Plastic and damaging: the exact derivative is (1-w) Lbar - evo Sbar (x) S0; the
    reference's tangent takes the symmetric half, so L - evo/2 (Sbar (x) S0 - S0 (x) Sbar)
    matches FD and L itself does not."""
    F = _stretch(0.03) + np.array([[0, 0.004, 0], [0, 0, 0.002], [0.001, 0, 0]])
    cn = np.array([1, 1, 1, 0, 0, 0.0])
    dt = 0.5
    r = O.j2d_update(J2D, F, cn, 0.0, 0.0, 0.0, dt)
    assert r["damaged"] == 1 and r["gam"] > 0
    dmu = dt * J2D.mu
    evo = dmu * r["H"] / (1 + dmu)
    Sbar = r["S"] / (1.0 - r["w"])
    S0 = r["S0"]
    skew = 0.5 * evo * (np.einsum("ij,kl->ijkl", Sbar, S0) - np.einsum("ij,kl->ijkl", S0, Sbar))
    stress = lambda G: O.j2d_update(J2D, G, cn, 0.0, 0.0, 0.0, dt, False)["S"]  # noqa: E731
    assert _fd_tangent_err(stress, r["L"] - skew, F) < 1e-6
    assert _fd_tangent_err(stress, r["L"], F) > 1e-5


# ------------------------------------------------------------------ linear deviatoric law
HE_LIN = P.Material(model=P.MODEL_HYPER_ELASTICITY, devPotFlag=2, volPotFlag=99, E=3.0, nu=0.25, G=1.2)


def test_he_linear_law_patch_closed_form():
    """devPotFlag 2 (devStress_Linear, S_dev = G (C - I)) with the quadratic volumetric potential
    (volPotFlag 99): F = diag(2, 1, 1) gives S_xx = G (l^2 - 1) + kappa J (J - 1) / l^2 = 3.6 + 1.0
    = 4.6, the reference's patch-test value (verification_tests/output/stress_tension_disp.out)
    for its material line (Ex 3, Gyz 1.2, nyz 0.25 -> kappa = 2)."""
    S, _, rc = O.he_update(HE_LIN, np.diag([2.0, 1.0, 1.0]))
    # lateral: S_yy = G (1 - 1) + kappa J (J - 1) = 4
    assert rc == 0 and abs(S[0, 0] - 4.6) < 1e-14 and abs(S[1, 1] - 4.0) < 1e-14 and abs(S[0, 1]) == 0


def test_he_linear_law_tangent_matches_fd():
    rng = np.random.default_rng(12)
    for _ in range(3):
        F = _rand_F(rng, 0.05)
        _, L, rc = O.he_update(HE_LIN, F)
        assert rc == 0
        stress = lambda G: O.he_update(HE_LIN, G, False)[0]  # noqa: E731
        assert _fd_tangent_err(stress, L, F) < 1e-7
