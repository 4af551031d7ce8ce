"""Pins of the oracle's CM3F formulation for J2 plasticity (orc_tf_element / orc_assemble_tf /
orc_tf_update with a J2 material: CM_J2P_PARAM has pF = I, so M = I and dM/du = 0 as for the
hyperelastic model, and update_elasticity_dev gives the deviatoric part of the return map's
stress and tangent, src/cm_j2_plasticity.cc:384-461):

  * every block is the derivative of its residual by central differences of the re-integrated
    element (the return map re-run from the same state at n), in the plastic range, on linear
    and quadratic tets and hexes;
  * below yield the element is the hyperelastic CM3F element of the same elasticity;
  * on linear tets at theta = J and P = kappa U'(J) the condensed residual and tangent are the
    displacement-formulation J2 element's (orc_assemble, pinned in test_oracle_materials.py
    and test_oracle_element.py), and a converged CM3F load step lands on the displacement
    solution with theta = J.
"""
import numpy as np
import pytest

import oracle as O
import pgfem3d_amd as P
from problems import BoxProblem

E, NU = 70e3, 0.33
J2 = P.j2_linear(E, NU, 250.0, 1000.0)
KAPPA = E / (3 * (1 - 2 * NU))


def _dU(J):  # Doll-Schweizerhof 7 (volPotFlag 2)
    return 0.5 * (np.exp(J - 1) - 1 / J)


def _element(nne, seed, strain):
    rng = np.random.default_rng(seed)
    prob = BoxProblem(1, nne, J2, jitter=0.1, seed=seed)
    nod = prob.mesh.elem_conn[prob.mesh.elem_ptr[0]:prob.mesh.elem_ptr[1]]
    x = prob.mesh.X[nod].reshape(-1, 3)
    Fm = np.eye(3) + strain * np.array([[1.0, 0.3, 0.0], [0.1, -0.4, 0.2], [0.0, 0.25, -0.35]])
    u = (x @ Fm.T - x) + 0.2 * strain * rng.standard_normal(x.shape)
    return x.reshape(-1), u.reshape(-1)


def _state_n(nne, x, u):
    """A plastic J2 state at n: one integration of the displacement element's return map per
    QP at u (orc_j2_update), from the virgin state."""
    nint = O.nint(nne)
    Cpn, epn = np.zeros((nint, 6)), np.zeros(nint)
    for ip in range(nint):
        _, dNdX, _, _ = O.basis(nne, x, ip)
        F = np.eye(3) + u.reshape(-1, 3).T @ dNdX.reshape(-1, 3)
        r = O.j2_update(J2, F, np.array([1.0, 1, 1, 0, 0, 0]), 0.0)
        Cpn[ip], epn[ip] = r["Cpinv"], r["ep"]
    assert epn.max() > 0
    return Cpn, epn


@pytest.mark.parametrize("nne", [4, 10, 8])
def test_blocks_are_derivatives_of_the_residuals(nne):
    x, u0 = _element(nne, 5, 8e-3)
    Cpn, epn = _state_n(nne, x, u0)
    _, u = _element(nne, 6, 1.2e-2)
    theta, Pp = 1.0 + 2e-3, 300.0

    def ev(uu, th, pp):
        r = O.tf_element(J2, nne, x, uu, th, pp, Cpn=Cpn, epn=epn)
        assert r["rc"] == 0
        return r

    r0 = ev(u, theta, Pp)
    nd = 3 * nne
    h = 1e-8
    dRu, dRp = np.zeros((nd, nd)), np.zeros(nd)
    for l in range(nd):
        up, um = u.copy(), u.copy()
        up[l] += h
        um[l] -= h
        a, b = ev(up, theta, Pp), ev(um, theta, Pp)
        dRu[:, l] = (a["Ru"] - b["Ru"]) / (2 * h)
        dRp[l] = (a["Rp"] - b["Rp"]) / (2 * h)
    rel = lambda A, B: np.abs(A - B).max() / np.abs(B).max()
    assert rel(r0["Kuu"], dRu) < 1e-6
    assert rel(r0["Kup"], dRp) < 1e-6
    ht = 1e-7
    a, b = ev(u, theta + ht, Pp), ev(u, theta - ht, Pp)
    assert np.abs(a["Ru"] - b["Ru"]).max() == 0.0  # Kut = 0: Ru does not see theta
    assert abs(r0["Ktt"] - (a["Rt"] - b["Rt"]) / (2 * ht)) < 1e-6 * abs(r0["Ktt"])
    assert abs(r0["Ktp"] - (a["Rp"] - b["Rp"]) / (2 * ht)) < 1e-8 * abs(r0["Ktp"])
    hp = 1e-2
    a, b = ev(u, theta, Pp + hp), ev(u, theta, Pp - hp)
    assert rel(r0["Kup"], (a["Ru"] - b["Ru"]) / (2 * hp)) < 1e-8
    # the plastic flow changes the element: the virgin state gives another Kuu
    virgin = O.tf_element(J2, nne, x, u, theta, Pp)
    assert rel(virgin["Kuu"], r0["Kuu"]) > 1e-3


@pytest.mark.parametrize("nne", [4, 10])
def test_below_yield_is_the_hyperelastic_element(nne):
    x, u = _element(nne, 8, 4e-4)
    j2 = P.j2_linear(E, NU, 1e9, 1000.0)
    he = P.neo_hookean(E, NU)
    r = O.tf_element(j2, nne, x, u, 1.0 + 3e-4, 80.0)
    ref = O.tf_element(he, nne, x, u, 1.0 + 3e-4, 80.0)
    rel = lambda A, B: np.abs(A - B).max() / np.abs(B).max()
    for k in ("Kuu", "Ru", "Kup", "Kc", "fc"):
        assert rel(r[k], ref[k]) < 1e-10, k


def test_linear_tets_at_theta_J_are_the_displacement_element():
    prob = BoxProblem(2, P.PGF_TET4, J2, seed=5)
    st1 = O.OracleState(prob.nqp, True)
    st3 = O.OracleState(prob.nqp, True)
    for s, stretch in enumerate((0.004, 0.01, 0.016)):
        r, defl = prob.displacement(stretch, 2e-4, seed=s)
        f1, v1, e1 = O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st1, r, defl=defl)
        assert e1 == 0
        J = np.linalg.det(st1.F_np1.T.reshape(-1, 3, 3))
        F3 = np.zeros((9, prob.nqp))
        f3, v3, e3 = O.assemble_tf(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, J, KAPPA * _dU(J), r,
                                   defl=defl, F_np1=F3, state=st3)
        assert e3 == 0
        assert np.abs(f3 - f1).max() < 1e-9 * np.abs(f1).max(), s
        assert np.abs(v3 - v1).max() < 1e-9 * np.abs(v1).max(), s
        assert np.abs(st3.P_np1 - st1.P_np1).max() < 1e-12 and np.abs(st3.S_np1[:2] - st1.S_np1[:2]).max() < 1e-12
        if s:
            assert st1.S_np1[0].max() > 0  # yielded
        st1.commit()
        st3.commit()


def _solve(prob, vals, f):
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    return sla.spsolve(sp.csr_matrix((vals, prob.Ai, prob.Ap), shape=(prob.ndof, prob.ndof)).tocsc(), f)


def test_condensed_newton_lands_on_the_displacement_solution():
    prob = BoxProblem(2, P.PGF_TET4, J2, seed=9)
    r0, defl = prob.displacement(0.01, 1e-4, seed=1)
    r = r0.copy()
    for _ in range(30):
        st1 = O.OracleState(prob.nqp, True)
        f, vals, err = O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st1, r, defl=defl)
        assert err == 0
        du = -_solve(prob, vals, f)
        r += du
        if np.abs(du).max() < 1e-13:
            break
    r_disp = r.copy()
    ne = prob.mesh.n_elem
    theta, Pe = np.ones(ne), np.zeros(ne)
    r = r0.copy()
    for it in range(30):
        st3 = O.OracleState(prob.nqp, True)
        f, vals, err = O.assemble_tf(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, theta, Pe, r,
                                     defl=defl, state=st3)
        assert err == 0
        du = -_solve(prob, vals, f)
        dV, dP, err = O.tf_update(prob.mesh, prob.mats, prob.ids, theta, Pe, du, r, defl=defl, state=st3)
        assert err == 0
        r += du
        theta += dV
        Pe += dP
        if np.abs(du).max() < 1e-12 and np.abs(dV).max() < 1e-14:
            break
    assert it < 12
    assert np.abs(r - r_disp).max() < 1e-9 * np.abs(r_disp).max()
    J = np.linalg.det(st1.F_np1.T.reshape(-1, 3, 3))
    assert np.abs(theta - J).max() < 1e-10
    assert np.abs(Pe - KAPPA * _dU(J)).max() < 1e-8 * np.abs(Pe).max()
    assert st1.S_np1[0].max() > 0  # plastic
