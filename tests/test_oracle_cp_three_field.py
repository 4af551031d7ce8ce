"""Pins of the oracle's CM3F formulation of crystal plasticity (orc_cp_tf_element,
orc_assemble_cp_tf, orc_cp_tf_update; CM_ThreeField, src/constitutive_model_3f.cc:189-470,
with the model integrated at (theta/J)^(1/3) F, src/constitutive_model.cc:810-818):

  * every block is the derivative of the residual it belongs to, by central differences of
    the re-integrated element (the return map re-solved from the same state at n):
    Kuu = dRu/du, Kup = dRu/dP = (dRp/du)^T, Ktu = dRt/du, Ktt = dRt/dtheta,
    Ktp = dRt/dP = dRp/dtheta, and Kut = dRu/dtheta = 0 (the flow rule does not see the
    volumetric scaling of Fb) -- on linear and quadratic tets and hexes, in the plastic range;
  * with slip switched off (gamma_dot_0 -> 0) the element is the hyperelastic CM3F element of
    the same law (orc_tf_element, pinned in test_oracle_three_field.py);
  * on linear tets (one QP) at theta = J and P = kappa U'(J) the condensed residual and
    tangent are those of the displacement-formulation crystal-plasticity element
    (orc_assemble_cp, pinned against the reference's own restart files and by central
    differences in test_oracle_cp.py), and a converged CM3F load step lands on the
    displacement solution.
"""
import numpy as np
import pytest

import oracle as O
import pgfem3d_amd as P
from problems import BoxProblem, CPBoxProblem, fcc_crystal_he

MAT = fcc_crystal_he()
KAPPA = MAT.E / (3 * (1 - 2 * MAT.nu))


def _grain(seed):
    return P.quat_to_rot(P.random_quaternions(1, seed))[0]


def _element(nne, seed, strain):
    rng = np.random.default_rng(seed)
    prob = BoxProblem(1, nne, P.neo_hookean(200e3, 0.3), jitter=0.1, seed=seed)
    nod = prob.mesh.elem_conn[prob.mesh.elem_ptr[0]:prob.mesh.elem_ptr[1]]
    x = prob.mesh.X[nod].reshape(-1, 3)
    Fm = np.eye(3) + strain * np.array([[1.0, 0.3, 0.0], [0.1, -0.4, 0.2], [0.0, 0.25, -0.35]])
    u = (x @ Fm.T - x) + 0.2 * strain * rng.standard_normal(x.shape)
    return x.reshape(-1), u.reshape(-1)


def _state_n(mat, gP, gL, nne, x, u, nint):
    """A plastic state at n: the element integrated once at u (theta = its mean J), from rest."""
    eye = np.repeat(np.eye(3)[None], nint, axis=0)
    r = O.cp_tf_element(mat, gP, gL, nne, x, u, 1.0, 0.0, 1.0, eye, np.full(nint, mat.g0), np.zeros(nint))
    assert r["rc"] == 0
    return dict(pFn=r["pF1"].reshape(-1, 3, 3), g_n=r["g1"], lam_n=r["lam1"], Fn=r["Fb"].reshape(-1, 3, 3))


@pytest.mark.parametrize("nne", [4, 10, 8])
def test_blocks_are_derivatives_of_the_residuals(nne):
    nint = O.nint(nne)
    R = _grain(3)
    gP, gL = O.cp_grain(MAT, R)
    x, u0 = _element(nne, 5, 1.5e-3)
    st = _state_n(MAT, gP, gL, nne, x, u0, nint)
    _, u = _element(nne, 6, 3e-3)
    theta, Pp = 1.0 + 2e-4, 150.0

    def ev(uu, th, pp):
        r = O.cp_tf_element(MAT, gP, gL, nne, x, uu, th, pp, 1.0, st["pFn"], st["g_n"], st["lam_n"], Fn=st["Fn"])
        assert r["rc"] == 0
        return r

    r0 = ev(u, theta, Pp)
    assert np.abs(r0["gdot"]).max() > 1e-4  # slipping
    nd = 3 * nne
    h = 1e-8
    dRu, dRt, dRp = np.zeros((nd, nd)), np.zeros(nd), np.zeros(nd)
    for l in range(nd):
        up, um = u.copy(), u.copy()
        up[l] += h
        um[l] -= h
        a, b = ev(up, theta, Pp), ev(um, theta, Pp)
        dRu[:, l] = (a["Ru"] - b["Ru"]) / (2 * h)
        dRt[l] = (a["Rt"] - b["Rt"]) / (2 * h)
        dRp[l] = (a["Rp"] - b["Rp"]) / (2 * h)
    rel = lambda A, B: np.abs(A - B).max() / np.abs(B).max()
    assert rel(r0["Kuu"], dRu) < 1e-6
    assert rel(r0["Ktu"], dRt) < 1e-5 or np.abs(r0["Ktu"] - dRt).max() < 1e-6 * abs(r0["Ktt"])
    assert rel(r0["Kup"], dRp) < 1e-6
    ht = 1e-8
    a, b = ev(u, theta + ht, Pp), ev(u, theta - ht, Pp)
    # the resolved shear stresses do not see the volumetric scaling of Fb (tr P_a = 0 and the
    # deviatoric law is isochoric), so M does not depend on theta: Kut = dRu/dtheta = 0
    scale = np.abs(r0["Ru"]).max()
    assert np.abs(r0["Kut"]).max() < 1e-9 * scale
    assert np.abs((a["Ru"] - b["Ru"]) / (2 * ht)).max() < 1e-5 * scale
    assert abs(r0["Ktt"] - (a["Rt"] - b["Rt"]) / (2 * ht)) < 1e-6 * abs(r0["Ktt"])
    assert abs(r0["Ktp"] - (a["Rp"] - b["Rp"]) / (2 * ht)) < 1e-6 * abs(r0["Ktp"])
    hp = 1e-3
    a, b = ev(u, theta, Pp + hp), ev(u, theta, Pp - hp)
    assert rel(r0["Kup"], (a["Ru"] - b["Ru"]) / (2 * hp)) < 1e-8
    assert abs(r0["Ktp"] - (a["Rt"] - b["Rt"]) / (2 * hp)) < 1e-8 * abs(r0["Ktp"])


@pytest.mark.parametrize("nne", [4, 10])
def test_without_slip_is_the_hyperelastic_cm3f_element(nne):
    mat = fcc_crystal_he(gamma_dot_0=1e-60)
    he = P.Material(model=P.MODEL_HYPER_ELASTICITY, devPotFlag=1, volPotFlag=mat.volPotFlag, E=mat.E, nu=mat.nu,
                    m10=mat.m10, m01=mat.m01)
    nint = O.nint(nne)
    gP, gL = O.cp_grain(mat, _grain(4))
    x, u = _element(nne, 8, 4e-3)
    eye = np.repeat(np.eye(3)[None], nint, axis=0)
    theta, Pp = 1.0 + 3e-4, 80.0
    r = O.cp_tf_element(mat, gP, gL, nne, x, u, theta, Pp, 1.0, eye, np.full(nint, mat.g0), np.zeros(nint))
    assert r["rc"] == 0
    ref = O.tf_element(he, nne, x, u, theta, Pp)
    rel = lambda A, B: np.abs(A - B).max() / np.abs(B).max()
    for k in ("Kuu", "Ru", "Kup", "Kc", "fc"):
        assert rel(r[k], ref[k]) < 1e-9, k
    for k in ("Rp", "Rt", "Ktt", "Ktp"):
        assert abs(r[k] - ref[k]) <= 1e-9 * max(abs(ref[k]), 1e-300), k
    assert np.abs(r["Kut"]).max() < 1e-9 * np.abs(r["Kup"]).max() * KAPPA


def test_linear_tets_at_theta_J_are_the_displacement_element():
    prob = CPBoxProblem(2, P.PGF_TET4, MAT, n_grains=3, seed=5)
    st1 = O.CPState(prob.nqp, MAT.g0)
    st3 = O.CPState(prob.nqp, MAT.g0)
    for s in range(1, 4):
        r, defl = prob.displacement(1.2e-3 * s, 2e-5, seed=s)
        f1, v1, e1 = prob.assemble(st1, r, defl)
        assert e1 == 0
        # theta = J and P = kappa U'(J) per element (one QP), from the displacement
        J = np.array([np.linalg.det(st1.F_np1[:, q].reshape(3, 3)) for q in range(prob.nqp)])
        dU = 0.5 * (np.exp(J - 1) - 1 / J)
        f3, v3, e3 = O.assemble_cp_tf(prob.mesh, prob.mats, prob.grain, prob.gP, prob.gL, prob.ids, prob.ndof,
                                      prob.Ap, prob.Ai, st3, J, KAPPA * dU, r, defl=defl)
        assert e3 == 0
        assert np.abs(st3.gdot).max() > 1e-5 or s == 1
        assert np.abs(f3 - f1).max() < 1e-9 * np.abs(f1).max(), s
        assert np.abs(v3 - v1).max() < 1e-9 * np.abs(v1).max(), s
        assert np.abs(st3.pF_np1 - st1.pF_np1).max() < 1e-12
        st1.commit()
        st3.commit()


def test_condensed_newton_lands_on_the_displacement_solution():
    """A CM3F load step on linear tets (condensed solve + element update) converges to the
    displacement formulation's nodal solution, with theta = J and P = kappa U'(J)."""
    prob = CPBoxProblem(2, P.PGF_TET4, MAT, n_grains=3, seed=9)
    # start from a smooth field that already meets the prescribed stretch
    r_prev, defl = prob.displacement(5e-4, 1e-6, seed=1)
    # displacement formulation: Newton on the oracle
    st1 = O.CPState(prob.nqp, MAT.g0)
    r = r_prev.copy()
    for _ in range(30):
        st1 = O.CPState(prob.nqp, MAT.g0)
        f, vals, err = prob.assemble(st1, r, defl)
        assert err == 0
        du = -_solve(prob, vals, f)
        r += du
        if np.abs(du).max() < 1e-13:
            break
    r_disp = r.copy()
    # CM3F: theta, P per element, condensed solve + update
    ne = prob.mesh.n_elem
    theta, Pe = np.ones(ne), np.zeros(ne)
    r = r_prev.copy()
    for it in range(30):
        st3 = O.CPState(prob.nqp, MAT.g0)
        f, vals, err = O.assemble_cp_tf(prob.mesh, prob.mats, prob.grain, prob.gP, prob.gL, prob.ids, prob.ndof,
                                        prob.Ap, prob.Ai, st3, theta, Pe, r, defl=defl)
        assert err == 0
        du = -_solve(prob, vals, f)
        dV, dP, err = O.cp_tf_update(prob.mesh, prob.mats, prob.grain, prob.gP, prob.gL, prob.ids, st3, theta, Pe, du,
                                     r, defl=defl)
        assert err == 0
        r += du
        theta += dV
        Pe += dP
        if np.abs(du).max() < 1e-12 and np.abs(dV).max() < 1e-14:
            break
    assert it < 12
    assert np.abs(r - r_disp).max() < 1e-9 * np.abs(r_disp).max()
    # theta equals J of the displacement solution's F (the model's F(n+1) is (theta/J)^(1/3) F,
    # whose determinant is theta by construction, so J comes from the displacement run)
    J = np.array([np.linalg.det(st1.F_np1[:, q].reshape(3, 3)) for q in range(prob.nqp)])
    assert np.abs(theta - J).max() < 1e-10
    assert np.abs(st3.F_np1 - st1.F_np1).max() < 1e-10
    assert np.abs(Pe - KAPPA * 0.5 * (np.exp(J - 1) - 1 / J)).max() < 1e-8 * np.abs(Pe).max()


def _solve(prob, vals, f):
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    return sla.spsolve(sp.csr_matrix((vals, prob.Ai, prob.Ap), shape=(prob.ndof, prob.ndof)).tocsc(), f)
