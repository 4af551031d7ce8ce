"""Pins of the oracle's element kinematics and assembly (CPU only): shape functions, F from an
affine field, K = FD of f, rigid rotations, patch test."""
import numpy as np
import pytest

import oracle as O
import pgfem3d_amd as P
from problems import BoxProblem

HE = P.neo_hookean(200e3, 0.3)
J2 = P.j2_linear(70e3, 0.33, 250.0, 1000.0)


def _free_problem(n, etype, mat, seed=3):
    """Box mesh with every DOF free (no supports)."""
    prob = BoxProblem(n, etype, mat, jitter=0.1, seed=seed)
    sup = np.zeros_like(prob.supports)
    prob.supports = sup
    prob.ids, prob.ndof = P.generate_dof_ids(prob.mesh, sup)
    prob.Ap, prob.Ai = P.psparse_apai(prob.mesh, prob.ids, prob.ndof)
    return prob


def _nodal_to_r(prob, u):
    r = np.zeros(prob.ndof)
    m = prob.ids > 0
    r[prob.ids[m] - 1] = u[m]
    return r


def _dense(prob, vals):
    K = np.zeros((prob.ndof, prob.ndof))
    for i in range(prob.ndof):
        K[i, prob.Ai[prob.Ap[i]:prob.Ap[i + 1]]] = vals[prob.Ap[i]:prob.Ap[i + 1]]
    return K


@pytest.mark.parametrize("nne", [4, 10, 8])
def test_shape_functions(nne):
    m = P.mesh_box(1, nne, 0.0, 0)
    rng = np.random.default_rng(nne)
    x = m.X[m.elem_conn[m.elem_ptr[0]:m.elem_ptr[1]]] + 0.05 * rng.standard_normal((nne, 3))
    A = rng.standard_normal((3, 3))
    vol = 0.0
    for ip in range(O.nint(nne)):
        N, dN, w, rc = O.basis(nne, x, ip)
        assert rc == 0
        assert abs(N.sum() - 1.0) < 1e-14  # partition of unity
        assert np.abs(dN.sum(axis=0)).max() < 1e-12
        # exact gradient of a linear field: sum_a (A x_a) (x) dN_a = A
        np.testing.assert_allclose(((x @ A.T).T @ dN), A, atol=1e-12)
        vol += w
    if nne == 4:
        v = abs(np.linalg.det(np.array([x[1] - x[0], x[2] - x[0], x[3] - x[0]]))) / 6
        assert abs(vol - v) < 1e-15


@pytest.mark.parametrize("etype,mat", [(4, HE), (10, HE), (8, HE), (10, J2), (8, J2)])
def test_tangent_is_fd_of_residual(etype, mat):
    prob = BoxProblem(1, etype, mat, seed=5)
    r, defl = prob.displacement(0.012, 1e-3, seed=9)
    st = O.OracleState(prob.nqp, True)
    f0, vals, err = O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r, defl=defl)
    assert err == 0
    K = _dense(prob, vals)
    h = 1e-7
    Kfd = np.zeros_like(K)
    for j in range(prob.ndof):
        rp, rm = r.copy(), r.copy()
        rp[j] += h
        rm[j] -= h
        fp = O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, O.OracleState(prob.nqp, True),
                        rp, defl=defl, tangent=False)[0]
        fm = O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, O.OracleState(prob.nqp, True),
                        rm, defl=defl, tangent=False)[0]
        Kfd[:, j] = (fp - fm) / (2 * h)
    assert np.abs(K - Kfd).max() / np.abs(K).max() < 1e-6


@pytest.mark.parametrize("etype", [4, 10, 8])
def test_rigid_rotation_gives_zero_force(etype):
    prob = _free_problem(2, etype, HE)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    u = prob.mesh.X @ R.T - prob.mesh.X
    st = O.OracleState(prob.nqp, False)
    f, vals, err = O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, _nodal_to_r(prob, u))
    assert err == 0
    assert np.abs(f).max() < 1e-9 * 200e3 * 0.1  # force scale E * area


@pytest.mark.parametrize("etype,mat", [(4, HE), (10, J2), (8, J2)])
def test_patch_affine_field(etype, mat):
    """Affine displacement on a jittered mesh: F is the same at every QP, interior nodal
    forces vanish and the net force of every component sums to zero."""
    prob = _free_problem(3, etype, mat)
    A = np.array([[0.01, 0.002, 0.0], [0.001, -0.003, 0.002], [0.0, 0.001, 0.012]])
    u = prob.mesh.X @ A.T
    st = O.OracleState(prob.nqp, True)
    f, vals, err = O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, _nodal_to_r(prob, u))
    assert err == 0
    F = st.F_np1.reshape(3, 3, -1)
    np.testing.assert_allclose(F.transpose(2, 0, 1) - (np.eye(3) + A), 0.0, atol=1e-14)
    X = prob.mesh.X
    interior = np.all((X > 1e-9) & (X < 1 - 1e-9), axis=1)
    fi = np.array([f[prob.ids[n] - 1] for n in np.where(interior)[0]])
    fb = np.array([f[prob.ids[n] - 1] for n in np.where(~interior)[0]])
    assert np.abs(fi).max() < 1e-10 * np.abs(fb).max()
    fn = f.reshape(-1, 3) if False else np.array([f[prob.ids[n] - 1] for n in range(prob.mesh.n_nodes)])
    assert np.abs(fn.sum(axis=0)).max() < 1e-10 * np.abs(fb).max()
