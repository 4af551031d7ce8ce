"""The CM3F element of crystal plasticity on the device (pgf_set_three_field with a
crystal-plasticity material on its hyperelastic law) against the oracle's restatement
(orc_assemble_cp_tf / orc_cp_tf_update, pinned in test_oracle_cp_three_field.py), at the same
per-element volume and pressure variables and the same state at n:

  * over a plastic load history, the condensed residual and tangent in both assembly modes
    within 1e-10 of the oracle's largest entry, the model's F(n+1) = (theta/J)^(1/3) F within
    1e-12, pF(n+1), g and lambda within 1e-10;
  * the element update (ddV, ddP for a Newton increment) within 1e-10 relative;
  * the prescribed-load pass against the oracle's condensed tangent over the expanded numbering;
  * Newton load steps on linear tets (device driver and linear solve) land on the displacement
    formulation's solution, with theta = J;
  * the cubic elastic law, which has no deviatoric split, is refused.
"""
import copy

import numpy as np
import pytest
import torch

import newton as NW
import oracle as O
import pgfem3d_amd as P
from conftest import rel_inf
from problems import CPBoxProblem, expand_prescribed, fcc_crystal_he

pytestmark = pytest.mark.gpu
MAT = fcc_crystal_he()
TOL = 1e-10


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(device)


def _handle(prob, assembly=P.ASSEMBLY_GATHER):
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.R, prob.grain)
    h.set_assembly(assembly)
    h.set_three_field(True)
    return h


@pytest.mark.parametrize("assembly", [P.ASSEMBLY_GATHER, P.ASSEMBLY_COLOURED])
@pytest.mark.parametrize("etype,n", [(4, 3), (10, 2), (8, 2)])
def test_condensed_pass_state_and_update(device, etype, n, assembly):
    print(f"cp cm3f: etype {etype} n {n} assembly {assembly}", flush=True)
    prob = CPBoxProblem(n, etype, MAT, n_grains=4, seed=2)
    ne = prob.mesh.n_elem
    h = _handle(prob, assembly)
    st = O.CPState(prob.nqp, MAT.g0)
    rng = np.random.default_rng(7)
    f = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.zeros(h.nnz, dtype=torch.float64, device=device)
    for step in range(1, 4):
        r, defl = prob.displacement(4e-4 * step, 2e-5, seed=step)
        theta = 1.0 + 2e-4 * rng.standard_normal(ne)
        Pe = 60.0 * rng.standard_normal(ne)
        for fld, v in ((P.TF_V_NP1, theta), (P.TF_P_NP1, Pe), (P.TF_DV, 0 * theta), (P.TF_DP, 0 * Pe)):
            h.tf_set(fld, v)
        assert h.residual_tangent(f, vals, _dev(r, device), defl=_dev(defl, device)) == 0
        fo, vo, err = O.assemble_cp_tf(prob.mesh, prob.mats, prob.grain, prob.gP, prob.gL, prob.ids, prob.ndof,
                                       prob.Ap, prob.Ai, st, theta, Pe, r, defl=defl)
        assert err == 0
        assert rel_inf(f.cpu().numpy(), fo) < TOL, step
        assert rel_inf(vals.cpu().numpy(), vo) < TOL, step
        assert np.abs(h.get_state(P.STATE_F_NP1) - st.F_np1).max() < 1e-12
        assert rel_inf(h.get_state(P.STATE_PLASTIC_NP1), st.pF_np1) < TOL
        sc = h.get_state(P.STATE_SCALAR_NP1)
        assert rel_inf(sc[0], st.g_np1) < TOL and np.abs(sc[1] - st.lam_np1).max() < 1e-10
        if step > 1:
            assert np.abs(st.gdot).max() > 1e-5  # slipping
        # the element update for a Newton increment
        du = 1e-5 * rng.standard_normal(prob.ndof)
        h.tf_update(_dev(du, device))
        ddV, ddP, err = O.cp_tf_update(prob.mesh, prob.mats, prob.grain, prob.gP, prob.gL, prob.ids, st, theta, Pe,
                                       du, r, defl=defl)
        assert err == 0
        assert np.abs(h.tf_get(P.TF_DDV) - ddV).max() <= TOL * np.abs(ddV).max()
        assert np.abs(h.tf_get(P.TF_DDP) - ddP).max() <= TOL * np.abs(ddP).max()
        h.commit_step()
        st.commit()
    h.close()


def test_prescribed_load(device):
    prob = CPBoxProblem(2, P.PGF_TET10, MAT, n_grains=3, seed=4)
    ne = prob.mesh.n_elem
    rng = np.random.default_rng(3)
    h = _handle(prob)
    st = O.CPState(prob.nqp, MAT.g0)
    r, defl = prob.displacement(6e-4, 2e-5, seed=1)
    theta, Pe = 1.0 + 2e-4 * rng.standard_normal(ne), 40.0 * rng.standard_normal(ne)
    h.tf_set(P.TF_V_NP1, theta)
    h.tf_set(P.TF_P_NP1, Pe)
    f = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    assert h.residual(f, _dev(r, device), defl=_dev(defl, device)) == 0
    fo, _, err = O.assemble_cp_tf(prob.mesh, prob.mats, prob.grain, prob.gP, prob.gL, prob.ids, prob.ndof, prob.Ap,
                                  prob.Ai, st, theta, Pe, r, defl=defl, tangent=False)
    assert err == 0 and rel_inf(f.cpu().numpy(), fo) < TOL
    defl_d = np.array([3e-4])
    fd = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    assert h.prescribed_load(fd, _dev(r, device), _dev(defl_d, device), defl=_dev(defl, device)) == 0
    h.close()
    import scipy.sparse as sp

    ids2, ndof2, Ap2, Ai2 = expand_prescribed(prob, len(defl))
    _, v2, err = O.assemble_cp_tf(prob.mesh, prob.mats, prob.grain, prob.gP, prob.gL, ids2, ndof2, Ap2, Ai2,
                                  copy.deepcopy(st), theta, Pe, np.concatenate([r, defl]), integrate=False)
    assert err == 0
    K2 = sp.csr_matrix((v2, Ai2, Ap2), shape=(ndof2, ndof2))
    ref = K2[:prob.ndof, prob.ndof:] @ defl_d
    assert np.abs(fd.cpu().numpy() - ref).max() <= TOL * np.abs(ref).max()


def test_newton_on_linear_tets_lands_on_the_displacement_solution(device):
    prob = CPBoxProblem(3, P.PGF_TET4, MAT, n_grains=4, seed=9)
    h = _handle(prob)
    # 1e-11: the assembled residual of these steps rounds at about 1e-12 of the step's first
    # residual (forces of the accumulated strain against the increment's), where a tighter
    # tolerance is met or missed by chance in the line search
    tol = 1e-11
    drv = NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=tol, max_iter=30, alpha_restart=None)
    h1 = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h1.set_orientations(prob.R, prob.grain)
    d1 = NW.NewtonDriver(h1, prob.Ap, prob.Ai, tol=tol, max_iter=30, alpha_restart=None)
    r = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    r1 = torch.zeros_like(r)
    for s in range(4):
        defl = _dev(np.array([1.5e-4 * (s + 1)]), device)
        res = drv.load_step(r, defl)
        assert res.converged, (s, res.norms)
        res1 = d1.load_step(r1, defl)
        assert res1.converged, (s, res1.norms)
        assert (r1 - r).abs().max().item() <= 1e-9 * r1.abs().max().item(), s
        # theta = J of each element's F (one QP per element: QP order = element order)
        J = np.linalg.det(h1.get_state(P.STATE_F_N).T.reshape(-1, 3, 3))
        assert np.abs(h.tf_get(P.TF_V_N) - J).max() < 1e-9
    assert np.abs(h1.get_state(P.STATE_SCALAR_N)[0] - MAT.g0).max() > 0  # hardened
    h.close()
    h1.close()


def test_cubic_law_is_refused(device):
    prob = CPBoxProblem(2, P.PGF_TET4, P.fcc_crystal(), n_grains=2, seed=1)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.R, prob.grain)
    with pytest.raises(P.PGFError):
        h.set_three_field(True)
    h.close()
