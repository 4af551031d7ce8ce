"""The CM3F element of J2 plasticity on the device (pgf_set_three_field with J2 materials)
against the oracle's restatement (orc_assemble_tf / orc_tf_update with the J2 state, pinned in
test_oracle_j2_three_field.py), at the same per-element volume and pressure and the same state
at n:

  * over a load history into the plastic range, the condensed residual and tangent in both
    assembly modes within 1e-10 of the oracle's largest entry, the model's F(n+1) =
    (theta/J)^(1/3) F within 1e-12, Cp^-1, ep and the plastic multiplier within 1e-10;
  * the element update (ddV, ddP) within 1e-10 relative;
  * Newton load steps on linear tets (device driver, device linear solve) land on the
    displacement formulation's J2 solution, with theta = J.
"""
import numpy as np
import pytest
import torch

import newton as NW
import oracle as O
import pgfem3d_amd as P
from conftest import rel_inf
from problems import BoxProblem

pytestmark = pytest.mark.gpu
J2 = P.j2_linear(70e3, 0.33, 250.0, 1000.0)
TOL = 1e-10


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(device)


@pytest.mark.parametrize("assembly", [P.ASSEMBLY_GATHER, P.ASSEMBLY_COLOURED])
@pytest.mark.parametrize("etype,n", [(4, 3), (10, 2), (8, 2)])
def test_condensed_pass_state_and_update(device, etype, n, assembly):
    print(f"j2 cm3f: etype {etype} n {n} assembly {assembly}", flush=True)
    prob = BoxProblem(n, etype, J2, seed=2)
    ne = prob.mesh.n_elem
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_assembly(assembly)
    h.set_three_field(True)
    st = O.OracleState(prob.nqp, True)
    rng = np.random.default_rng(5)
    f = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.zeros(h.nnz, dtype=torch.float64, device=device)
    for step, stretch in enumerate((0.004, 0.01, 0.016)):
        r, defl = prob.displacement(stretch, 2e-4, seed=step)
        theta = 1.0 + 1e-3 * rng.standard_normal(ne)
        Pe = 200.0 * rng.standard_normal(ne)
        for fld, v in ((P.TF_V_NP1, theta), (P.TF_P_NP1, Pe), (P.TF_DV, 0 * theta), (P.TF_DP, 0 * Pe)):
            h.tf_set(fld, v)
        assert h.residual_tangent(f, vals, _dev(r, device), defl=_dev(defl, device)) == 0
        Fo = np.zeros((9, prob.nqp))
        fo, vo, err = O.assemble_tf(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, theta, Pe, r,
                                    defl=defl, F_np1=Fo, state=st)
        assert err == 0
        assert rel_inf(f.cpu().numpy(), fo) < TOL, step
        assert rel_inf(vals.cpu().numpy(), vo) < TOL, step
        assert np.abs(h.get_state(P.STATE_F_NP1) - Fo).max() < 1e-12
        assert rel_inf(h.get_state(P.STATE_PLASTIC_NP1), st.P_np1) < TOL
        sc = h.get_state(P.STATE_SCALAR_NP1)
        assert np.abs(sc[0] - st.S_np1[0]).max() < 1e-12 and np.abs(sc[1] - st.S_np1[1]).max() < 1e-10
        if step:
            assert st.S_np1[0].max() > 0  # yielded
        du = 1e-4 * rng.standard_normal(prob.ndof)
        h.tf_update(_dev(du, device))
        ddV, ddP, err = O.tf_update(prob.mesh, prob.mats, prob.ids, theta, Pe, du, r, defl=defl, state=st)
        assert err == 0
        assert np.abs(h.tf_get(P.TF_DDV) - ddV).max() <= TOL * np.abs(ddV).max()
        assert np.abs(h.tf_get(P.TF_DDP) - ddP).max() <= TOL * np.abs(ddP).max()
        h.commit_step()
        st.commit()
    h.close()


def test_newton_on_linear_tets_lands_on_the_displacement_solution(device):
    prob = BoxProblem(3, P.PGF_TET4, J2, seed=9)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_three_field(True)
    drv = NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=1e-12, max_iter=30, alpha_restart=None)
    h1 = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    d1 = NW.NewtonDriver(h1, prob.Ap, prob.Ai, tol=1e-12, max_iter=30, alpha_restart=None)
    r = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    r1 = torch.zeros_like(r)
    for s in range(4):
        defl = _dev(np.array([0.004 * (s + 1)]), device)
        res = drv.load_step(r, defl)
        assert res.converged, (s, res.norms)
        assert d1.load_step(r1, defl).converged, s
        assert (r1 - r).abs().max().item() <= 1e-9 * r1.abs().max().item(), s
        J = np.linalg.det(h1.get_state(P.STATE_F_N).T.reshape(-1, 3, 3))
        assert np.abs(h.tf_get(P.TF_V_N) - J).max() < 1e-9
    assert h1.get_state(P.STATE_SCALAR_N)[0].max() > 0  # plastic
    h.close()
    h1.close()
