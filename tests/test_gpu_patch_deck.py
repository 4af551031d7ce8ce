"""The hyperelastic model's linear deviatoric law (devPotFlag 2) on the device, and the
reference's box patch test run end to end there:
  * residual, CSR values and F against the oracle within 1e-10 (tet4, tet10, hex8);
  * the patch deck (24 linear or quadratic tets, u_x = 1 prescribed on x = 1 in one step)
    solved by the Newton driver over the device path: the volume average of the device's output
    PK2 S_xx is the reference's 4.6 (verification_tests/output/stress_tension_disp.out) within
    1e-12, and the field is homogeneous, F = diag(2, 1, 1)."""
import numpy as np
import pytest
import torch

import newton as NW
import oracle as O
import patch_deck
import pgfem3d_amd as P
from conftest import rel_inf
from problems import BoxProblem

pytestmark = pytest.mark.gpu
LIN = P.Material(model=P.MODEL_HYPER_ELASTICITY, devPotFlag=2, volPotFlag=99, E=200e3, nu=0.3, G=76.9e3)


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


@pytest.mark.parametrize("etype,n", [(P.PGF_TET4, 3), (P.PGF_TET10, 2), (P.PGF_HEX8, 3)])
def test_linear_law_against_oracle(device, etype, n):
    prob = BoxProblem(n, etype, LIN)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    st = O.OracleState(prob.nqp, False)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(int(prob.Ap[-1]), dtype=torch.float64, device=device)
    for step, s in enumerate([0.02, 0.1]):
        r, defl = prob.displacement(s, 1e-3, seed=30 + step)
        assert h.residual_tangent(f, vals, _dev(r, device), defl=_dev(defl, device)) == 0
        fo, vo, err = O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r, defl=defl)
        assert err == 0
        assert rel_inf(f.cpu().numpy(), fo) < 1e-10 and rel_inf(vals.cpu().numpy(), vo) < 1e-10
        assert rel_inf(h.get_state(P.STATE_F_NP1), st.F_np1) < 1e-12
        h.commit_step()
        st.commit()
    h.close()


@pytest.mark.parametrize("lt", ["LT", "QT"])
def test_patch_deck_on_device(device, lt):
    prob = patch_deck.load(lt)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    drv = NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=1e-14, max_iter=40, computer_zero=1e-13, solver="host")
    r = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    res = drv.load_step(r, _dev(prob.deck.defl, device))
    assert res.converged, (res.norms, res.info)
    S, E, sig, eps = h.output_variables(1)  # the committed step
    sxx = patch_deck.average_S_xx(prob, S)
    assert abs(sxx - patch_deck.expected()) < 1e-12, sxx
    F = h.get_state(P.STATE_F_N).reshape(3, 3, -1)
    assert np.abs(F - np.diag([2.0, 1.0, 1.0])[:, :, None]).max() < 1e-12
    # the element Cauchy stress of F = diag(2, 1, 1): sigma_xx = 2 S_xx = 9.2, sigma_yy = S_yy / 2 = 2
    assert np.abs(sig[:, 0] - 9.2).max() < 1e-11 and np.abs(sig[:, 1] - 2.0).max() < 1e-11
    h.close()
