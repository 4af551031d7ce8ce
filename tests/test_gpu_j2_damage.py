"""J2 with continuum damage on the device (the elem_kernel damage instantiation) against the
oracle's restatement of CM_J2P_PARAM's damage step (orc_j2d_update, pinned in
test_oracle_materials.py): residual, CSR values and every state variable -- F, Cp^-1, ep,
gam, w, X, H and the damaged flag -- within 1e-10 of the oracle's largest entry per call, over
load steps that cross the damage threshold, on tet4, tet10 and hex8, in both assembly modes;
the tangent at a displacement other than the integrated one; and the physics-based
subdivision measure max (w(n+1) - w(n)) / 0.05 against the oracle's states."""
import numpy as np
import pytest
import torch

import oracle as O
import pgfem3d_amd as P
from conftest import rel_inf
from problems import BoxProblem

pytestmark = pytest.mark.gpu
TOL = 1e-10
J2D = P.j2_damage(70e3, 0.33, 250.0, 1000.0, mu=2.0, p1=2.0, p2=1.5, Yin=2.0)


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


@pytest.mark.parametrize("assembly", [P.ASSEMBLY_GATHER, P.ASSEMBLY_COLOURED])
@pytest.mark.parametrize("etype,n", [(P.PGF_TET10, 3), (P.PGF_TET4, 4), (P.PGF_HEX8, 3)])
def test_damage_steps_against_oracle(device, etype, n, assembly):
    prob = BoxProblem(n, etype, J2D)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_assembly(assembly)
    st = O.OracleState(prob.nqp, True)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(int(prob.Ap[-1]), dtype=torch.float64, device=device)
    dt = 0.5
    damaged_seen = False
    for step, s in enumerate([0.004, 0.012, 0.02, 0.03]):
        r, defl = prob.displacement(s, 2e-4, seed=20 + step)
        assert h.residual_tangent(f, vals, _dev(r, device), defl=_dev(defl, device), dt=dt) == 0
        fo, vo, err = O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r, defl=defl,
                                 dt=dt)
        assert err == 0
        assert rel_inf(f.cpu().numpy(), fo) < TOL, step
        assert rel_inf(vals.cpu().numpy(), vo) < TOL, step
        assert rel_inf(h.get_state(P.STATE_F_NP1), st.F_np1) < TOL
        assert rel_inf(h.get_state(P.STATE_PLASTIC_NP1), st.P_np1) < TOL
        sc = h.get_state(P.STATE_SCALAR_NP1)
        assert sc.shape[0] == 6
        for c in range(6):
            assert rel_inf(sc[c], st.S_np1[c]) < TOL or np.abs(st.S_np1[c]).max() == 0, (step, c)
        alpha = h.subdivision_param(dt)
        assert abs(alpha - ((st.S_np1[2] - st.S_n[2]) / 0.05).max()) < 1e-10 * max(1.0, alpha)
        damaged_seen |= bool(st.S_np1[5].any())
        h.commit_step()
        st.commit()
    assert damaged_seen and st.S_n[2].max() > 1e-3  # the load crossed the damage threshold
    h.close()


def test_damage_tangent_at_other_displacement(device):
    """The tangent at a displacement other than the integrated one recomputes the damage step
    from the state at n, as the oracle does."""
    prob = BoxProblem(3, P.PGF_TET10, J2D)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    st = O.OracleState(prob.nqp, True)
    r1, d1 = prob.displacement(0.02, 2e-4, seed=7)
    r2, d2 = prob.displacement(0.021, 3e-4, seed=8)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(int(prob.Ap[-1]), dtype=torch.float64, device=device)
    assert h.residual(f, _dev(r1, device), defl=_dev(d1, device), run_integration=True) == 0
    O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r1, defl=d1, tangent=False)
    assert st.S_np1[5].any()
    assert h.tangent(vals, _dev(r2, device), defl=_dev(d2, device)) == 0
    _, vo, _ = O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r2, defl=d2,
                          integrate=False, residual=False)
    assert rel_inf(vals.cpu().numpy(), vo) < TOL
    h.close()
