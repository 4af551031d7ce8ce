"""Output variables on the device (pgf_output_variables) against the oracle's restatement of
constitutive_model_update_output_variables (orc_output_qp / output_variables, pinned in
test_oracle_output.py): per QP the PK2 stress and the Green-Lagrange strain, per element the
volume averages of the Cauchy stress and of the Almansi strain, within 1e-10 relative of each
array's largest entry and of each QP's (element's) own largest entry -- for the hyperelastic
model, J2 in the plastic range, J2 with damage and crystal plasticity with both elastic laws,
at n+1 after an integration and at n after the commit."""
import numpy as np
import pytest
import torch

import oracle as O
import pgfem3d_amd as P
from conftest import rel_cols, rel_inf
from problems import BoxProblem, CPBoxProblem, fcc_crystal_he

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _check(h, ref, level):
    got = h.output_variables(level)
    for name, a, b in zip(("S", "E", "sigma", "eps"), got, ref):
        assert np.abs(b).max() > 0, name
        assert rel_inf(a, b) < TOL, (name, level, rel_inf(a, b))
        assert rel_cols(np.asarray(a).T, np.asarray(b).T) < TOL, (name, level)  # each QP / element


@pytest.mark.parametrize("mat,etype,n", [
    (P.neo_hookean(200e3, 0.3), P.PGF_TET4, 3),
    (P.j2_linear(70e3, 0.33, 250.0, 1000.0), P.PGF_TET10, 2),
    (P.j2_damage(70e3, 0.33, 250.0, 1000.0, mu=2.0, p1=2.0, p2=1.5, Yin=2.0), P.PGF_HEX8, 3),
])
def test_output_variables_displacement_models(device, mat, etype, n):
    prob = BoxProblem(n, etype, mat)
    print(f"output: model {mat.model} nne {etype}", flush=True)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    st = O.OracleState(prob.nqp, True)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    plastic = mat.model == P.MODEL_J2_PLASTICITY
    for step, s in enumerate([0.006, 0.02]):
        r, defl = prob.displacement(s, 5e-4, seed=3 + step)
        assert h.residual(f, _dev(r, device), defl=_dev(defl, device)) == 0
        O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r, defl=defl, tangent=False)
        ref = O.output_variables(prob.mesh, prob.mats, st.F_np1, st.P_np1 if plastic else None,
                                 st.S_np1[2] if plastic else None)
        _check(h, ref, 0)
        h.commit_step()
        st.commit()
        _check(h, ref, 1)
    if plastic:
        assert st.S_n[1].max() > 0  # the second step yielded
    h.close()


@pytest.mark.parametrize("law", ["cubic", "hyperelastic"])
def test_output_variables_crystal_plasticity(device, law):
    mat = P.fcc_crystal() if law == "cubic" else fcc_crystal_he()
    prob = CPBoxProblem(3, P.PGF_TET4, mat, n_grains=4, seed=2)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.R, prob.grain)
    st = O.CPState(prob.nqp, mat.g0)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    for step in range(1, 4):
        r, defl = prob.displacement(1e-3 * step, 2e-5, seed=step)
        assert h.residual(f, _dev(r, device), defl=_dev(defl, device)) == 0
        O.assemble_cp(prob.mesh, prob.mats, prob.grain, prob.gP, prob.gL, prob.ids, prob.ndof, prob.Ap, prob.Ai, st,
                      r, defl=defl, tangent=False)
        h.commit_step()
        st.commit()
    ref = O.output_variables(prob.mesh, prob.mats, st.F_n, st.pF_n, None, prob.grain,
                             prob.gL.reshape(-1, 81))
    _check(h, ref, 1)
    h.close()
