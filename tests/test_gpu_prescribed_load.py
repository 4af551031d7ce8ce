"""GPU parity of pgf_prescribed_load, the load vector of a prescribed increment (f_defl of the
reference's momentum_equation_load4pBCs, src/load.cc:102-220), against the oracle's restatement
(tests/problems.py oracle_prescribed_load: the oracle tangent over free and prescribed DOFs,
its free-row / prescribed-column block times the increment).  Each case first runs an
integrating pass at (r, defl) -- the state a converged step leaves -- then compares f_defl at
(r, defl) for the increment defl_d, in both assembly modes.  Tolerance 1e-10 relative (FP64)."""
import numpy as np
import pytest
import torch

import oracle as O
import pgfem3d_amd as P
from conftest import rel_inf
from problems import BoxProblem, CPBoxProblem, InterfaceBoxProblem, fcc_crystal_he, oracle_prescribed_load

pytestmark = pytest.mark.gpu
TOL = 1e-10
COH = P.Cohesive(sig_c=500.0, chi_c=1e-5, beta=0.5, alpha=1.0)


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _check(prob, device, r, defl, dd, st, cp_tables=None, cst=None, dt=1.0):
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    if cp_tables is not None:
        h.set_orientations(prob.R, prob.grain)
    if cst is not None:
        h.set_cohesive_props(prob.coh)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(int(prob.Ap[-1]), dtype=torch.float64, device=device)
    rd, dfl, ddd = _dev(r, device), _dev(defl, device), _dev(dd, device)
    assert h.residual_tangent(f, vals, rd, defl=dfl, dt=dt) == 0
    fo = oracle_prescribed_load(prob, st, r, defl, dd, cp_tables=cp_tables, cst=cst, dt=dt)
    assert np.abs(fo).max() > 0
    for mode in (P.ASSEMBLY_GATHER, P.ASSEMBLY_COLOURED):
        h.set_assembly(mode)
        out = torch.full((prob.ndof,), np.nan, dtype=torch.float64, device=device)
        assert h.prescribed_load(out, rd, ddd, defl=dfl, dt=dt) == 0
        assert rel_inf(out.cpu().numpy(), fo) < TOL, (mode, rel_inf(out.cpu().numpy(), fo))
    h.close()


@pytest.mark.parametrize("etype", [P.PGF_TET4, P.PGF_TET10, P.PGF_HEX8])
def test_neo_hookean(device, etype):
    prob = BoxProblem(2, etype, P.neo_hookean(200e3, 0.3), seed=3)
    st = O.OracleState(prob.nqp, False)
    r, defl = prob.displacement(0.02, 1e-3, seed=4)
    assert O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r, defl=defl)[2] == 0
    _check(prob, device, r, defl, np.array([3e-3]), st)


def test_j2_plastic(device):
    prob = BoxProblem(2, P.PGF_TET10, P.j2_linear(70e3, 0.33, 250.0, 1e3), seed=2)
    st = O.OracleState(prob.nqp, True)
    r, defl = prob.displacement(0.01, 1e-4, seed=6)
    assert O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r, defl=defl)[2] == 0
    assert np.abs(st.S_np1[0]).max() > 0  # yielded
    _check(prob, device, r, defl, np.array([1e-3]), st)


@pytest.mark.parametrize("law", ["cubic", "hyperelastic"])
def test_crystal_plasticity(device, law):
    mat = P.fcc_crystal() if law == "cubic" else fcc_crystal_he()
    prob = CPBoxProblem(2, P.PGF_TET10 if law == "cubic" else P.PGF_TET4, mat, n_grains=3, seed=5)
    st = O.CPState(prob.nqp, mat.g0)
    r, defl = prob.displacement(1.5e-3, 2e-5, seed=8)
    assert prob.assemble(st, r, defl)[2] == 0
    assert np.abs(st.gdot).max() > 1e-5
    _check(prob, device, r, defl, np.array([2e-4]), st, cp_tables=(prob.gP, prob.gL))


def test_interfaces(device):
    prob = InterfaceBoxProblem(2, P.PGF_TET4, P.neo_hookean(200e3, 0.3), COH, n_grains=3, seed=2)
    st = O.OracleState(prob.nqp, False)
    cst = O.CohState(prob.ncqp)
    r, defl = prob.displacement(1e-3, 1e-4, seed=5)
    assert O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r, defl=defl)[2] == 0
    _check(prob, device, r, defl, np.array([1e-4]), st, cst=cst)
