"""GPU parity of the interface (cohesive) elements: volume elements plus zero-thickness
interfaces on Voronoi grain boundaries, residual, tangent and interface history against the
CPU oracle on the same seeded inputs, over a load history with commits (loading, softening,
unloading, contact and killed QPs all occur).  Tolerance 1e-10 relative (FP64)."""
import numpy as np
import pytest
import torch

import oracle as O
import pgfem3d_amd as P
from conftest import rel_inf
from problems import InterfaceBoxProblem

TOL = 1e-10
COH = P.Cohesive(sig_c=500.0, chi_c=1e-5, beta=0.5, alpha=1.0)
STEPS = [(0.002, 2e-4), (0.004, 3e-4), (0.003, 1e-4), (0.001, 1e-4)]

pytestmark = pytest.mark.gpu


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _run(prob, device, cp=False):
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_cohesive_props(prob.coh)
    assert h.n_cqp == prob.ncqp
    if cp:
        h.set_orientations(prob.R, prob.grain)
        st = O.CPState(prob.nqp, prob.mats[0].g0)
        gtab = [O.cp_grain(prob.mats[0], prob.R[g]) for g in range(prob.R.shape[0])]
        gP = np.ascontiguousarray(np.stack([t[0] for t in gtab]).reshape(-1, 108))
        gL = np.ascontiguousarray(np.stack([t[1] for t in gtab]).reshape(-1, 81))
    else:
        st = O.OracleState(prob.nqp, False)
    cst = O.CohState(prob.ncqp)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(int(prob.Ap[-1]), dtype=torch.float64, device=device)
    killed = False
    for step, (s, noise) in enumerate(STEPS):
        r, defl = prob.displacement(s * (0.1 if cp else 1.0), noise * (0.1 if cp else 1.0), seed=300 + step)
        rd, dd = _dev(r, device), _dev(defl, device)
        err = h.residual_tangent(f, vals, rd, defl=dd)
        if cp:
            fo, vo, erro = O.assemble_cp(prob.mesh, prob.mats, prob.grain, gP, gL, prob.ids, prob.ndof, prob.Ap,
                                         prob.Ai, st, r, defl=defl)
        else:
            fo, vo, erro = O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r, defl=defl)
        erro += prob.assemble_coh(cst, r, defl, fo, vo)
        assert err == erro == 0, (step, err, erro)
        assert rel_inf(f.cpu().numpy(), fo) < TOL, (step, rel_inf(f.cpu().numpy(), fo))
        assert rel_inf(vals.cpu().numpy(), vo) < TOL, (step, rel_inf(vals.cpu().numpy(), vo))
        cv = h.get_state(P.STATE_COHESIVE_NP1)
        assert rel_inf(cv, cst.vars_np1) < TOL, step
        killed |= bool(np.any(cst.vars_np1[0] > 12.5 * COH.chi_c))
        # residual-only and tangent-only passes reproduce the fused pass
        f2 = torch.empty_like(f)
        v2 = torch.empty_like(vals)
        assert h.residual(f2, rd, defl=dd) == 0
        assert h.tangent(v2, rd, defl=dd) == 0
        assert rel_inf(f2.cpu().numpy(), fo) < TOL and rel_inf(v2.cpu().numpy(), vo) < TOL
        h.commit_step()
        if cp:
            st.commit()
        else:
            st.commit()
        cst.commit()
        assert rel_inf(h.get_state(P.STATE_COHESIVE_N), cst.vars_n) < TOL
    h.close()
    return killed


def test_interfaces_with_neo_hookean_tets(device):
    prob = InterfaceBoxProblem(3, P.PGF_TET4, P.neo_hookean(200e3, 0.3), COH, n_grains=4, seed=3)
    assert prob.mesh.n_coel > 0
    assert _run(prob, device)  # some QPs pass the kill criterion


def test_interfaces_with_neo_hookean_hexes(device):
    """4 + 4 node interfaces on hexahedral grain boundaries (2x2 Gauss)."""
    prob = InterfaceBoxProblem(3, P.PGF_HEX8, P.neo_hookean(200e3, 0.3), COH, n_grains=3, seed=5)
    assert prob.mesh.n_coel > 0
    _run(prob, device)


def test_interfaces_with_crystal_plasticity(device):
    """Polycrystal with interfaces on the grain boundaries (the mixed workload's physics)."""
    prob = InterfaceBoxProblem(3, P.PGF_TET4, P.fcc_crystal(), COH, n_grains=4, seed=7)
    _run(prob, device, cp=True)


def test_interface_props_required(device):
    prob = InterfaceBoxProblem(2, P.PGF_TET4, P.neo_hookean(200e3, 0.3), COH, n_grains=2, seed=1)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    r, defl = prob.displacement(0.001, 0.0, seed=1)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    with pytest.raises(P.PGFError):
        h.residual(f, _dev(r, device), defl=_dev(defl, device))
    h.close()
