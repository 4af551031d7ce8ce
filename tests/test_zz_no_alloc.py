"""The allocation-free fused pass (pgf_residual_tangent_static) against pgf_residual_tangent.

Both run the same kernels in the same order; the static entry point only puts the gather
assembly's staging blocks in the library's fixed device array instead of a buffer it
allocates.  So f, the CSR values and the integrated state must be bit-identical.  Two calls
in a row share the array: the second, on a smaller mesh, must not see what the first left.
"""
import numpy as np
import pytest
import torch

import oracle as O
import pgfem3d_amd as P
from conftest import rel_inf
from problems import BoxProblem, MixedProblem

pytestmark = pytest.mark.gpu
COH = P.Cohesive(sig_c=500.0, chi_c=1e-5, beta=0.5, alpha=1.0)


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _handle(prob):
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    if hasattr(prob, "R"):
        h.set_orientations(prob.R, prob.grain)
    if prob.mesh.n_coel:
        h.set_cohesive_props(prob.coh)
    return h


def _fields(h):
    out = [h.get_state(P.STATE_F_NP1)]
    for fld in (P.STATE_PLASTIC_NP1, P.STATE_SCALAR_NP1):
        try:
            out.append(h.get_state(fld))
        except P.PGFError:
            pass
    return out


def _compare(prob, device, stretches, oracle_check=None):
    print(f"no-alloc pass: {prob.mesh.n_elem} elements, {prob.mesh.n_coel} interfaces", flush=True)
    h = _handle(prob)
    assert 0 < h.stage_doubles() <= P.STATIC_STAGE_DOUBLES
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(int(prob.Ap[-1]), dtype=torch.float64, device=device)
    fs = torch.full_like(f, float("nan"))
    vs = torch.full_like(vals, float("nan"))
    for step, s in enumerate(stretches):
        r, defl = prob.displacement(s, s * 0.02, seed=11 + step)
        rd, dd = _dev(r, device), _dev(defl, device)
        assert h.residual_tangent(f, vals, rd, defl=dd) == 0
        ref_state = _fields(h)
        assert h.residual_tangent_static(fs, vs, rd, defl=dd) == 0
        assert torch.equal(fs, f), step
        assert torch.equal(vs, vals), step
        for a, b in zip(_fields(h), ref_state):
            assert np.array_equal(a, b), step
        if oracle_check is not None:
            oracle_check(r, defl, fs.cpu().numpy(), vs.cpu().numpy())
        h.commit_step()
    h.close()


def test_static_pass_matches_allocating_pass_then_smaller_mesh(device):
    # a J2 P2 box, then a smaller neo-Hookean P1 box whose staging is a prefix of the first's
    big = BoxProblem(5, P.PGF_TET10, P.j2_linear(70e3, 0.33, 250.0, 1000.0))
    _compare(big, device, [0.004, 0.012])
    small = BoxProblem(3, P.PGF_TET4, P.neo_hookean(200e3, 0.3))
    st = O.OracleState(small.nqp, False)

    def against_oracle(r, defl, f, vals):
        fo, vo, err = O.assemble(small.mesh, small.mats, small.ids, small.ndof, small.Ap, small.Ai, st, r, defl=defl)
        assert err == 0
        assert rel_inf(f, fo) < 1e-10 and rel_inf(vals, vo) < 1e-10
        st.commit()

    _compare(small, device, [0.01, 0.02], oracle_check=against_oracle)


def test_static_pass_mixed_crystal_plasticity_with_interfaces(device):
    big = MixedProblem(3, 2, 3, P.fcc_crystal(), COH, n_grains=8, seed=3)
    _compare(big, device, [1.5e-4, 3e-4])
    small = MixedProblem(2, 1, 2, P.fcc_crystal(), COH, n_grains=4, seed=4)
    _compare(small, device, [1.5e-4])


def test_static_pass_rejects_oversized_staging(device):
    prob = BoxProblem(12, P.PGF_TET10, P.neo_hookean(200e3, 0.3))
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    assert h.stage_doubles() > P.STATIC_STAGE_DOUBLES
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(int(prob.Ap[-1]), dtype=torch.float64, device=device)
    r, defl = prob.displacement(0.01, 2e-4, seed=1)
    with pytest.raises(P.PGFError, match="PGF_ESCRATCH"):
        h.residual_tangent_static(f, vals, _dev(r, device), defl=_dev(defl, device))
    h.close()
