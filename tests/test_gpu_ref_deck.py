"""The reference's crystal-plasticity tension deck (8 partitions merged, 257 linear tets, the
hyperelastic elastic law, one orientation per element, 50 load steps) solved end to end on the
device: the Newton driver over the device element path, the local return map and the device
linear solve.

Two bars:
  * against the same Newton over the CPU oracle, both run to a 1e-12 normalised residual:
    nodal displacements within 1e-9 relative, F, pF and g within 1e-9 (relative for g), and
    Newton iterations per step within one;
  * against the reference's own step-49 restart files, the bars of the oracle's test of the
    same deck (test_ref_deck.py): |du| < 1e-5, |dF|, |dpF| < 5e-5, relative dg < 2e-4, set by
    the reference's local tolerances of 1e-5.
The second test runs the deck as the reference's Newton does with the solver file's own
settings (nor_min 1e-5, at most iter_max = 5 iterations): each step opens with the predictor
K du = -f_defl at (u_n, defl_n).  The first step is elastic, so its predictor leaves a residual
near round-off and NORM (taken from it) is tiny; the absolute criterion ||R|| <= 1e-9 ends such
steps on both sides.  Its bars against the oracle run the same way: u within 1e-7 relative,
F, pF within 1e-9, g within 1e-9 relative.
"""
import numpy as np
import pytest
import torch

import cp_deck
import newton as NW
import pgfem3d_amd as P

pytestmark = pytest.mark.gpu
TOL = 1e-9


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def test_reference_tension_deck_on_device(device):
    prob = cp_deck.load()
    nsteps = cp_deck.FIXTURE_STEP + 1
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.rot, prob.grain)
    drv = NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=1e-12, max_iter=40)
    r = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    defl = np.zeros(len(prob.deck.defl))
    its = []
    for step in range(nsteps):
        defl = defl + prob.sf.increment(step, prob.deck.defl)
        res = drv.load_step(r, _dev(defl, device), dt=prob.sf.times[step + 1] - prob.sf.times[step])
        assert res.converged, (step, res.norms, res.info)
        its.append(res.iterations)
    u = cp_deck.nodal_u(prob, r.cpu().numpy(), defl)
    F = h.get_state(P.STATE_F_N)
    pF = h.get_state(P.STATE_PLASTIC_N)
    g = h.get_state(P.STATE_SCALAR_N)[0]
    h.close()

    u_o, st, defl_o, its_o = cp_deck.run_oracle(prob, prob.rot, nsteps, tol=1e-12)
    assert np.array_equal(defl, defl_o)
    assert np.abs(u - u_o).max() / np.abs(u_o).max() < TOL
    assert np.abs(F - st.F_n).max() < TOL and np.abs(pF - st.pF_n).max() < TOL
    assert np.abs(g / st.g_n - 1.0).max() < TOL
    assert max(abs(a - b) for a, b in zip(its, its_o)) <= 1, (its, its_o)

    du, dF, dpF, dg = cp_deck.compare_with_fixture(prob, u, F, pF, g)
    assert du < 1e-5 and dF < 5e-5 and dpF < 5e-5 and dg < 2e-4, (du, dF, dpF, dg)


def test_reference_tension_deck_with_solver_file_settings(device):
    prob = cp_deck.load()
    sf = prob.sf
    nsteps = cp_deck.FIXTURE_STEP + 1
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.rot, prob.grain)
    drv = NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=sf.nor_min, max_iter=40, iter_max=sf.iter_max,
                          computer_zero=1e-9)
    r = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    defl = np.zeros(len(prob.deck.defl))
    its = []
    for step in range(nsteps):
        inc = sf.increment(step, prob.deck.defl)
        res = drv.load_step(r, _dev(defl, device), defl_d=_dev(inc, device), dt=sf.times[step + 1] - sf.times[step])
        assert res.converged, (step, res.norms, res.info)
        defl = defl + inc
        its.append(res.iterations)
    u = cp_deck.nodal_u(prob, r.cpu().numpy(), defl)
    F = h.get_state(P.STATE_F_N)
    pF = h.get_state(P.STATE_PLASTIC_N)
    g = h.get_state(P.STATE_SCALAR_N)[0]
    h.close()

    u_o, st, _, its_o = cp_deck.run_oracle(prob, prob.rot, nsteps, tol=sf.nor_min, predictor=True,
                                           iter_max=sf.iter_max, computer_zero=1e-9)
    assert np.abs(u - u_o).max() / np.abs(u_o).max() < 1e-7
    assert np.abs(F - st.F_n).max() < TOL and np.abs(pF - st.pF_n).max() < TOL
    assert np.abs(g / st.g_n - 1.0).max() < TOL
    assert max(abs(a - b) for a, b in zip(its, its_o)) <= 1, (its, its_o)
    assert max(its) <= sf.iter_max

    du, dF, dpF, dg = cp_deck.compare_with_fixture(prob, u, F, pF, g)
    assert du < 1e-5 and dF < 5e-5 and dpF < 5e-5 and dg < 2e-4, (du, dF, dpF, dg)
