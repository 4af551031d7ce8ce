"""Load steps solved end to end: the Newton driver over the device element path and the device
linear solve against the same Newton over the CPU oracle with a sparse direct solve.  Each step
starts from the converged field of the step before plus the homogeneous increment of the
prescribed stretch, on both sides.  The nodal displacements must agree within 1e-9 relative
after every load step (both Newton loops run to a 1e-12 normalised residual), and the two
line searches must take the same step lengths (1e-6; the step length is a smooth function of
residual norms that agree to ~1e-12) wherever the residual is above rounding level, and
iteration counts within one.  The driver's
subdivision of a failing step is exercised separately.

With the reference's predictor (load_step(r, defl_n, defl_d=increment): K du = -f_defl at
(u_n, defl_n) first), NORM is the residual left by the first step's predictor, which for an
elastic first step is near round-off; those runs stop on the absolute criterion
||R|| <= computer_zero as well, on both sides."""
import numpy as np
import pytest
import torch

import newton as NW
import oracle as O
import pgfem3d_amd as P
from problems import BoxProblem, CPBoxProblem, oracle_newton_step

pytestmark = pytest.mark.gpu
TOL_U = 1e-9


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _run(prob, device, stretches, solver="device", cp=False, predictor=False, computer_zero=1e-15):
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    tables = None
    if cp:
        h.set_orientations(prob.R, prob.grain)
        st = O.CPState(prob.nqp, prob.mats[0].g0)
        tables = (prob.gP, prob.gL)
    else:
        st = O.OracleState(prob.nqp, prob.mats[0].model == P.MODEL_J2_PLASTICITY)
    # the step sequence is fixed on both sides: no physics-based rejection (at 1e-3 per step
    # with dt = 1 these boxes reach dt sum |gamma_dot| / 0.005 = 1.4, which would subdivide).
    # Crystal plasticity: 1e-11, above the rounding of its assembled residual (about 1e-12 of a
    # step's first residual, the stiff flow rule amplifying the rounding of the local solves),
    # where a tighter tolerance is met or missed by chance in the line search on either side
    tol = 1e-11 if cp else 1e-12
    drv = NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=tol, solver=solver, computer_zero=computer_zero,
                          alpha_restart=None)
    r_gpu = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    r_cpu = np.zeros(prob.ndof)
    g_prev = np.zeros(prob.ndof)
    ost = dict(NORM=0.0, first=True)
    iters = []
    defl_n = None
    for s in stretches:
        guess, defl = prob.displacement(s, 0.0, seed=0)
        if predictor:  # from u(n) at defl(n), the predictor takes the increment
            dn = np.zeros_like(defl) if defl_n is None else defl_n
            res = drv.load_step(r_gpu, _dev(dn, device), defl_d=_dev(defl - dn, device))
            r_cpu, it_cpu, gam_cpu = oracle_newton_step(prob, st, r_cpu, dn, ost, tol=tol, cp_tables=tables,
                                                         defl_d=defl - dn, computer_zero=computer_zero)
            defl_n = defl
        else:
            r_gpu += _dev(guess - g_prev, device)  # converged u(n) + homogeneous increment
            res = drv.load_step(r_gpu, _dev(defl, device))
            r_cpu, it_cpu, gam_cpu = oracle_newton_step(prob, st, r_cpu + guess - g_prev, defl, ost, tol=tol,
                                                         cp_tables=tables, computer_zero=computer_zero)
            g_prev = guess
        assert res.converged, (s, res.norms, res.info, res.solver_iters, res.gammas)
        u = r_gpu.cpu().numpy()
        rel = np.abs(u - r_cpu).max() / np.abs(r_cpu).max()
        assert rel < TOL_U, (s, rel)
        # an iterative solve at 1e-13 may move the last norm across the tolerance
        assert abs(res.iterations - it_cpu) <= 1, (s, res.norms, res.gammas, gam_cpu)
        # step lengths taken from a residual at rounding level (< 1e-9 of the first) are
        # decided by the rounding of the sums, which differs between the two assemblies
        k = min(len(res.gammas), len(gam_cpu))
        live = np.asarray(res.norms[:k]) > 1e-9
        np.testing.assert_allclose(np.asarray(res.gammas[:k])[live], np.asarray(gam_cpu[:k])[live], rtol=1e-6)
        iters.append((res.iterations, it_cpu, min(gam_cpu, default=1.0)))
    h.close()
    return iters


def test_neo_hookean_cfg1_five_steps(device):
    """Config 1: 10^3 cells x 6 P1 tets, neo-Hookean E=200 GPa nu=0.3, 1% stretch per step."""
    prob = BoxProblem(10, P.PGF_TET4, P.neo_hookean(200e3, 0.3))
    iters = _run(prob, device, [0.01 * k for k in range(1, 6)])
    assert all(a <= 8 for a, _, _ in iters), iters  # quadratic convergence


def test_j2_tet10_into_plasticity(device):
    prob = BoxProblem(3, P.PGF_TET10, P.j2_linear(70e3, 0.33, 250.0, 1000.0))
    _run(prob, device, [0.003, 0.006, 0.01])


def test_host_direct_solve_option(device):
    prob = BoxProblem(4, P.PGF_HEX8, P.neo_hookean(200e3, 0.3))
    _run(prob, device, [0.01, 0.02], solver="host")


def test_crystal_plasticity_steps(device):
    """Four grains of FCC crystal plasticity; without the line search the full Newton steps
    cycle at ~0.1 of the first residual.  The device BiCGStab solves the tangent."""
    prob = CPBoxProblem(3, P.PGF_TET4, P.fcc_crystal(), n_grains=4, seed=2)
    iters = _run(prob, device, [1e-3, 2e-3, 3e-3, 4e-3], cp=True)
    assert all(g < 1.0 for _, _, g in iters), iters  # every step needed a cut


def test_predictor_j2_tet10_into_plasticity(device):
    """The reference's predictor on the J2 steps; the first (elastic) step's predictor leaves
    a residual near round-off, so the absolute criterion 1e-7 (forces of order 1e3) ends it."""
    prob = BoxProblem(3, P.PGF_TET10, P.j2_linear(70e3, 0.33, 250.0, 1000.0))
    iters = _run(prob, device, [0.003, 0.006, 0.01], predictor=True, computer_zero=1e-7)
    assert all(a <= 8 for a, _, _ in iters), iters


def test_predictor_crystal_plasticity(device):
    prob = CPBoxProblem(3, P.PGF_TET4, P.fcc_crystal(), n_grains=4, seed=2)
    _run(prob, device, [1e-3, 2e-3, 3e-3, 4e-3], cp=True, predictor=True, computer_zero=1e-9)


def test_subdivision_recovers_a_large_step(device):
    """A 4% crystal-plasticity stretch from rest applied at once fails in the local Newton of
    its first pass (7 QPs on the oracle); the driver cuts it into halves until every piece
    converges (the oracle with the same halving needs 9 cuts, none deeper than 3 levels)."""
    prob = CPBoxProblem(3, P.PGF_TET4, P.fcc_crystal(), n_grains=4, seed=2)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.R, prob.grain)
    drv = NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=1e-10)
    r = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    _, d0 = prob.displacement(0.0, 0.0, seed=0)
    _, d1 = prob.displacement(4e-2, 0.0, seed=0)
    res = drv.increment(r, _dev(d0, device), _dev(d1, device), max_cuts=6, predictor=False)
    assert res.converged, res
    assert res.cuts >= 1
    h.close()


def test_failed_linear_solve_fails_the_step(device):
    """A linear solve that stops short (here: a 2-iteration BiCGStab limit) leaves a relative
    residual far above 500 x its tolerance; the step must fail with INFO = 1 for subdivision
    (Newton_Raphson.cc:996-1011) instead of line-searching a partial update."""
    prob = BoxProblem(4, P.PGF_TET4, P.neo_hookean(200e3, 0.3))
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    drv = NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=1e-10, lin_maxit=2)
    r = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    _, defl = prob.displacement(0.01, 0.0, seed=0)
    res = drv.load_step(r, _dev(defl, device))
    assert not res.converged and res.info == 1, res
    assert res.solver_iters and res.solver_iters[-1][1] > 500.0 * drv.lin_tol
    # with the default limit the same step converges
    h.reset_step()
    r.zero_()
    drv2 = NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=1e-10)
    assert drv2.load_step(r, _dev(defl, device)).converged
    h.close()
