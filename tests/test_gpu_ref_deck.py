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


def test_checkpoint_and_resume(device, tmp_path):
    """The deck stopped after step 24, its state written as the reference's restart files
    (restart/Mechanical/STEP_000024/box_<p>_24.res and the time-step info), read into a fresh
    handle and driver and run on to step 49, ends bit for bit where the uninterrupted run ends;
    and the step-49 files written from the device read back within the fixture bars of the
    reference's own step-49 files at n.  (Their values at n-1 belong to t = 9.998e-3: the
    reference split its last step there, as its time_step_info.res records.)"""
    import ref_io as R
    import restart as RS

    prob = cp_deck.load()
    sf = prob.sf
    nsteps, stop = cp_deck.FIXTURE_STEP + 1, 25

    def new():
        h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
        h.set_orientations(prob.rot, prob.grain)
        return h, NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=1e-12, max_iter=40)

    def steps(h, drv, r, defl, s0, s1):
        r_prev = r.clone()
        for step in range(s0, s1):
            defl = defl + sf.increment(step, prob.deck.defl)
            r_prev.copy_(r)
            res = drv.load_step(r, _dev(defl, device), dt=sf.times[step + 1] - sf.times[step])
            assert res.converged, (step, res.norms)
        return defl, r_prev

    h, drv = new()
    r = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    defl, _ = steps(h, drv, r, np.zeros(2), 0, nsteps)
    r_full, F_full, g_full = r.cpu().numpy(), h.get_state(P.STATE_F_N), h.get_state(P.STATE_SCALAR_N)
    h.close()

    h, drv = new()
    r = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    defl, r_prev = steps(h, drv, r, np.zeros(2), 0, stop)
    u_nm1 = cp_deck.nodal_u(prob, r_prev.cpu().numpy(), defl - sf.increment(stop - 1, prob.deck.defl))
    u_n = cp_deck.nodal_u(prob, r.cpu().numpy(), defl)
    t3 = (sf.times[stop - 1], sf.times[stop], sf.times[stop + 1])
    RS.save(str(tmp_path), "box", stop - 1, prob.deck, h, u_nm1, u_n, t3, drv.NORM)
    norm = drv.NORM
    h.close()

    h, drv = new()
    _, u_n2, t3b, norm2 = RS.load(str(tmp_path), "box", stop - 1, prob.deck, h)
    assert t3b == t3 and norm2 == norm and np.array_equal(u_n2, u_n)
    drv.NORM, drv.first = norm2, False
    r = _dev(cp_deck.free_dofs(prob, u_n2), device)
    defl2, r_prev = steps(h, drv, r, defl, stop, nsteps)
    assert np.array_equal(r.cpu().numpy(), r_full)
    assert np.array_equal(h.get_state(P.STATE_F_N), F_full)
    assert np.array_equal(h.get_state(P.STATE_SCALAR_N), g_full)

    u_nm1 = cp_deck.nodal_u(prob, r_prev.cpu().numpy(), defl2 - sf.increment(nsteps - 1, prob.deck.defl))
    u_n = cp_deck.nodal_u(prob, r.cpu().numpy(), defl2)
    RS.save(str(tmp_path), "box", nsteps - 1, prob.deck, h, u_nm1, u_n,
            (sf.times[nsteps - 2], sf.times[nsteps - 1], sf.times[nsteps]), drv.NORM)
    h.close()
    for p, part in enumerate(prob.deck.parts):
        a = R.read_restart_cp(R.restart_file(str(tmp_path), "box", p, nsteps - 1), part.nn, 3, part.ne)
        b = R.read_restart_cp(f"{cp_deck.DIR}/ref_out/box_{p}_49.res", part.nn, 3, part.ne)
        assert np.abs(a[1] - b[1]).max() < 1e-5
        for ea, eb in zip(a[2], b[2]):
            ra, rb = ea[0], eb[0]
            assert np.abs(ra.F_n - rb.F_n).max() < 5e-5 and np.abs(ra.pF_n - rb.pF_n).max() < 5e-5
            assert abs(ra.g_n / rb.g_n - 1) < 2e-4
