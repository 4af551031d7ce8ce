"""The learned warm start on the GPU: trained with the library's MLP kernels on QP states
logged from plain crystal-plasticity passes, it must leave the converged state unchanged
(1e-10 relative against the plain start); an untrained network must give the same state
through the fallback to the plain start.  The iteration counts of both starts are printed:
with slip increments predicted to a few percent (what a BF16 network reaches) the learned
start takes MORE local iterations than the analytic plastic predictor -- the local Newton
needs increments accurate to about 0.3% to gain even half an iteration."""
import numpy as np
import pytest
import torch

import oracle as O
import pgfem3d_amd as P
import surrogate as S
from conftest import rel_inf
from problems import CPBoxProblem

pytestmark = pytest.mark.gpu
TOL = 1e-10
MAT = P.fcc_crystal()
STEP = 1e-3


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _history(prob, device, n_steps, log=None):
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.R, prob.grain)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    row = 0
    for s in range(1, n_steps + 1):
        r, defl = prob.displacement(STEP * s, 0.0, seed=s)
        assert h.residual(f, _dev(r, device), defl=_dev(defl, device)) == 0
        if log is not None:
            row = S.log_step(h, log[0], log[1], row)
        h.commit_step()
    return h, f


def _step_states(h, prob, device, s):
    r, defl = prob.displacement(STEP * s, 0.0, seed=s)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(int(prob.Ap[-1]), dtype=torch.float64, device=device)
    assert h.residual_tangent(f, vals, _dev(r, device), defl=_dev(defl, device)) == 0
    out = {k: h.get_state(k) for k in (P.STATE_PLASTIC_NP1, P.STATE_SCALAR_NP1, P.STATE_TAU, P.STATE_ITERS)}
    out["f"] = f.cpu().numpy()
    out["vals"] = vals.cpu().numpy()
    return out


def _oracle_history(prob, n_steps):
    """The same load history over the CPU oracle: (CPState after n_steps, f, vals at step
    n_steps + 1 with its integration)."""
    st = O.CPState(prob.nqp, prob.mats[0].g0)
    for s in range(1, n_steps + 1):
        r, defl = prob.displacement(STEP * s, 0.0, seed=s)
        O.assemble_cp(prob.mesh, prob.mats, prob.grain, prob.gP, prob.gL, prob.ids, prob.ndof, prob.Ap, prob.Ai, st,
                      r, defl=defl, tangent=False)
        st.commit()
    r, defl = prob.displacement(STEP * (n_steps + 1), 0.0, seed=n_steps + 1)
    fo, vo, err = O.assemble_cp(prob.mesh, prob.mats, prob.grain, prob.gP, prob.gL, prob.ids, prob.ndof, prob.Ap,
                                prob.Ai, st, r, defl=defl)
    assert err == 0
    return st, fo, vo


def test_trained_warm_start_matches_the_oracle(device):
    """Config 5 at small size: the learned start (trained on 8 logged plain steps) and the
    plain start both land on the oracle's step-9 state -- f, K, pF, g and tau within 1e-10 of
    the oracle's largest entry -- and on each other.  The iteration counts are printed, not
    asserted: the learned start does not save iterations here (see the module docstring)."""
    prob = CPBoxProblem(8, P.PGF_TET4, MAT, n_grains=8, seed=4)
    n_log = 8
    X = torch.empty((n_log * prob.nqp, S.N_FEAT), dtype=torch.float32, device=device)
    T = torch.empty((n_log * prob.nqp, S.N_OUT), dtype=torch.float32, device=device)
    h, _ = _history(prob, device, n_log, (X, T))
    assert torch.isfinite(X).all() and torch.isfinite(T).all() and float(T.abs().max()) > 0
    mlp, nrm, hist = S.train(X, T, epochs=25, batch=1024, lr=2e-3, seed=0)
    assert hist[-1][2] < 0.2 * hist[0][2] or hist[-1][2] < 0.05  # validation loss fell
    plain = _step_states(h, prob, device, n_log + 1)
    h.reset_step()
    S.attach(h, mlp, nrm)
    warm = _step_states(h, prob, device, n_log + 1)
    n_learned, n_fallback = h.cp_start_stats()
    st, fo, vo = _oracle_history(prob, n_log)
    ref = {P.STATE_PLASTIC_NP1: st.pF_np1, P.STATE_TAU: st.tau, "f": fo, "vals": vo}
    for run in (plain, warm):
        for k, v in ref.items():
            assert rel_inf(run[k], v) < TOL, k
        assert rel_inf(run[P.STATE_SCALAR_NP1][0], st.g_np1) < TOL
    it_p, it_w = plain[P.STATE_ITERS].mean(), warm[P.STATE_ITERS].mean()
    print(f"[surrogate] mean local iterations: plain {it_p:.3f}, learned start {it_w:.3f}; {n_learned} of "
          f"{prob.nqp} QPs from the learned start, {n_fallback} fell back; validation loss "
          f"{hist[0][2]:.3e} -> {hist[-1][2]:.3e}")
    assert 0 <= n_fallback <= n_learned + n_fallback <= prob.nqp
    h.close()
    mlp.close()


def test_warm_start_full_size_same_state(device):
    """Config 5 at its size: 55^3 cells x 6 linear tets (998,250 QPs), 64 grains, trained on
    5 logged plain steps (5.0M states, seeded 90/10 split); at step 6 the learned-start state
    equals the plain-start state within 1e-10 (pF, g, tau, f); the starts taken are reported."""
    prob = CPBoxProblem(55, P.PGF_TET4, MAT, n_grains=64, seed=2)
    assert prob.nqp == 998250
    n_log = 5
    X = torch.empty((n_log * prob.nqp, S.N_FEAT), dtype=torch.float32, device=device)
    T = torch.empty((n_log * prob.nqp, S.N_OUT), dtype=torch.float32, device=device)
    h, _ = _history(prob, device, n_log, (X, T))
    mlp, nrm, hist = S.train(X, T, epochs=2, batch=8192, lr=2e-3, seed=0)
    del X, T
    r, defl = prob.displacement(STEP * (n_log + 1), 0.0, seed=n_log + 1)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    out = []
    for learned in (False, True):
        if learned:
            S.attach(h, mlp, nrm)
        assert h.residual(f, _dev(r, device), defl=_dev(defl, device)) == 0
        out.append({k: h.get_state(k) for k in (P.STATE_PLASTIC_NP1, P.STATE_SCALAR_NP1, P.STATE_TAU,
                                                P.STATE_ITERS)})
        out[-1]["f"] = f.cpu().numpy()
        h.reset_step()
    n_learned, n_fallback = h.cp_start_stats()
    plain, warm = out
    for k in (P.STATE_PLASTIC_NP1, P.STATE_TAU, "f"):
        assert rel_inf(warm[k], plain[k]) < TOL, k
    assert rel_inf(warm[P.STATE_SCALAR_NP1][0], plain[P.STATE_SCALAR_NP1][0]) < TOL
    print(f"[surrogate 55^3] mean local iterations: plain {plain[P.STATE_ITERS].mean():.3f}, learned "
          f"{warm[P.STATE_ITERS].mean():.3f}; {n_learned} QPs from the learned start, {n_fallback} fell back; "
          f"validation loss {hist[0][2]:.3e} -> {hist[-1][2]:.3e}")
    h.set_surrogate(None)
    h.close()
    mlp.close()


def test_untrained_network_falls_back_to_plain_start(device):
    prob = CPBoxProblem(4, P.PGF_TET4, MAT, n_grains=4, seed=6)
    h, _ = _history(prob, device, 4)
    plain = _step_states(h, prob, device, 5)
    h.reset_step()
    mlp = P.MLP(S.WIDTHS, seed=3)
    # a wild output scale makes the predicted increments far too large
    S.attach(h, mlp, S.Normalisation(np.zeros(34, np.float32), np.ones(34, np.float32), np.full(12, 1e-2, np.float32)))
    warm = _step_states(h, prob, device, 5)
    for k in (P.STATE_PLASTIC_NP1, P.STATE_SCALAR_NP1, P.STATE_TAU, "f", "vals"):
        assert rel_inf(warm[k], plain[k]) < TOL, k
    h.set_surrogate(None)
    h.close()
    mlp.close()
