"""The mixed batch: linear tets, quadratic tets and hexes in crystal-plasticity grains with
interface elements on grain boundaries, one pass over all of them (the local Newton and the
tangent data of every element type in single launches).  Checked against the oracle (1e-10),
and the hipGraph-captured pass against eager launches (bitwise, replayed several times with
eager work in between), and the autotuned work queue against the default (bitwise)."""
import numpy as np
import pytest
import torch

import oracle as O
import pgfem3d_amd as P
from conftest import rel_inf
from problems import MixedProblem

pytestmark = pytest.mark.gpu
TOL = 1e-10
COH = P.Cohesive(sig_c=500.0, chi_c=1e-5, beta=0.5, alpha=1.0)


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _handle(prob):
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.R, prob.grain)
    if prob.mesh.n_coel:
        h.set_cohesive_props(prob.coh)
    return h


def test_mixed_batch_against_oracle(device):
    prob = MixedProblem(3, 2, 3, P.fcc_crystal(), COH, n_grains=8, seed=3)
    assert prob.mesh.n_coel > 0 and set(np.unique(np.diff(prob.mesh.elem_ptr))) == {4, 8, 10}
    h = _handle(prob)
    st = O.CPState(prob.nqp, prob.mats[0].g0)
    cst = O.CohState(prob.ncqp)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(int(prob.Ap[-1]), dtype=torch.float64, device=device)
    for step in range(1, 5):
        r, defl = prob.displacement(1.5e-4 * step, 2e-5, seed=step)
        assert h.residual_tangent(f, vals, _dev(r, device), defl=_dev(defl, device)) == 0
        fo, vo, err = O.assemble_cp(prob.mesh, prob.mats, prob.grain, prob.gP, prob.gL, prob.ids, prob.ndof, prob.Ap,
                                    prob.Ai, st, r, defl=defl)
        err += prob.assemble_coh(cst, r, defl, fo, vo)
        assert err == 0
        assert rel_inf(f.cpu().numpy(), fo) < TOL, step
        assert rel_inf(vals.cpu().numpy(), vo) < TOL, step
        assert rel_inf(h.get_state(P.STATE_PLASTIC_NP1), st.pF_np1) < TOL
        h.commit_step()
        st.commit()
        cst.commit()
    h.close()


def test_graph_replay_and_autotune_are_bitwise_eager(device):
    prob = MixedProblem(4, 2, 3, P.fcc_crystal(), COH, n_grains=8, seed=5)
    h = _handle(prob)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(int(prob.Ap[-1]), dtype=torch.float64, device=device)
    r, defl = prob.displacement(3e-4, 2e-5, seed=1)
    rd, dd = _dev(r, device), _dev(defl, device)
    assert h.residual(f, rd, defl=dd) == 0
    h.commit_step()
    r, defl = prob.displacement(6e-4, 2e-5, seed=2)
    rd, dd = _dev(r, device), _dev(defl, device)
    assert h.residual_tangent(f, vals, rd, defl=dd) == 0
    f0, v0 = f.cpu().numpy().copy(), vals.cpu().numpy().copy()
    h.set_graph(True)
    for rep in range(3):
        f.fill_(np.nan)
        vals.fill_(np.nan)
        assert h.residual_tangent(f, vals, rd, defl=dd) == 0
        assert np.array_equal(f.cpu().numpy(), f0) and np.array_equal(vals.cpu().numpy(), v0), rep
        # eager work of the program between replays
        g = torch.empty_like(f)
        h.set_graph(False)
        assert h.residual(g, rd, defl=dd) == 0
        assert np.array_equal(g.cpu().numpy(), f0)
        h.set_graph(True)
    h.set_graph(False)
    chunk = h.autotune(rd, defl=dd)
    assert chunk in (1, 64, 128, 256, 512, 1024, 2048)  # 1: one global queue
    assert h.residual_tangent(f, vals, rd, defl=dd) == 0
    assert np.array_equal(f.cpu().numpy(), f0) and np.array_equal(vals.cpu().numpy(), v0)
    h.close()


def test_setters_drop_the_captured_pass(device):
    """New orientations and interface properties replace device tables that a captured pass
    reads: with the graph on, the pass after each setter equals an eager pass of a fresh
    handle holding the new data (the setters drop the graph; the next pass captures again)."""
    prob = MixedProblem(3, 2, 3, P.fcc_crystal(), COH, n_grains=8, seed=7)
    h = _handle(prob)
    h.set_graph(True)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(int(prob.Ap[-1]), dtype=torch.float64, device=device)
    r, defl = prob.displacement(3e-4, 2e-5, seed=1)
    rd, dd = _dev(r, device), _dev(defl, device)
    assert h.residual_tangent(f, vals, rd, defl=dd) == 0
    assert h.residual_tangent(f, vals, rd, defl=dd) == 0  # a replay
    rng = np.random.default_rng(9)
    R2 = np.stack([np.linalg.qr(rng.standard_normal((3, 3)))[0] for _ in range(len(prob.R))])
    R2 *= np.sign(np.linalg.det(R2))[:, None, None]
    coh2 = P.Cohesive(sig_c=300.0, chi_c=2e-5, beta=0.5, alpha=1.0)
    h.set_orientations(R2, prob.grain)
    h.set_cohesive_props([coh2])
    f.fill_(np.nan)
    vals.fill_(np.nan)
    assert h.residual_tangent(f, vals, rd, defl=dd) == 0
    fresh = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    fresh.set_orientations(R2, prob.grain)
    fresh.set_cohesive_props([coh2])
    f2 = torch.empty_like(f)
    v2 = torch.empty_like(vals)
    assert fresh.residual_tangent(f2, v2, rd, defl=dd) == 0
    assert torch.equal(f, f2) and torch.equal(vals, v2)
    h.set_graph(False)
    h.close()
    fresh.close()
