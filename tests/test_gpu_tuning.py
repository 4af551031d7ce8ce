"""The launch configuration (pgf_autotune, pgf_get_tuning, pgf_set_tuning): every gather tile
(nodes per workgroup x staging blocks in flight) and every points-per-wavefront of the
crystal-plasticity work queue gives the same f and K bit for bit, eager and under graph
replay, the autotuner's choices are valid configurations, and invalid ones are refused."""
import numpy as np
import pytest
import torch

import pgfem3d_amd as P
from problems import BoxProblem, MixedProblem

pytestmark = pytest.mark.gpu
COH = P.Cohesive(sig_c=500.0, chi_c=1e-5, beta=0.5, alpha=1.0)


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _pass(h, f, vals, rd, dd):
    f.fill_(np.nan)
    vals.fill_(np.nan)
    assert h.residual_tangent(f, vals, rd, defl=dd) == 0
    return f.cpu().numpy().copy(), vals.cpu().numpy().copy()


def test_every_gather_tile_gives_the_same_sums(device):
    prob = BoxProblem(3, P.PGF_TET10, P.j2_linear(70e3, 0.33, 250.0, 1000.0), seed=4)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(h.nnz, dtype=torch.float64, device=device)
    r, defl = prob.displacement(0.012, 2e-4, seed=3)
    rd, dd = _dev(r, device), _dev(defl, device)
    f0, v0 = _pass(h, f, vals, rd, dd)
    assert h.tuning()["gather_batch"] == 4
    for graph in (False, True):
        h.set_graph(graph)
        for w in (1, 2, 4):
            for b in (2, 4, 8):
                h.set_tuning(gather_waves=w, gather_batch=b)
                t = h.tuning()
                assert t["gather_waves"] <= w and t["gather_batch"] == b
                f1, v1 = _pass(h, f, vals, rd, dd)
                assert np.array_equal(f1, f0) and np.array_equal(v1, v0), (graph, w, b)
    h.set_graph(False)
    chunk = h.autotune(rd, defl=dd)
    t = h.tuning()
    assert chunk == t["cp_chunk"] and t["gather_waves"] in (1, 2, 4) and t["gather_batch"] in (2, 4, 8)
    f1, v1 = _pass(h, f, vals, rd, dd)
    assert np.array_equal(f1, f0) and np.array_equal(v1, v0)
    for bad in (dict(cp_chunk=8), dict(gather_waves=3), dict(gather_batch=5)):
        with pytest.raises(P.PGFError):
            h.set_tuning(**bad)
    h.close()


def test_points_per_wavefront_and_gather_tile_on_a_mixed_batch(device):
    prob = MixedProblem(3, 2, 2, P.fcc_crystal(), COH, n_grains=6, seed=8)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.R, prob.grain)
    h.set_cohesive_props(prob.coh)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(h.nnz, dtype=torch.float64, device=device)
    r, defl = prob.displacement(4e-4, 2e-5, seed=2)
    rd, dd = _dev(r, device), _dev(defl, device)
    f0, v0 = _pass(h, f, vals, rd, dd)
    for c, w, b in ((64, 1, 2), (2048, 2, 8), (1, 2, 4), (512, 4, 4)):  # 1: one global queue
        h.set_tuning(cp_chunk=c, gather_waves=w, gather_batch=b)
        f1, v1 = _pass(h, f, vals, rd, dd)
        assert np.array_equal(f1, f0) and np.array_equal(v1, v0), (c, w, b)
    h.autotune(rd, defl=dd)
    f1, v1 = _pass(h, f, vals, rd, dd)
    assert np.array_equal(f1, f0) and np.array_equal(v1, v0)
    h.close()
