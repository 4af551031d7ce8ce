"""The CM3F three-field element on the device (pgf_set_three_field) against the oracle's
restatement of the reference (orc_assemble_tf / orc_tf_update / orc_tf_initial, pinned in
test_oracle_three_field.py), at the same per-element volume and pressure variables:

  * the condensed residual and tangent and the model's stored F(n+1) = (theta/J)^1/3 Fr, in
    both assembly modes, within 1e-10 of the oracle's largest entry (1e-12 for F);
  * the element update (ddV, ddP for a Newton increment) within 1e-10 relative, and the
    initial conditions within 1e-12;
  * the prescribed-load pass: f_defl = K(free rows, prescribed columns) defl_d of the oracle's
    condensed tangent over the expanded numbering;
  * Newton load steps (device driver, device linear solve) against the oracle's CM3F Newton:
    displacements within 1e-9 relative, theta and P within 1e-9; on linear tets the converged
    steps equal the displacement formulation's;
  * commit and reset: n+1 <- n+1 + increments, n <- n+1, n-1 <- n, increments zeroed;
  * a mesh of all three element types (three device blocks: block offsets, device order);
  * checkpoint/resume through the reference's restart files, and hipGraph replays;
  * load steps opened by the reference's predictor land on the steps without it.
"""
import numpy as np
import pytest
import torch

import newton as NW
import oracle as O
import pgfem3d_amd as P
from problems import BoxProblem, expand_prescribed, oracle_tf_newton_step

pytestmark = pytest.mark.gpu
HE = P.neo_hookean(200e3, 0.3)


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(device)


def _handle(prob, device, assembly=P.ASSEMBLY_GATHER):
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_assembly(assembly)
    h.set_three_field(True)
    return h


def _fields(ne, seed):
    rng = np.random.default_rng(seed)
    return (1.0 + 0.01 * rng.standard_normal(ne), 300.0 * rng.standard_normal(ne),
            1e-3 * rng.standard_normal(ne), 10.0 * rng.standard_normal(ne))


@pytest.mark.parametrize("assembly", [P.ASSEMBLY_GATHER, P.ASSEMBLY_COLOURED])
@pytest.mark.parametrize("etype,n", [(4, 3), (10, 2), (8, 3)])
def test_condensed_pass_and_update(device, etype, n, assembly):
    prob = BoxProblem(n, etype, HE, seed=2)
    r, defl = prob.displacement(0.03, 0.01, seed=5)
    ne = prob.mesh.n_elem
    V, Pr, dV, dP = _fields(ne, 7)
    h = _handle(prob, device, assembly)
    for fld, v in ((P.TF_V_NP1, V), (P.TF_P_NP1, Pr), (P.TF_DV, dV), (P.TF_DP, dP)):
        h.tf_set(fld, v)
    f = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.zeros(h.nnz, dtype=torch.float64, device=device)
    assert h.residual_tangent(f, vals, _dev(r, device), defl=_dev(defl, device)) == 0
    Fo = np.zeros((9, prob.nqp))
    fo, vo, err = O.assemble_tf(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, V + dV, Pr + dP, r,
                                defl=defl, F_np1=Fo)
    assert err == 0
    fg, vg = f.cpu().numpy(), vals.cpu().numpy()
    assert np.abs(fg - fo).max() <= 1e-10 * np.abs(fo).max()
    assert np.abs(vg - vo).max() <= 1e-10 * np.abs(vo).max()
    assert np.abs(h.get_state(P.STATE_F_NP1) - Fo).max() < 1e-12
    # the tangent-only and residual-only passes give the fused pass's values
    v2 = torch.zeros_like(vals)
    f2 = torch.zeros_like(f)
    assert h.tangent(v2, _dev(r, device), defl=_dev(defl, device)) == 0
    assert h.residual(f2, _dev(r, device), defl=_dev(defl, device)) == 0
    assert torch.equal(v2, vals) and torch.equal(f2, f)
    # element update for a Newton increment
    du = 1e-3 * np.random.default_rng(9).standard_normal(prob.ndof)
    h.tf_update(_dev(du, device))
    ddV, ddP, err = O.tf_update(prob.mesh, prob.mats, prob.ids, V + dV, Pr + dP, du, r, defl=defl)
    assert err == 0
    assert np.abs(h.tf_get(P.TF_DDV) - ddV).max() <= 1e-10 * np.abs(ddV).max()
    assert np.abs(h.tf_get(P.TF_DDP) - ddP).max() <= 1e-10 * np.abs(ddP).max()
    # update_increments_from_NR(gamma): ddV *= gamma, dV += ddV; a trial taken back with
    # gamma = 0 restores dV exactly
    full_V, full_P = h.tf_get(P.TF_DDV), h.tf_get(P.TF_DDP)
    h.tf_accept(0.5)
    assert np.array_equal(h.tf_get(P.TF_DDV), 0.5 * full_V) and np.array_equal(h.tf_get(P.TF_DDP), 0.5 * full_P)
    assert np.array_equal(h.tf_get(P.TF_DV), dV + 0.5 * full_V)
    assert np.array_equal(h.tf_get(P.TF_DP), dP + 0.5 * full_P)
    h.tf_accept(0.0)
    assert np.array_equal(h.tf_get(P.TF_DV), dV) and np.array_equal(h.tf_get(P.TF_DP), dP)
    h.tf_accept(1.0)
    assert np.array_equal(h.tf_get(P.TF_DDV), full_V) and np.array_equal(h.tf_get(P.TF_DV), dV + full_V)
    h.close()


def test_prescribed_load(device):
    prob = BoxProblem(2, 10, HE, seed=3)
    r, defl = prob.displacement(0.02, 0.01, seed=1)
    ne = prob.mesh.n_elem
    V, Pr, dV, dP = _fields(ne, 11)
    h = _handle(prob, device)
    for fld, v in ((P.TF_V_NP1, V), (P.TF_P_NP1, Pr), (P.TF_DV, dV), (P.TF_DP, dP)):
        h.tf_set(fld, v)
    defl_d = np.array([0.004])
    fd = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    assert h.prescribed_load(fd, _dev(r, device), _dev(defl_d, device), defl=_dev(defl, device)) == 0
    h.close()
    import scipy.sparse as sp
    ids2, ndof2, Ap2, Ai2 = expand_prescribed(prob, len(defl))
    _, v2, err = O.assemble_tf(prob.mesh, prob.mats, ids2, ndof2, Ap2, Ai2, V + dV, Pr + dP,
                               np.concatenate([r, defl]), integrate=False)
    assert err == 0
    K2 = sp.csr_matrix((v2, Ai2, Ap2), shape=(ndof2, ndof2))
    ref = K2[:prob.ndof, prob.ndof:] @ defl_d
    assert np.abs(fd.cpu().numpy() - ref).max() <= 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("etype", [4, 10, 8])
def test_initial_conditions(device, etype):
    prob = BoxProblem(2, etype, HE, seed=4)
    r, defl = prob.displacement(0.03, 0.01, seed=2)
    r0, defl0 = prob.displacement(0.01, 0.005, seed=3)
    h = _handle(prob, device)
    h.tf_initial(_dev(r, device), _dev(defl, device), _dev(r0, device), _dev(defl0, device))

    def nodal(rr, dd):
        u = np.zeros_like(prob.mesh.X)
        m = prob.ids > 0
        u[m] = rr[prob.ids[m] - 1]
        u[prob.ids < 0] = dd[-prob.ids[prob.ids < 0] - 1]
        return u

    V, Pr, _ = O.tf_initial(prob.mesh, prob.mats, nodal(r, defl))
    V0, P0, _ = O.tf_initial(prob.mesh, prob.mats, nodal(r0, defl0))
    for fld, ref in ((P.TF_V_NP1, V), (P.TF_V_N, V), (P.TF_V_NM1, V0)):
        assert np.abs(h.tf_get(fld) - ref).max() < 1e-12
    for fld, ref in ((P.TF_P_NP1, Pr), (P.TF_P_N, Pr), (P.TF_P_NM1, P0)):
        assert np.abs(h.tf_get(fld) - ref).max() < 1e-12 * np.abs(Pr).max()
    assert not h.tf_get(P.TF_DV).any() and not h.tf_get(P.TF_DP).any()
    # the model's F(n), F(n-1) from the same displacements (compute_cm_initial_conditions)
    for fld, rr, dd in ((P.STATE_F_N, r, defl), (P.STATE_F_NM1, r0, defl0)):
        st = O.OracleState(prob.nqp, False)
        O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, rr, defl=dd)
        assert np.abs(h.get_state(fld) - st.F_np1).max() < 1e-12
        assert np.abs(st.F_np1 - np.eye(3).reshape(9, 1)).max() > 1e-3  # a nonzero displacement
    h.close()


def test_commit_and_reset(device):
    prob = BoxProblem(2, 4, HE, seed=1)
    ne = prob.mesh.n_elem
    V, Pr, dV, dP = _fields(ne, 5)
    h = _handle(prob, device)
    V0 = h.tf_get(P.TF_V_N)
    assert np.all(V0 == 1.0) and not h.tf_get(P.TF_P_N).any()
    for fld, v in ((P.TF_V_NP1, V), (P.TF_P_NP1, Pr), (P.TF_DV, dV), (P.TF_DP, dP), (P.TF_DDV, dV)):
        h.tf_set(fld, v)
    h.reset_step()
    assert not h.tf_get(P.TF_DV).any() and not h.tf_get(P.TF_DDV).any() and not h.tf_get(P.TF_DP).any()
    h.tf_set(P.TF_DV, dV)
    h.tf_set(P.TF_DP, dP)
    h.commit_step()
    assert np.array_equal(h.tf_get(P.TF_V_NP1), V + dV) and np.array_equal(h.tf_get(P.TF_V_N), V + dV)
    assert np.array_equal(h.tf_get(P.TF_P_N), Pr + dP) and np.array_equal(h.tf_get(P.TF_V_NM1), V0)
    assert not h.tf_get(P.TF_DV).any() and not h.tf_get(P.TF_DP).any()
    h.close()


@pytest.mark.parametrize("mat", [P.j2_damage(200e3, 0.3, 300.0, 1000.0, mu=2.0, p1=2.0, p2=1.5, Yin=2.0),
                                 P.Material(model=P.MODEL_HYPER_ELASTICITY, devPotFlag=2, volPotFlag=2, E=200e3,
                                            nu=0.3, G=80e3)],
                         ids=["j2-damage", "linear-law"])
def test_three_field_refuses_laws_without_its_split(device, mat):
    prob = BoxProblem(2, 4, mat, seed=1)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    with pytest.raises(P.PGFError):
        h.set_three_field(True)
    h.close()


@pytest.mark.parametrize("etype,n", [(10, 2), (8, 3), (4, 3)])
def test_newton_load_steps(device, etype, n):
    prob = BoxProblem(n, etype, HE, seed=3)
    h = _handle(prob, device)
    drv = NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=1e-12, max_iter=30)
    r = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    ro = np.zeros(prob.ndof)
    tf = dict(V=np.ones(prob.mesh.n_elem), P=np.zeros(prob.mesh.n_elem))
    state = {"NORM": 0.0, "first": True}
    for s in range(3):
        defl = np.array([0.02 * (s + 1)])
        res = drv.load_step(r, _dev(defl, device))
        assert res.converged, (s, res.norms)
        ro, its = oracle_tf_newton_step(prob, ro, defl, tf, tol=1e-12, state=state)
        assert abs(res.iterations - its) <= 1, (res.iterations, its)
        u = r.cpu().numpy()
        assert np.abs(u - ro).max() <= 1e-9 * np.abs(ro).max()
        assert np.abs(h.tf_get(P.TF_V_N) - tf["V"]).max() < 1e-9
        assert np.abs(h.tf_get(P.TF_P_N) - tf["P"]).max() < 1e-9 * np.abs(tf["P"]).max()
    h.close()
    if etype == 4:  # one QP per element: the displacement formulation's solution
        h1 = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
        d1 = NW.NewtonDriver(h1, prob.Ap, prob.Ai, tol=1e-12, max_iter=30)
        r1 = torch.zeros_like(r)
        for s in range(3):
            assert d1.load_step(r1, _dev(np.array([0.02 * (s + 1)]), device)).converged
        h1.close()
        assert (r1 - r).abs().max().item() <= 1e-9 * r1.abs().max().item()


def test_checkpoint_and_resume(device, tmp_path):
    """Three CM3F load steps, stopped after the second: the state written as the reference's
    restart files (with the volume and pressure records), read into a fresh handle and driver
    and run on, ends bit for bit on the uninterrupted run."""
    from types import SimpleNamespace

    import restart as RS

    prob = BoxProblem(2, 10, HE, seed=5)
    m = prob.mesh
    deck = SimpleNamespace(parts=[SimpleNamespace(nn=m.n_nodes, ne=m.n_elem)], node_of=[np.arange(m.n_nodes)],
                           elem_of=[np.arange(m.n_elem)], elem_ptr=m.elem_ptr, X=m.X)

    def new():
        h = _handle(prob, device)
        return h, NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=1e-12, max_iter=30)

    def steps(drv, r, s0, s1):
        for s in range(s0, s1):
            assert drv.load_step(r, _dev(np.array([0.02 * (s + 1)]), device)).converged

    h, drv = new()
    r = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    steps(drv, r, 0, 3)
    full = (r.cpu().numpy(), h.tf_get(P.TF_V_N), h.tf_get(P.TF_P_N), h.get_state(P.STATE_F_N))
    h.close()

    def nodal(rr, d):
        u = np.zeros_like(m.X)
        msk = prob.ids > 0
        u[msk] = rr[prob.ids[msk] - 1]
        u[prob.ids < 0] = d
        return u

    h, drv = new()
    r = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    steps(drv, r, 0, 1)
    r1 = r.cpu().numpy()
    steps(drv, r, 1, 2)
    RS.save(str(tmp_path), "box", 1, deck, h, nodal(r1, 0.02), nodal(r.cpu().numpy(), 0.04), (0.0, 1.0, 2.0), drv.NORM)
    norm = drv.NORM
    h.close()

    h, drv = new()
    _, u_n, _, norm2 = RS.load(str(tmp_path), "box", 1, deck, h)
    assert norm2 == norm
    drv.NORM, drv.first = norm2, False
    r = _dev(u_n[prob.ids > 0][np.argsort(prob.ids[prob.ids > 0])], device)
    steps(drv, r, 2, 3)
    assert np.array_equal(r.cpu().numpy(), full[0])
    assert np.array_equal(h.tf_get(P.TF_V_N), full[1]) and np.array_equal(h.tf_get(P.TF_P_N), full[2])
    assert np.array_equal(h.get_state(P.STATE_F_N), full[3])
    h.close()


class _MixedHE(BoxProblem):
    """Linear tets, quadratic tets and hexes side by side in one neo-Hookean mesh: three device
    blocks, so the per-element variables sit at block offsets and in device order."""

    def __init__(self):
        from problems import mixed_mesh

        self.mesh, _, _ = mixed_mesh(2, 2, 2, 3, seed=4, interfaces=False)
        self.supports = P.box_supports(self.mesh)
        self.ids, self.ndof = P.generate_dof_ids(self.mesh, self.supports)
        self.Ap, self.Ai = P.psparse_apai(self.mesh, self.ids, self.ndof)
        self.mats = [HE]
        self.nqp = int(sum(O.nint(int(k)) for k in np.diff(self.mesh.elem_ptr)))


def test_mixed_element_types(device):
    prob = _MixedHE()
    r, defl = prob.displacement(0.03, 0.01, seed=6)
    V, Pr, dV, dP = _fields(prob.mesh.n_elem, 13)
    h = _handle(prob, device)
    for fld, v in ((P.TF_V_NP1, V), (P.TF_P_NP1, Pr), (P.TF_DV, dV), (P.TF_DP, dP)):
        h.tf_set(fld, v)
    assert np.array_equal(h.tf_get(P.TF_V_NP1), V) and np.array_equal(h.tf_get(P.TF_DP), dP)
    f = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.zeros(h.nnz, dtype=torch.float64, device=device)
    assert h.residual_tangent(f, vals, _dev(r, device), defl=_dev(defl, device)) == 0
    Fo = np.zeros((9, prob.nqp))
    fo, vo, err = O.assemble_tf(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, V + dV, Pr + dP, r,
                                defl=defl, F_np1=Fo)
    assert err == 0
    assert np.abs(f.cpu().numpy() - fo).max() <= 1e-10 * np.abs(fo).max()
    assert np.abs(vals.cpu().numpy() - vo).max() <= 1e-10 * np.abs(vo).max()
    assert np.abs(h.get_state(P.STATE_F_NP1) - Fo).max() < 1e-12
    du = 1e-3 * np.random.default_rng(2).standard_normal(prob.ndof)
    h.tf_update(_dev(du, device))
    ddV, ddP, _ = O.tf_update(prob.mesh, prob.mats, prob.ids, V + dV, Pr + dP, du, r, defl=defl)
    assert np.abs(h.tf_get(P.TF_DDV) - ddV).max() <= 1e-10 * np.abs(ddV).max()
    assert np.abs(h.tf_get(P.TF_DDP) - ddP).max() <= 1e-10 * np.abs(ddP).max()
    h.close()
    # Newton load steps from rest
    h = _handle(prob, device)
    drv = NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=1e-12, max_iter=30)
    rr = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    ro = np.zeros(prob.ndof)
    tf = dict(V=np.ones(prob.mesh.n_elem), P=np.zeros(prob.mesh.n_elem))
    state = {"NORM": 0.0, "first": True}
    for s in range(2):
        d = np.array([0.02 * (s + 1)])
        assert drv.load_step(rr, _dev(d, device)).converged
        ro, _ = oracle_tf_newton_step(prob, ro, d, tf, tol=1e-12, state=state)
        assert np.abs(rr.cpu().numpy() - ro).max() <= 1e-9 * np.abs(ro).max()
        assert np.abs(h.tf_get(P.TF_V_N) - tf["V"]).max() < 1e-9
        assert np.abs(h.tf_get(P.TF_P_N) - tf["P"]).max() < 1e-9 * np.abs(tf["P"]).max()
    h.close()


def test_graph_replays_match_eager(device):
    """The condensed pass captured in a hipGraph and replayed (with eager work between the
    replays) gives the eager pass's values on every replay."""
    prob = BoxProblem(2, 10, HE, seed=8)
    r, defl = prob.displacement(0.02, 0.01, seed=3)
    V, Pr, dV, dP = _fields(prob.mesh.n_elem, 3)
    h = _handle(prob, device)
    for fld, v in ((P.TF_V_NP1, V), (P.TF_P_NP1, Pr), (P.TF_DV, dV), (P.TF_DP, dP)):
        h.tf_set(fld, v)
    rd, dd = _dev(r, device), _dev(defl, device)
    f0 = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
    v0 = torch.zeros(h.nnz, dtype=torch.float64, device=device)
    assert h.residual_tangent(f0, v0, rd, defl=dd) == 0
    h.set_graph(True)
    f = torch.zeros_like(f0)
    v = torch.zeros_like(v0)
    for _ in range(3):
        f.fill_(7.0)
        v.fill_(-1.0)
        assert h.residual_tangent(f, v, rd, defl=dd) == 0
        torch.cuda.synchronize()
        assert torch.equal(f, f0) and torch.equal(v, v0)
    h.close()


def test_predictor_steps_reach_the_same_solution(device):
    """Load steps that open with the reference's predictor (K du = -f_defl at (u_n, defl_n);
    theta and P follow in the iterations after it) converge to the steps taken from the full
    prescribed values."""
    prob = BoxProblem(2, 10, HE, seed=9)
    out = []
    for pred in (False, True):
        h = _handle(prob, device)
        drv = NW.NewtonDriver(h, prob.Ap, prob.Ai, tol=1e-12, max_iter=30, computer_zero=1e-9)
        r = torch.zeros(prob.ndof, dtype=torch.float64, device=device)
        for s in range(3):
            if pred:
                res = drv.load_step(r, _dev(np.array([0.02 * s]), device), defl_d=_dev(np.array([0.02]), device))
            else:
                res = drv.load_step(r, _dev(np.array([0.02 * (s + 1)]), device))
            assert res.converged, (pred, s, res.norms)
        out.append((r.cpu().numpy(), h.tf_get(P.TF_V_N), h.tf_get(P.TF_P_N)))
        h.close()
    (r0, V0, P0), (r1, V1, P1) = out
    assert np.abs(r1 - r0).max() <= 1e-8 * np.abs(r0).max()
    assert np.abs(V1 - V0).max() < 1e-9 and np.abs(P1 - P0).max() <= 1e-8 * np.abs(P0).max()
