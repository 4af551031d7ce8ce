"""Oracle parity at the workloads' full sizes (the smaller tests hold the same comparisons on
meshes of up to 10^3 cells):

  * J2 plasticity on 48^3 cells x 6 quadratic tets (663,552 elements, 2,654,208 QPs, the
    headline workload), one fused integrate + residual + tangent pass into plasticity from the
    virgin state, in the gather assembly and in the coloured one;
  * crystal plasticity on 55^3 cells x 6 linear tets (998,250 QPs) in 64 Voronoi grains, three
    load steps of 1e-3 with the state committed between them;
  * the mixed batch: 41^3 x 6 linear tets, 32^3 x 6 quadratic tets and 46^3 hexes in 64 grains
    with interface elements on the grain boundaries, two load steps;
  * the CM3F three-field element (one volume and one pressure per element, condensed): a
    neo-Hookean material on the J2 mesh (48^3 x 6 quadratic tets) and crystal plasticity with
    the hyperelastic law on the CP mesh, one pass each at seeded element volumes and pressures.

Each compares f to its largest entry, K to each CSR row's own largest entry and the state per
QP to its own size, within 1e-10 of the oracle (orc_assemble / orc_assemble_cp /
orc_assemble_coh on the host's cores: OpenMP over element colours, the same results as one
thread bit for bit).  The CSR pattern is the oracle's Psparse_ApAi restatement (pinned
bit-exactly in test_pattern.py)."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
import pgfem3d_amd as P
from conftest import rel_cols, rel_csr_rows, rel_inf
from problems import BoxProblem, CPBoxProblem, MixedProblem, fcc_crystal_he

pytestmark = pytest.mark.gpu
TOL = 1e-10
HOST_THREADS = 16  # the host share of one GPU


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    L = O.lib()
    L.orc_set_threads.restype = C.c_int
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_threads(HOST_THREADS)


def _dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _check(tag, f, vals, fo, vo, Ap):
    # f against its largest entry: an interior entry is a sum of element contributions that
    # cancel to near zero, so its own size is no scale for its rounding
    ef = rel_inf(f.cpu().numpy(), fo)
    ek = rel_csr_rows(vals.cpu().numpy(), vo, Ap)
    print(f"{tag}: f {ef:.2e} of max |f|, K {ek:.2e} of each row's max", flush=True)
    assert ef < TOL and ek < TOL, (tag, ef, ek)


def test_j2_p2_48cube(device):
    print("full size: J2 48^3 P2 tets", flush=True)
    prob = BoxProblem(48, P.PGF_TET10, P.j2_linear(70e3, 0.33, 250.0, 1000.0))
    assert prob.nqp == 2654208
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(h.nnz, dtype=torch.float64, device=device)
    r, defl = prob.displacement(0.015, 2e-4, seed=11)
    rd, dd = _dev(r, device), _dev(defl, device)
    st = O.OracleState(prob.nqp, True)
    fo, vo, err = O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r, defl=defl)
    assert err == 0
    for mode in (P.ASSEMBLY_GATHER, P.ASSEMBLY_COLOURED):
        h.set_assembly(mode)
        f.fill_(np.nan)
        vals.fill_(np.nan)
        assert h.residual_tangent(f, vals, rd, defl=dd) == 0
        _check(("j2", mode), f, vals, fo, vo, prob.Ap)
    assert rel_inf(h.get_state(P.STATE_F_NP1), st.F_np1) < TOL
    assert rel_cols(h.get_state(P.STATE_PLASTIC_NP1), st.P_np1) < TOL
    sc = h.get_state(P.STATE_SCALAR_NP1)
    # ep per QP to its own size, with an absolute floor of 1e-4 in strain: F is summed in another
    # order on the device and agrees with the oracle's to ~1e-15, and ep follows the strain with a
    # sensitivity of order one, so a point that has only just yielded (ep ~ 1e-5) carries that
    # 1e-15 as ~1e-10 of its own size
    ep, ep_o = sc[0], st.S_np1[0]
    assert (np.abs(ep - ep_o) / np.maximum(np.abs(ep_o), 1e-4)).max() < TOL
    assert np.abs(ep - ep_o).max() < 1e-14
    assert np.mean(st.S_np1[0] > 0) > 0.5  # most points yielded
    h.close()


def test_cp_p1_55cube_64_grains(device):
    print("full size: CP 55^3 P1 tets, 64 grains", flush=True)
    mat = P.fcc_crystal()
    prob = CPBoxProblem(55, P.PGF_TET4, mat, n_grains=64, seed=2)
    assert prob.nqp == 998250
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.R, prob.grain)
    st = O.CPState(prob.nqp, mat.g0)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(h.nnz, dtype=torch.float64, device=device)
    for s in range(1, 4):
        r, defl = prob.displacement(1e-3 * s, 0.0, seed=s)
        assert h.residual_tangent(f, vals, _dev(r, device), defl=_dev(defl, device)) == 0
        fo, vo, err = prob.assemble(st, r, defl)
        assert err == 0
        _check(("cp", s), f, vals, fo, vo, prob.Ap)
        assert rel_cols(h.get_state(P.STATE_PLASTIC_NP1), st.pF_np1) < TOL, s
        assert rel_cols(h.get_state(P.STATE_TAU), st.tau) < TOL, s
        assert rel_inf(h.get_state(P.STATE_SCALAR_NP1)[0], st.g_np1) < TOL, s
        its = h.get_state(P.STATE_ITERS)[0]
        assert np.mean(its == st.iters) > 0.95, s
        h.commit_step()
        st.commit()
    assert np.mean(np.abs(st.gdot).max(axis=0) > 1e-5) > 0.5  # most points slip
    h.close()


def test_mixed_batch_with_interfaces(device):
    print("full size: mixed batch 41^3 P1 + 32^3 P2 + 46^3 hexes with interfaces", flush=True)
    coh = P.Cohesive(sig_c=500.0, chi_c=1e-3, beta=0.5, alpha=1.0)
    prob = MixedProblem(41, 32, 46, P.fcc_crystal(), coh, n_grains=64, seed=2)
    assert prob.nqp == 1978646 and prob.mesh.n_coel > 0
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.R, prob.grain)
    h.set_cohesive_props(prob.coh)
    st = O.CPState(prob.nqp, prob.mats[0].g0)
    cst = O.CohState(prob.ncqp)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(h.nnz, dtype=torch.float64, device=device)
    for s in range(1, 3):
        r, defl = prob.displacement(1e-3 * s, 0.0, seed=s)
        assert h.residual_tangent(f, vals, _dev(r, device), defl=_dev(defl, device)) == 0
        fo, vo, err = O.assemble_cp(prob.mesh, prob.mats, prob.grain, prob.gP, prob.gL, prob.ids, prob.ndof, prob.Ap,
                                    prob.Ai, st, r, defl=defl)
        err += prob.assemble_coh(cst, r, defl, fo, vo)
        assert err == 0
        _check(("mixed", s), f, vals, fo, vo, prob.Ap)
        assert rel_cols(h.get_state(P.STATE_PLASTIC_NP1), st.pF_np1) < TOL, s
        h.commit_step()
        st.commit()
        cst.commit()
    h.close()


def _tf_fields(h, theta, Pe):
    for fld, v in ((P.TF_V_NP1, theta), (P.TF_P_NP1, Pe), (P.TF_DV, 0 * theta), (P.TF_DP, 0 * Pe)):
        h.tf_set(fld, v)


def test_cm3f_neo_hookean_p2_48cube(device):
    print("full size: CM3F neo-Hookean 48^3 P2 tets", flush=True)
    prob = BoxProblem(48, P.PGF_TET10, P.neo_hookean(70e3, 0.33))
    ne = prob.mesh.n_elem
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_three_field(True)
    rng = np.random.default_rng(21)
    theta = 1.0 + 1e-3 * rng.standard_normal(ne)
    Pe = 100.0 * rng.standard_normal(ne)
    _tf_fields(h, theta, Pe)
    r, defl = prob.displacement(0.015, 2e-4, seed=11)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(h.nnz, dtype=torch.float64, device=device)
    assert h.residual_tangent(f, vals, _dev(r, device), defl=_dev(defl, device)) == 0
    Fo = np.zeros((9, prob.nqp))
    fo, vo, err = O.assemble_tf(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, theta, Pe, r,
                                defl=defl, F_np1=Fo)
    assert err == 0
    _check("cm3f he", f, vals, fo, vo, prob.Ap)
    assert np.abs(h.get_state(P.STATE_F_NP1) - Fo).max() < 1e-12
    h.close()


def test_cm3f_crystal_plasticity_p1_55cube(device):
    print("full size: CM3F crystal plasticity 55^3 P1 tets, 64 grains", flush=True)
    mat = fcc_crystal_he()
    prob = CPBoxProblem(55, P.PGF_TET4, mat, n_grains=64, seed=2)
    ne = prob.mesh.n_elem
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.R, prob.grain)
    h.set_three_field(True)
    st = O.CPState(prob.nqp, mat.g0)
    rng = np.random.default_rng(22)
    theta = 1.0 + 2e-4 * rng.standard_normal(ne)
    Pe = 60.0 * rng.standard_normal(ne)
    _tf_fields(h, theta, Pe)
    r, defl = prob.displacement(1e-3, 0.0, seed=1)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=device)
    vals = torch.empty(h.nnz, dtype=torch.float64, device=device)
    assert h.residual_tangent(f, vals, _dev(r, device), defl=_dev(defl, device)) == 0
    fo, vo, err = O.assemble_cp_tf(prob.mesh, prob.mats, prob.grain, prob.gP, prob.gL, prob.ids, prob.ndof, prob.Ap,
                                   prob.Ai, st, theta, Pe, r, defl=defl)
    assert err == 0
    _check("cm3f cp", f, vals, fo, vo, prob.Ap)
    assert np.abs(h.get_state(P.STATE_F_NP1) - st.F_np1).max() < 1e-12
    assert rel_cols(h.get_state(P.STATE_PLASTIC_NP1), st.pF_np1) < TOL
    assert np.mean(np.abs(st.gdot).max(axis=0) > 1e-5) > 0.5  # most points slip
    h.close()
