"""The reference's crystal-plasticity tension deck (tests/golden/cp_box_tension, copied from
verification_tests/crystal_plasticity by tests/golden/collect_cp_box.py) set up as a problem
for the oracle and the device: 8 partition files merged into one mesh of 89 nodes and 257
linear tets, one crystal-plasticity material whose elastic part is the hyperelastic law of the
material line (src/homogen.cc:183-190 maps Ex, Ey, Ez, Gyz, nyz to E, m10, m01, G, nu), one
orientation per element, 50 load steps of a prescribed stretch."""
import os
import types

import numpy as np

import oracle as O
import pgfem3d_amd as P
import ref_io as R
from problems import oracle_newton_step

DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cp_box_tension")
NPARTS = 8
# the last step the restart fixtures hold (ref_out/box_<p>_49.res)
FIXTURE_STEP = 49


def load(tol_M=1e-13, max_itr_M=80):
    deck = R.read_partitioned(DIR, "box_", NPARTS)
    sf = R.read_solver_file(os.path.join(DIR, "box.in.st"), len(deck.defl))
    mp = R.read_model_params(os.path.join(DIR, "model_params.in"))[0]
    m0 = deck.parts[0].materials[0]
    mat = P.Material(model=P.MODEL_CRYSTAL_PLASTICITY, devPotFlag=m0["devPotFlag"], volPotFlag=m0["volPotFlag"],
                     cp_elastic=1, E=m0["Ex"], nu=m0["nyz"], G=m0["Gyz"], m10=m0["Ey"], m01=m0["Ez"],
                     gamma_dot_0=mp["gamma_dot_0"], m_rate=mp["m"], G0=mp["G0"], g0=mp["g0"], gs_0=mp["gs_0"],
                     gamma_dot_s=mp["gamma_dot_s"], w_sat=mp["w"], tol_M=tol_M, max_itr_M=max_itr_M)
    ne = len(deck.elem_ptr) - 1
    mesh = P.MeshArrays(X=deck.X, elem_ptr=deck.elem_ptr, elem_conn=deck.elem_conn,
                        elem_mat=np.zeros(ne, np.int32))
    ids, ndof = O.generate_dof_ids(deck.X, deck.elem_ptr, deck.elem_conn, deck.support)
    Ap, Ai = O.psparse_apai(deck.X, deck.elem_ptr, deck.elem_conn, ids, ndof)
    ang = np.zeros((ne, 3))
    for p in range(NPARTS):
        _, a = R.read_orientations(os.path.join(DIR, "CO", f"co_{p}.in"))
        for (e, _ip), v in a.items():
            ang[deck.elem_of[p][e]] = v
    rot = np.stack([R.euler_rotation(*ang[e]) for e in range(ne)])
    return types.SimpleNamespace(deck=deck, sf=sf, mp=mp, mat=mat, mats=[mat], mesh=mesh, ids=ids, ndof=ndof,
                                 Ap=Ap, Ai=Ai, rot=rot, grain=np.arange(ne, dtype=np.int32), ne=ne)


def nodal_u(prob, r, defl):
    """Total displacement per node from free DOFs r and prescribed values defl."""
    u = np.zeros((prob.deck.X.shape[0], 3))
    ids = prob.ids
    free = ids > 0
    u[free] = r[ids[free] - 1]
    pres = ids < 0
    u[pres] = defl[-ids[pres] - 1]
    return u


def compare_with_fixture(prob, u, F_n, pF_n, g_n):
    """Largest differences from the step-49 restart files: (|du|, |dF|, |dpF|, relative dg),
    each a max over every node or QP of every partition.  F_n, pF_n are [9, ne], g_n [ne]."""
    du = dF = dpF = dg = 0.0
    deck = prob.deck
    for p in range(NPARTS):
        part = deck.parts[p]
        _, un, el = R.read_restart_cp(os.path.join(DIR, "ref_out", f"box_{p}_{FIXTURE_STEP}.res"), part.nn, 3,
                                      part.ne)
        du = max(du, float(np.abs(u[deck.node_of[p]] - un).max()))
        for e in range(part.ne):
            q = deck.elem_of[p][e]
            rec = el[e][0]
            dF = max(dF, float(np.abs(F_n[:, q] - rec.F_n).max()))
            dpF = max(dpF, float(np.abs(pF_n[:, q] - rec.pF_n).max()))
            dg = max(dg, abs(g_n[q] - rec.g_n) / rec.g_n)
    return du, dF, dpF, dg


def run_oracle(prob, rot, nsteps, tol=1e-10, predictor=False, iter_max=None, computer_zero=1e-15, alphas=None):
    """The deck's load steps over the oracle (Newton to tol on the normalised residual, sparse
    direct solve; with predictor each step opens with K du = -f_defl at (u_n, defl_n), as the
    reference's Newton does): (nodal u, CPState, prescribed values, iterations per step).
    alphas (a list) receives each step's physics-based subdivision measure,
    max over QPs of dt sum |gamma_dot| / 0.005 (CP_PARAM::get_subdiv_param)."""
    tabs = [O.cp_grain(prob.mat, x) for x in rot]
    gP = np.stack([t[0] for t in tabs])
    gL = np.stack([t[1] for t in tabs])
    st = O.CPState(prob.ne, prob.mat.g0)
    defl = np.zeros(len(prob.deck.defl))
    r = np.zeros(prob.ndof)
    state = dict(NORM=0.0, first=True)
    sf = prob.sf
    its = []
    for step in range(nsteps):
        inc = sf.increment(step, prob.deck.defl)
        dt = sf.times[step + 1] - sf.times[step]
        if predictor:
            r, it, _ = oracle_newton_step(prob, st, r, defl, state, tol=tol, max_iter=40, cp_tables=(gP, gL),
                                          dt=dt, defl_d=inc, iter_max=iter_max, computer_zero=computer_zero)
            defl = defl + inc
        else:
            defl = defl + inc
            r, it, _ = oracle_newton_step(prob, st, r, defl, state, tol=tol, max_iter=40, cp_tables=(gP, gL),
                                          dt=dt, iter_max=iter_max, computer_zero=computer_zero)
        its.append(it)
        if alphas is not None:
            alphas.append(float((np.abs(st.gdot).sum(axis=0) * dt / 0.005).max()))
    return nodal_u(prob, r, defl), st, defl, its


def free_dofs(prob, u):
    """The free-DOF vector of nodal values u [n_nodes, 3]."""
    r = np.zeros(prob.ndof)
    free = prob.ids > 0
    r[prob.ids[free] - 1] = u[free]
    return r
