"""The reference's box patch test (tests/golden/patch_tension, copied from
verification_tests/box_patch_{LT,QT}/tension_disp by tests/golden/collect_patch.py) set up as a
problem for the oracle and the device: a unit box of 24 linear (LT) or quadratic (QT) tets,
u_y = u_z = 0 on every node, u_x = 0 on x = 0 and the prescribed u_x = 1 on x = 1 (one load
step), and the material line Ex 3, Gyz 1.2, nyz 0.25 with the linear deviatoric law
(devPotFlag 2) and the quadratic volumetric potential (volPotFlag 99), mapped as
src/homogen.cc:183-190 maps them.  The reference's check (verification_tests/src/tension_disp.cc,
post_processing_compute_stress, src/post_processing.cc:103-181) is the volume average of the
PK2 stress S_xx over the box: 4.6 (verification_tests/output/stress_tension_disp.out)."""
import os
import types

import numpy as np

import oracle as O
import pgfem3d_amd as P
import ref_io as R

DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "patch_tension")


def expected():
    with open(os.path.join(DIR, "stress_tension_disp.out")) as fh:
        return float(fh.read().split()[0])


def load(lt):
    deck = R.read_partitioned(DIR, f"box_{lt}_", 1)
    sf = R.read_solver_file(os.path.join(DIR, f"box_{lt}_0.in.st"), len(deck.defl))
    m0 = deck.parts[0].materials[0]
    mat = P.Material(model=P.MODEL_HYPER_ELASTICITY, devPotFlag=m0["devPotFlag"], volPotFlag=m0["volPotFlag"],
                     E=m0["Ex"], nu=m0["nyz"], G=m0["Gyz"], m10=m0["Ey"], m01=m0["Ez"])
    ne = len(deck.elem_ptr) - 1
    mesh = P.MeshArrays(X=deck.X, elem_ptr=deck.elem_ptr, elem_conn=deck.elem_conn, elem_mat=np.zeros(ne, np.int32))
    ids, ndof = O.generate_dof_ids(deck.X, deck.elem_ptr, deck.elem_conn, deck.support)
    Ap, Ai = O.psparse_apai(deck.X, deck.elem_ptr, deck.elem_conn, ids, ndof)
    nqp = sum(O.nint(int(deck.elem_ptr[e + 1] - deck.elem_ptr[e])) for e in range(ne))
    return types.SimpleNamespace(deck=deck, sf=sf, mat=mat, mats=[mat], mesh=mesh, ids=ids, ndof=ndof, Ap=Ap, Ai=Ai,
                                 nqp=nqp, ne=ne)


def average_S_xx(prob, S_qp):
    """The detJ w average of the per-QP PK2 S_xx (post_processing_compute_stress); S_qp is
    [n_qp, 6] in mesh QP order, S_xx first."""
    num = vol = 0.0
    q = 0
    for e in range(prob.ne):
        conn = prob.mesh.elem_conn[prob.mesh.elem_ptr[e]:prob.mesh.elem_ptr[e + 1]]
        x = prob.mesh.X[conn]
        for ip in range(O.nint(len(conn))):
            w = O.basis(len(conn), x, ip)[2]
            num += w * S_qp[q, 0]
            vol += w
            q += 1
    return num / vol
