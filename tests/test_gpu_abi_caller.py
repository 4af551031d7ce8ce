"""The C caller (tests/abi_caller.c, built against include/pgfem3d_amd.h alone) runs the
boundary end to end on the GPU -- pgf_mesh_box, pgf_generate_dof_ids, pgf_psparse_apai,
pgf_create, pgf_residual_tangent -- and its f and CSR values match the oracle on the same
problem within 1e-10 of the largest entry."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import pgfem3d_amd as P
from conftest import rel_inf
from problems import BoxProblem
from test_abi import CALLER

pytestmark = pytest.mark.gpu


def test_c_caller_against_oracle(device, tmp_path):
    out = tmp_path / "caller.bin"
    print("abi_caller: J2 tet10 2^3 box", flush=True)
    run = subprocess.run([CALLER, str(out)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, (run.returncode, run.stdout, run.stderr)
    raw = out.read_bytes()
    ndof, nnz, status = np.frombuffer(raw[:24], dtype=np.int64)
    prob = BoxProblem(2, P.PGF_TET10, P.j2_linear(70e3, 0.33, 250.0, 1000.0))
    assert ndof == prob.ndof and nnz == prob.Ap[-1] and status == 0
    f = np.frombuffer(raw[24:24 + 8 * ndof], dtype=np.float64)
    vals = np.frombuffer(raw[24 + 8 * ndof:], dtype=np.float64)
    r, defl = prob.displacement(0.012, 0.0, seed=0)
    st = O.OracleState(prob.nqp, True)
    fo, vo, err = O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r, defl=defl)
    assert err == 0
    assert np.abs(st.S_np1[0]).max() > 0  # the stretch is plastic: the J2 return map ran
    assert rel_inf(f, fo) < 1e-10 and rel_inf(vals, vo) < 1e-10
