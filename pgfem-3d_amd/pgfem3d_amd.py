"""ctypes binding of libpgfem3d_amd.so (the C ABI declared in include/pgfem3d_amd.h).

The library is the product: HIP kernels for gfx950 behind a C ABI.  This module only moves
arrays across the boundary; device vectors are torch tensors on the GPU whose data pointers
are handed to the library.  There is no fallback: if the library is missing or no GPU is
present, the calls raise.

Naming follows the reference (C-SWARM/pgfem-3d): ``generate_dof_ids`` mirrors
generate_local_dof_ids (src/generate_dof_ids.cc:69-171), ``psparse_apai`` mirrors
Psparse_ApAi (src/Psparse_ApAi.cc:238-670), ``Handle.residual`` mirrors fd_residuals_MP
(src/fd_residuals.cc:468-636) and ``Handle.tangent`` mirrors stiffmat_fd_MP
(src/stiffmat_fd.cc:831-993).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PGFEM3D_AMD_LIB") or os.path.join(_HERE, "_build", "libpgfem3d_amd.so")

PGF_TET4, PGF_TET10, PGF_HEX8 = 4, 10, 8
MODEL_HYPER_ELASTICITY, MODEL_CRYSTAL_PLASTICITY, MODEL_J2_PLASTICITY = 0, 1, 4
ASSEMBLY_GATHER, ASSEMBLY_COLOURED = 0, 1
STATE_F_NP1, STATE_F_N, STATE_F_NM1 = 0, 1, 2
STATE_PLASTIC_NP1, STATE_PLASTIC_N, STATE_PLASTIC_NM1 = 3, 4, 5
STATE_SCALAR_NP1, STATE_SCALAR_N, STATE_SCALAR_NM1 = 6, 7, 8
STATE_TAU, STATE_TAU_N, STATE_GDOT, STATE_GDOT_N, STATE_ITERS = 9, 10, 11, 12, 13
STATE_COHESIVE_NP1, STATE_COHESIVE_N = 14, 15
TF_V_NP1, TF_V_N, TF_V_NM1, TF_P_NP1, TF_P_N, TF_P_NM1, TF_DV, TF_DDV, TF_DP, TF_DDP = range(10)

ERRORS = {-1: "PGF_EINVAL", -2: "PGF_ENOMEM", -3: "PGF_EHIP", -4: "PGF_ENOTSUP", -5: "PGF_ENODEV",
          -6: "PGF_ESCRATCH"}
STATIC_STAGE_DOUBLES = 8388608  # PGF_STATIC_STAGE_DOUBLES


class PGFError(RuntimeError):
    pass


class Material(C.Structure):
    """pgf_material (include/pgfem3d_amd.h)."""

    _fields_ = [
        ("model", C.c_int32), ("devPotFlag", C.c_int32), ("volPotFlag", C.c_int32), ("cp_elastic", C.c_int32),
        ("E", C.c_double), ("nu", C.c_double), ("G", C.c_double), ("m10", C.c_double), ("m01", C.c_double),
        ("k0", C.c_double), ("hp", C.c_double),
        ("gamma_dot_0", C.c_double), ("m_rate", C.c_double), ("G0", C.c_double), ("g0", C.c_double),
        ("gs_0", C.c_double), ("gamma_dot_s", C.c_double), ("w_sat", C.c_double),
        ("C11", C.c_double), ("C12", C.c_double), ("C44", C.c_double),
        ("tol_M", C.c_double), ("tol_hardening", C.c_double),
        ("max_itr_M", C.c_int32), ("max_itr_stag", C.c_int32),
        ("mu", C.c_double), ("w_max", C.c_double), ("p1", C.c_double), ("p2", C.c_double), ("Yin", C.c_double),
    ]


class Mesh(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int64), ("X", C.c_void_p), ("n_elem", C.c_int64), ("elem_ptr", C.c_void_p),
        ("elem_conn", C.c_void_p), ("elem_mat", C.c_void_p), ("n_coel", C.c_int64),
        ("coel_ptr", C.c_void_p), ("coel_conn", C.c_void_p),
    ]


class Cohesive(C.Structure):
    """pgf_cohesive: exponential traction-separation law (sig_c, chi_c, beta, alpha)."""

    _fields_ = [("sig_c", C.c_double), ("chi_c", C.c_double), ("beta", C.c_double), ("alpha", C.c_double)]


class NodalInput(C.Structure):
    _fields_ = [("r", C.c_void_p), ("d_r", C.c_void_p), ("defl", C.c_void_p), ("defl_d", C.c_void_p)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PGFError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P, I64, I32, D = C.c_void_p, C.c_int64, C.c_int32, C.c_double
        sig = {
            "pgf_generate_dof_ids": [P, P, P],
            "pgf_psparse_apai": [P, P, I64, P, P, P],
            "pgf_mesh_box": [I32, I32, D, C.c_uint64, P, P, P, P, P, P],
            "pgf_create": [P, P, I64, P, P, I32, P, P],
            "pgf_destroy": [P],
            "pgf_set_stream": [P, P],
            "pgf_set_profiling": [P, C.c_int],
            "pgf_last_pass_ms": [P, P, P],
            "pgf_synchronize": [P],
            "pgf_get_sizes": [P, P, P, P, P],
            "pgf_residual": [P, P, D, C.c_int, P],
            "pgf_tangent": [P, P, D, P],
            "pgf_prescribed_load": [P, P, D, P],
            "pgf_residual_tangent": [P, P, D, P, P],
            "pgf_residual_tangent_static": [P, P, D, P, P],
            "pgf_stage_doubles": [P, P],
            "pgf_subdivision_param": [P, D, P],
            "pgf_cp_start_stats": [P, P, P],
            "pgf_output_variables": [P, I32, P, P, P, P],
            "pgf_commit_step": [P],
            "pgf_reset_step": [P],
            "pgf_state_ncomp": [P, I32, P],
            "pgf_get_state": [P, I32, P],
            "pgf_set_state": [P, I32, P],
            "pgf_set_orientations": [P, I32, P, P],
            "pgf_set_time_step": [P, D],
            "pgf_set_warm_start": [P, P],
            "pgf_get_qp_order": [P, P],
            "pgf_state_device_ptr": [P, I32, P],
            "pgf_set_cohesive_props": [P, I32, P, P],
            "pgf_get_cohesive_sizes": [P, P, P],
            "pgf_mlp_create": [I32, P, C.c_uint64, P],
            "pgf_mlp_destroy": [P],
            "pgf_mlp_param_count": [P, P],
            "pgf_mlp_get_params": [P, P],
            "pgf_mlp_set_params": [P, P],
            "pgf_mlp_get_grads": [P, P],
            "pgf_mlp_forward": [P, P, I64, P],
            "pgf_mlp_infer": [P, P, I64, P],
            "pgf_mlp_backward": [P, P, P, I64, P],
            "pgf_mlp_adam_step": [P, C.c_float, C.c_float, C.c_float, C.c_float],
            "pgf_cp_log_samples": [P, P, P],
            "pgf_solve": [P, P, P, P, D, I32, P, P],
            "pgf_set_graph": [P, C.c_int],
        "pgf_set_assembly": [P, C.c_int32],
            "pgf_autotune": [P, P, D, P],
            "pgf_get_tuning": [P, P, P, P],
            "pgf_set_tuning": [P, I64, I32, I32],
            "pgf_set_surrogate": [P, P, P, P, P],
            "pgf_set_three_field": [P, I32],
            "pgf_tf_initial": [P, P, P],
            "pgf_tf_update": [P, P],
            "pgf_tf_accept": [P, D],
            "pgf_tf_get": [P, I32, P],
            "pgf_tf_set": [P, I32, P],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        _lib = L
    return _lib


def exported_symbols():
    """Names the C ABI exports (all functions declared in include/pgfem3d_amd.h)."""
    return [
        "pgf_generate_dof_ids", "pgf_psparse_apai", "pgf_mesh_box", "pgf_create", "pgf_destroy",
        "pgf_set_stream", "pgf_set_profiling", "pgf_last_pass_ms", "pgf_synchronize", "pgf_get_sizes", "pgf_residual",
        "pgf_tangent", "pgf_prescribed_load", "pgf_residual_tangent", "pgf_commit_step", "pgf_reset_step", "pgf_state_ncomp",
        "pgf_get_state", "pgf_set_state", "pgf_set_orientations", "pgf_set_time_step",
        "pgf_set_warm_start", "pgf_get_qp_order", "pgf_state_device_ptr", "pgf_set_cohesive_props",
        "pgf_get_cohesive_sizes", "pgf_mlp_create", "pgf_mlp_destroy", "pgf_mlp_param_count",
        "pgf_mlp_get_params", "pgf_mlp_set_params", "pgf_mlp_get_grads", "pgf_mlp_forward", "pgf_mlp_backward",
        "pgf_mlp_adam_step", "pgf_cp_log_samples", "pgf_set_surrogate", "pgf_solve", "pgf_set_graph", "pgf_set_assembly",
        "pgf_set_three_field", "pgf_tf_initial", "pgf_tf_update", "pgf_tf_accept", "pgf_tf_get", "pgf_tf_set",
        "pgf_autotune", "pgf_residual_tangent_static", "pgf_stage_doubles",
        "pgf_subdivision_param", "pgf_cp_start_stats", "pgf_output_variables", "pgf_mlp_infer",
        "pgf_get_tuning", "pgf_set_tuning",
    ]


def _check(rc: int, what: str) -> int:
    if rc < 0:
        raise PGFError(f"{what} failed: {ERRORS.get(rc, rc)}")
    return rc


def _ptr(a) -> int:
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return a.data_ptr()  # torch tensor


# ----------------------------------------------------------------------------- meshes
@dataclass
class MeshArrays:
    X: np.ndarray          # [n_nodes, 3]
    elem_ptr: np.ndarray   # [n_elem+1] int64
    elem_conn: np.ndarray  # int64
    elem_mat: np.ndarray | None = None
    coel_ptr: np.ndarray | None = None
    coel_conn: np.ndarray | None = None

    @property
    def n_nodes(self) -> int:
        return self.X.shape[0]

    @property
    def n_elem(self) -> int:
        return self.elem_ptr.shape[0] - 1

    @property
    def n_coel(self) -> int:
        return 0 if self.coel_ptr is None else self.coel_ptr.shape[0] - 1

    def struct(self) -> Mesh:
        m = Mesh()
        m.n_nodes = self.n_nodes
        m.X = _ptr(self.X)
        m.n_elem = self.n_elem
        m.elem_ptr = _ptr(self.elem_ptr)
        m.elem_conn = _ptr(self.elem_conn)
        m.elem_mat = _ptr(self.elem_mat) if self.elem_mat is not None else None
        if self.coel_ptr is not None:
            m.n_coel = self.coel_ptr.shape[0] - 1
            m.coel_ptr = _ptr(self.coel_ptr)
            m.coel_conn = _ptr(self.coel_conn)
        else:
            m.n_coel = 0
        return m


def mesh_box(n: int, etype: int, jitter: float = 0.1, seed: int = 0) -> MeshArrays:
    """Unit cube of n^3 cells (Kuhn 6-tet split for TET4/TET10, or HEX8)."""
    L = lib()
    nn, ne, cl = C.c_int64(), C.c_int64(), C.c_int64()
    _check(L.pgf_mesh_box(n, etype, jitter, seed, C.byref(nn), C.byref(ne), C.byref(cl), None, None, None),
           "pgf_mesh_box")
    X = np.zeros((nn.value, 3))
    ep = np.zeros(ne.value + 1, dtype=np.int64)
    ec = np.zeros(cl.value, dtype=np.int64)
    _check(L.pgf_mesh_box(n, etype, jitter, seed, C.byref(nn), C.byref(ne), C.byref(cl), _ptr(X), _ptr(ep),
                          _ptr(ec)), "pgf_mesh_box")
    return MeshArrays(X, ep, ec)


def generate_dof_ids(mesh: MeshArrays, supports: np.ndarray):
    """Support codes [n_nodes,3] (0 free, 1 fixed, -k prescribed) -> (DOF ids, ndof)."""
    ids = np.ascontiguousarray(supports, dtype=np.int64).copy()
    nd = C.c_int64()
    m = mesh.struct()
    _check(lib().pgf_generate_dof_ids(C.byref(m), _ptr(ids), C.byref(nd)), "pgf_generate_dof_ids")
    return ids, nd.value


def psparse_apai(mesh: MeshArrays, ids: np.ndarray, ndof: int):
    Ap = np.zeros(ndof + 1, dtype=np.int32)
    nnz = C.c_int64()
    m = mesh.struct()
    _check(lib().pgf_psparse_apai(C.byref(m), _ptr(ids), ndof, _ptr(Ap), None, C.byref(nnz)), "pgf_psparse_apai")
    Ai = np.zeros(nnz.value, dtype=np.int32)
    _check(lib().pgf_psparse_apai(C.byref(m), _ptr(ids), ndof, _ptr(Ap), _ptr(Ai), C.byref(nnz)),
           "pgf_psparse_apai")
    return Ap, Ai


def box_supports(mesh: MeshArrays, tol: float = 1e-12) -> np.ndarray:
    """Bottom face z=0 fixed in x, y, z; top face z=1 has u_z prescribed by value 1 (code -1)."""
    sup = np.zeros((mesh.n_nodes, 3), dtype=np.int64)
    z = mesh.X[:, 2]
    sup[z < tol] = 1
    sup[z > 1.0 - tol, 2] = -1
    return sup


# ------------------------------------------------------------------------- materials
def neo_hookean(E: float, nu: float, vol_pot: int = 2) -> Material:
    """Compressible neo-Hookean: Mooney-Rivlin with m01 = 0, m10 = G/2 (src/homogen.cc:184-186)."""
    G = E / (2.0 * (1.0 + nu))
    return Material(model=MODEL_HYPER_ELASTICITY, devPotFlag=1, volPotFlag=vol_pot, E=E, nu=nu, G=G,
                    m10=0.5 * G, m01=0.0)


def j2_linear(E: float, nu: float, sigma_y: float, H: float, vol_pot: int = 2) -> Material:
    G = E / (2.0 * (1.0 + nu))
    return Material(model=MODEL_J2_PLASTICITY, devPotFlag=1, volPotFlag=vol_pot, E=E, nu=nu, G=G,
                    m10=0.5 * G, m01=0.0, k0=sigma_y, hp=H)


def j2_damage(E: float, nu: float, sigma_y: float, H: float, mu: float, p1: float, p2: float, Yin: float,
              w_max: float = 1.0, vol_pot: int = 2) -> Material:
    """J2 with the viscous Weibull continuum damage (CM_J2P_PARAM damage properties mu, w_max,
    P1, P2, Yin; src/cm_j2_plasticity.cc:865-932)."""
    m = j2_linear(E, nu, sigma_y, H, vol_pot)
    m.mu, m.w_max, m.p1, m.p2, m.Yin = mu, w_max, p1, p2, Yin
    return m


# ---------------------------------------------------------------------------- handle
class Handle:
    """Device-side element hot path for one mesh (pgf_create ... pgf_destroy)."""

    def __init__(self, mesh: MeshArrays, ids: np.ndarray, ndof: int, Ap: np.ndarray, Ai: np.ndarray,
                 materials):
        self._lib = lib()
        mats = (Material * len(materials))(*materials)
        self._mats = mats
        self.materials = list(materials)
        h = C.c_void_p()
        m = mesh.struct()
        _check(self._lib.pgf_create(C.byref(m), _ptr(ids), ndof, _ptr(Ap), _ptr(Ai), len(materials), mats,
                                    C.byref(h)), "pgf_create")
        self.h = h
        nq, nz, nd, nc = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        _check(self._lib.pgf_get_sizes(h, C.byref(nq), C.byref(nz), C.byref(nd), C.byref(nc)), "pgf_get_sizes")
        self.n_qp, self.nnz, self.ndof, self.n_colors = nq.value, nz.value, nd.value, nc.value
        ncq, ncc = C.c_int64(), C.c_int32()
        _check(self._lib.pgf_get_cohesive_sizes(h, C.byref(ncq), C.byref(ncc)), "pgf_get_cohesive_sizes")
        self.n_cqp, self.n_ccolors = ncq.value, ncc.value
        self.n_prescribed = int(max(0, -int(np.min(ids)))) if ids.size else 0
        self.n_elem = mesh.n_elem
        self.three_field = False

    def close(self):
        if self.h:
            self._lib.pgf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _check_vec(t, n, what, exact=True):
        if t is None:
            return
        import torch

        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise PGFError(f"{what} must be a contiguous float64 GPU tensor")
        if (t.numel() != n) if exact else (t.numel() < n):
            raise PGFError(f"{what} has {t.numel()} entries, expected {n}")

    def _input(self, r, d_r=None, defl=None, defl_d=None) -> NodalInput:
        if r is None:
            raise PGFError("r is required")
        self._check_vec(r, self.ndof, "r")
        self._check_vec(d_r, self.ndof, "d_r")
        self._check_vec(defl, self.n_prescribed, "defl", exact=False)
        self._check_vec(defl_d, self.n_prescribed, "defl_d", exact=False)
        if self.n_prescribed and defl is None and defl_d is None:
            pass  # prescribed values default to zero
        return NodalInput(_ptr(r), _ptr(d_r), _ptr(defl), _ptr(defl_d))

    def set_profiling(self, on: bool):
        _check(self._lib.pgf_set_profiling(self.h, int(on)), "pgf_set_profiling")

    def last_pass_ms(self):
        ms, nl = C.c_double(), C.c_int32()
        _check(self._lib.pgf_last_pass_ms(self.h, C.byref(ms), C.byref(nl)), "pgf_last_pass_ms")
        return ms.value, nl.value

    def set_graph(self, on: bool):
        _check(self._lib.pgf_set_graph(self.h, int(on)), "pgf_set_graph")

    def set_assembly(self, mode: int):
        """ASSEMBLY_GATHER (default: staged element rows, one gather per node) or
        ASSEMBLY_COLOURED (colour-by-colour adds into the CSR values, no staging memory)."""
        _check(self._lib.pgf_set_assembly(self.h, int(mode)), "pgf_set_assembly")

    def autotune(self, r, d_r=None, defl=None, defl_d=None, dt=1.0) -> int:
        inp = self._input(r, d_r, defl, defl_d)
        c = C.c_int64()
        _check(self._lib.pgf_autotune(self.h, C.byref(inp), dt, C.byref(c)), "pgf_autotune")
        return c.value

    def tuning(self) -> dict:
        """The launch configuration in use: CP points per wavefront, gather nodes per workgroup
        and gather blocks in flight."""
        c, w, b = C.c_int64(), C.c_int32(), C.c_int32()
        _check(self._lib.pgf_get_tuning(self.h, C.byref(c), C.byref(w), C.byref(b)), "pgf_get_tuning")
        return {"cp_chunk": c.value, "gather_waves": w.value, "gather_batch": b.value}

    def set_tuning(self, cp_chunk=0, gather_waves=0, gather_batch=0):
        _check(self._lib.pgf_set_tuning(self.h, int(cp_chunk), int(gather_waves), int(gather_batch)),
               "pgf_set_tuning")

    def set_stream(self, stream_ptr: int | None):
        _check(self._lib.pgf_set_stream(self.h, stream_ptr), "pgf_set_stream")

    def residual(self, f, r, d_r=None, defl=None, defl_d=None, dt=1.0, run_integration=True) -> int:
        inp = self._input(r, d_r, defl, defl_d)
        self._check_vec(f, self.ndof, "f")
        return _check(self._lib.pgf_residual(self.h, C.byref(inp), dt, int(run_integration), _ptr(f)),
                      "pgf_residual")

    def tangent(self, vals, r, d_r=None, defl=None, defl_d=None, dt=1.0) -> int:
        inp = self._input(r, d_r, defl, defl_d)
        self._check_vec(vals, self.nnz, "vals")
        return _check(self._lib.pgf_tangent(self.h, C.byref(inp), dt, _ptr(vals)), "pgf_tangent")

    def prescribed_load(self, f_defl, r, defl_d, d_r=None, defl=None, dt=1.0) -> int:
        """f_defl = K(r + d_r, defl) . (defl_d on the prescribed DOFs), on the free rows."""
        if defl_d is None:
            raise PGFError("defl_d is required")
        inp = self._input(r, d_r, defl, defl_d)
        self._check_vec(f_defl, self.ndof, "f_defl")
        return _check(self._lib.pgf_prescribed_load(self.h, C.byref(inp), dt, _ptr(f_defl)), "pgf_prescribed_load")

    def residual_tangent(self, f, vals, r, d_r=None, defl=None, defl_d=None, dt=1.0) -> int:
        inp = self._input(r, d_r, defl, defl_d)
        self._check_vec(f, self.ndof, "f")
        self._check_vec(vals, self.nnz, "vals")
        return _check(self._lib.pgf_residual_tangent(self.h, C.byref(inp), dt, _ptr(f), _ptr(vals)),
                      "pgf_residual_tangent")

    def residual_tangent_static(self, f, vals, r, d_r=None, defl=None, defl_d=None, dt=1.0) -> int:
        """residual_tangent with the staging blocks in the library's fixed device array
        (no device allocation); PGF_ESCRATCH when the mesh's staging does not fit it."""
        inp = self._input(r, d_r, defl, defl_d)
        self._check_vec(f, self.ndof, "f")
        self._check_vec(vals, self.nnz, "vals")
        return _check(self._lib.pgf_residual_tangent_static(self.h, C.byref(inp), dt, _ptr(f), _ptr(vals)),
                      "pgf_residual_tangent_static")

    def output_variables(self, level=1):
        """(S, E per QP [n_qp, 6]; Cauchy stress and Almansi strain element averages
        [n_elem, 6]) at the time level (0 = n+1, 1 = n, 2 = n-1), components XX YY ZZ YZ XZ XY."""
        S = np.zeros((self.n_qp, 6))
        E = np.zeros((self.n_qp, 6))
        sig = np.zeros((self.n_elem, 6))
        eps = np.zeros((self.n_elem, 6))
        _check(self._lib.pgf_output_variables(self.h, int(level), _ptr(S), _ptr(E), _ptr(sig), _ptr(eps)),
               "pgf_output_variables")
        return S, E, sig, eps

    def cp_start_stats(self):
        """(QPs converged from the learned start, QPs that fell back to the plain start) of the
        last crystal-plasticity integration."""
        a, b = C.c_int64(), C.c_int64()
        _check(self._lib.pgf_cp_start_stats(self.h, C.byref(a), C.byref(b)), "pgf_cp_start_stats")
        return a.value, b.value

    def subdivision_param(self, dt=0.0) -> float:
        """Largest per-QP physics-based subdivision measure of the last integration
        (crystal plasticity: dt sum |gamma_dot| / 0.005)."""
        a = C.c_double()
        _check(self._lib.pgf_subdivision_param(self.h, float(dt), C.byref(a)), "pgf_subdivision_param")
        return a.value

    def stage_doubles(self) -> int:
        n = C.c_int64()
        _check(self._lib.pgf_stage_doubles(self.h, C.byref(n)), "pgf_stage_doubles")
        return n.value

    def commit_step(self):
        _check(self._lib.pgf_commit_step(self.h), "pgf_commit_step")

    # -- three-field (CM3F) analysis of hyperelastic meshes
    def set_three_field(self, on=True):
        """Switch the CM3F mode (per-element volume and pressure, condensed) on or off."""
        _check(self._lib.pgf_set_three_field(self.h, int(bool(on))), "pgf_set_three_field")
        self.three_field = bool(on)

    def tf_initial(self, r, defl=None, r_nm1=None, defl_nm1=None):
        """compute_3f_initial_conditions at (r, defl) for n and n+1, (r_nm1, defl_nm1) for n-1."""
        a = self._input(r, None, defl, None)
        b = self._input(r_nm1, None, defl_nm1, None) if r_nm1 is not None else None
        _check(self._lib.pgf_tf_initial(self.h, C.byref(a), C.byref(b) if b is not None else None),
               "pgf_tf_initial")

    def tf_update(self, du):
        """Element volume and pressure increments for the Newton increment du (device [ndof])."""
        self._check_vec(du, self.ndof, "du")
        _check(self._lib.pgf_tf_update(self.h, _ptr(du)), "pgf_tf_update")

    def tf_accept(self, gamma=1.0):
        _check(self._lib.pgf_tf_accept(self.h, float(gamma)), "pgf_tf_accept")

    def tf_get(self, field: int) -> np.ndarray:
        out = np.zeros(self.n_elem)
        _check(self._lib.pgf_tf_get(self.h, field, _ptr(out)), "pgf_tf_get")
        return out

    def tf_set(self, field: int, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.shape != (self.n_elem,):
            raise PGFError(f"tf_set: expected {self.n_elem} values")
        _check(self._lib.pgf_tf_set(self.h, field, _ptr(v)), "pgf_tf_set")

    def reset_step(self):
        _check(self._lib.pgf_reset_step(self.h), "pgf_reset_step")

    def get_state(self, field: int) -> np.ndarray:
        nc = C.c_int32()
        _check(self._lib.pgf_state_ncomp(self.h, field, C.byref(nc)), "pgf_state_ncomp")
        npts = self.n_cqp if field in (STATE_COHESIVE_NP1, STATE_COHESIVE_N) else self.n_qp
        out = np.zeros((nc.value, npts))
        _check(self._lib.pgf_get_state(self.h, field, _ptr(out)), "pgf_get_state")
        return out

    def set_state(self, field: int, values: np.ndarray):
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(self._lib.pgf_set_state(self.h, field, _ptr(v)), "pgf_set_state")

    def set_cohesive_props(self, props, coel_mat=None):
        """Interface properties (a list of Cohesive) and the property index of each interface
        element (None = all 0)."""
        arr = (Cohesive * len(props))(*props)
        cm = None if coel_mat is None else np.ascontiguousarray(coel_mat, dtype=np.int32)
        _check(self._lib.pgf_set_cohesive_props(self.h, len(props), arr, _ptr(cm)), "pgf_set_cohesive_props")
        self._coh_keep = (arr, cm)

    def solve(self, vals, b, x, tol=1e-12, maxit=5000):
        """Device BiCGStab: x <- K^-1 b (x holds the initial guess); returns (iters, relres)."""
        it, rr = C.c_int32(), C.c_double()
        rc = self._lib.pgf_solve(self.h, _ptr(vals), _ptr(b), _ptr(x), tol, maxit, C.byref(it), C.byref(rr))
        if rc < 0:
            _check(rc, "pgf_solve")
        return it.value, rr.value

    def log_samples(self, X, T):
        """Features [n_qp, 34] and targets [n_qp, 12] of the last integration into float32 GPU
        tensors (device QP order)."""
        self._check_f32(X, self.n_qp * 34, "X")
        self._check_f32(T, self.n_qp * 12, "T")
        _check(self._lib.pgf_cp_log_samples(self.h, _ptr(X), _ptr(T)), "pgf_cp_log_samples")

    def set_surrogate(self, mlp, shift=None, scale=None, out_scale=None):
        """Attach a trained MLP as the local-Newton warm start (None detaches)."""
        if mlp is None:
            _check(self._lib.pgf_set_surrogate(self.h, None, None, None, None), "pgf_set_surrogate")
            self._sur = None
            return
        nrm = [np.ascontiguousarray(v, dtype=np.float32) for v in (shift, scale, out_scale)]
        _check(self._lib.pgf_set_surrogate(self.h, mlp.m, _ptr(nrm[0]), _ptr(nrm[1]), _ptr(nrm[2])),
               "pgf_set_surrogate")
        self._sur = (mlp, nrm)

    @staticmethod
    def _check_f32(t, n, what):
        import torch

        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
            raise PGFError(f"{what} must be a contiguous float32 GPU tensor")
        if t.numel() != n:
            raise PGFError(f"{what} has {t.numel()} entries, expected {n}")

    def state_device_ptr(self, field: int) -> int:
        """Device address of a state field ([ncomp][n_qp], device QP order)."""
        p = C.c_void_p()
        _check(self._lib.pgf_state_device_ptr(self.h, field, C.byref(p)), "pgf_state_device_ptr")
        return p.value

    def qp_order(self) -> np.ndarray:
        """mesh QP index (element-major, QP-minor) of every device QP."""
        out = np.zeros(self.n_qp, dtype=np.int64)
        _check(self._lib.pgf_get_qp_order(self.h, _ptr(out)), "pgf_get_qp_order")
        return out

    def set_orientations(self, R: np.ndarray, elem_grain: np.ndarray | None):
        """Per-grain rotations (crystal -> sample, [n_grains,3,3]) and the grain of each element."""
        Rm = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 9)
        eg = None if elem_grain is None else np.ascontiguousarray(elem_grain, dtype=np.int32)
        _check(self._lib.pgf_set_orientations(self.h, Rm.shape[0], _ptr(Rm), _ptr(eg)), "pgf_set_orientations")
        self._grain_keep = (Rm, eg)

    def set_time_step(self, dt: float):
        _check(self._lib.pgf_set_time_step(self.h, float(dt)), "pgf_set_time_step")

    def set_warm_start(self, warm):
        """Start the next integration's local Newton from warm ([11][n_qp] float64 GPU tensor in
        device QP order: M row-major, lambda, g); None returns to the plain start."""
        if warm is not None:
            self._check_vec(warm, 11 * self.n_qp, "warm")
        self._warm = warm
        _check(self._lib.pgf_set_warm_start(self.h, _ptr(warm)), "pgf_set_warm_start")


# ------------------------------------------------------------------ crystal plasticity inputs
def fcc_crystal(C11=168.4e3, C12=121.4e3, C44=75.4e3, gamma_dot_0=1e-3, m=0.05, g0=16.0, gs=148.0, h0=180.0,
                w=0.0, gamma_dot_s=1.0, tol=1e-12, max_itr=60) -> Material:
    """FCC crystal plasticity (CP_PARAM, src/plasticity_model.cc:89-101): power-law flow and
    saturation hardening with the parameter names of the reference (G0 = h0, gs_0 = gs)."""
    return Material(model=MODEL_CRYSTAL_PLASTICITY, gamma_dot_0=gamma_dot_0, m_rate=m, G0=h0, g0=g0, gs_0=gs,
                    gamma_dot_s=gamma_dot_s, w_sat=w, C11=C11, C12=C12, C44=C44, tol_M=tol, max_itr_M=max_itr)


def random_quaternions(n: int, seed: int) -> np.ndarray:
    """n uniformly distributed rotations of SO(3) (Shoemake's method), as (w, x, y, z)."""
    u = np.random.default_rng(seed).random((n, 3))
    a, b, c = np.sqrt(1 - u[:, 0]), np.sqrt(u[:, 0]), 2 * np.pi * u[:, 1:]
    q = np.stack([b * np.cos(c[:, 1]), a * np.sin(c[:, 0]), a * np.cos(c[:, 0]), b * np.sin(c[:, 1])], axis=1)
    return q


def quat_to_rot(q: np.ndarray) -> np.ndarray:
    """Rotation matrices (crystal -> sample) of unit quaternions (w, x, y, z)."""
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                  2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                  2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], axis=-1)
    return R.reshape(q.shape[:-1] + (3, 3))


def voronoi_grains(mesh: MeshArrays, n_grains: int, seed: int):
    """Element -> grain id by the nearest of n_grains seeded uniform points in the unit cube
    (element centroid), and one uniform random orientation quaternion per grain."""
    rng = np.random.default_rng(seed)
    pts = rng.random((n_grains, 3))
    cen = np.zeros((mesh.n_elem, 3))
    nne = np.diff(mesh.elem_ptr)
    for k in np.unique(nne):
        sel = np.where(nne == k)[0]
        idx = mesh.elem_ptr[sel][:, None] + np.arange(k)[None, :]
        cen[sel] = mesh.X[mesh.elem_conn[idx]][:, :4 if k == 10 else k].mean(axis=1)
    grain = np.empty(mesh.n_elem, dtype=np.int32)
    for s in range(0, mesh.n_elem, 65536):
        d = ((cen[s:s + 65536, None, :] - pts[None, :, :]) ** 2).sum(-1)
        grain[s:s + 65536] = d.argmin(axis=1)
    return grain, random_quaternions(n_grains, seed + 1)


# ------------------------------------------------------------------ interface elements
_TET_FACES = np.array([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]])
_TET_OPP = np.array([3, 2, 1, 0])
_HEX_FACES = np.array([[0, 1, 2, 3], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]])


def split_grains(mesh: MeshArrays, grain: np.ndarray) -> MeshArrays:
    """Insert zero-thickness interface elements on every face shared by two elements of
    different grains: each node gets one copy per grain that uses it, and each such face
    becomes an interface element (- side: the copies of one grain in an order whose normal
    points into the other grain; + side: the other grain's copies in the same order).  Linear
    tets give 3 + 3 node interfaces, hexes 4 + 4; a mesh of one element type is required."""
    nne = np.diff(mesh.elem_ptr)
    if not np.all(nne == nne[0]) or nne[0] not in (4, 8):
        raise PGFError("split_grains needs a mesh of 4-node tets or 8-node hexes only")
    k = int(nne[0])
    conn = mesh.elem_conn.reshape(-1, k)
    grain = np.asarray(grain, dtype=np.int64)
    pairs = np.stack([conn.ravel(), np.repeat(grain, k)], axis=1)
    uniq, inv = np.unique(pairs, axis=0, return_inverse=True)
    new_X = np.ascontiguousarray(mesh.X[uniq[:, 0]])
    new_conn = inv.reshape(-1, k).astype(np.int64)
    copy = {(int(a), int(b)): i for i, (a, b) in enumerate(uniq)}
    faces = _TET_FACES if k == 4 else _HEX_FACES
    fn = conn[:, faces]                                   # [ne, nf, nv] original node ids
    key = np.sort(fn, axis=2).reshape(-1, faces.shape[1])
    order = np.lexsort(key.T[::-1])
    ks = key[order]
    same = np.all(ks[1:] == ks[:-1], axis=1)
    idx = np.where(same)[0]
    nf = faces.shape[0]
    coel = []
    for i in idx:
        fa, fb = order[i], order[i + 1]
        ea, eb = fa // nf, fb // nf
        if grain[ea] == grain[eb]:
            continue
        nodes = fn[ea, fa % nf].copy()
        x = mesh.X[nodes]
        nrm = np.cross(x[1] - x[0], x[2] - x[0])
        inner = mesh.X[conn[ea]].mean(axis=0)          # centroid of the - side element
        if np.dot(nrm, x.mean(axis=0) - inner) < 0:     # make the normal point out of ea
            nodes = nodes[::-1].copy()
        minus = [copy[(int(v), int(grain[ea]))] for v in nodes]
        plus = [copy[(int(v), int(grain[eb]))] for v in nodes]
        coel.append(minus + plus)
    coel = np.asarray(coel, dtype=np.int64).reshape(-1, 2 * faces.shape[1])
    cptr = np.arange(coel.shape[0] + 1, dtype=np.int64) * coel.shape[1]
    return MeshArrays(new_X, mesh.elem_ptr.copy(), np.ascontiguousarray(new_conn.ravel()),
                      None if mesh.elem_mat is None else mesh.elem_mat.copy(), cptr, np.ascontiguousarray(coel.ravel()))


# ------------------------------------------------------------------------------- MLP
class MLP:
    """The library's matrix-core MLP (pgf_mlp_*): BF16 weights/activations, FP32 accumulation,
    GELU hidden layers, trained by its own backward and Adam kernels."""

    def __init__(self, widths, seed=0):
        self._lib = lib()
        self.widths = list(widths)
        w = (C.c_int32 * len(widths))(*widths)
        m = C.c_void_p()
        _check(self._lib.pgf_mlp_create(len(widths) - 1, w, seed, C.byref(m)), "pgf_mlp_create")
        self.m = m
        n = C.c_int64()
        _check(self._lib.pgf_mlp_param_count(m, C.byref(n)), "pgf_mlp_param_count")
        self.n_param = n.value

    def close(self):
        if self.m:
            self._lib.pgf_mlp_destroy(self.m)
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.n_param, dtype=np.float32)
        _check(self._lib.pgf_mlp_get_params(self.m, _ptr(p)), "pgf_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_param:
            raise PGFError("parameter vector has the wrong length")
        _check(self._lib.pgf_mlp_set_params(self.m, _ptr(p)), "pgf_mlp_set_params")

    def get_grads(self):
        g = np.zeros(self.n_param, dtype=np.float32)
        _check(self._lib.pgf_mlp_get_grads(self.m, _ptr(g)), "pgf_mlp_get_grads")
        return g

    def forward(self, X, Y):
        """X [B, in], Y [B, out]: float32 GPU tensors."""
        _check(self._lib.pgf_mlp_forward(self.m, _ptr(X), X.shape[0], _ptr(Y)), "pgf_mlp_forward")

    def infer(self, X, Y):
        """The fused inference pass (pgf_mlp_infer): X [B, in], Y [B, out] float32 GPU tensors."""
        _check(self._lib.pgf_mlp_infer(self.m, _ptr(X), X.shape[0], _ptr(Y)), "pgf_mlp_infer")

    def backward(self, X, T):
        """Adds the gradient of the batch MSE to the gradient buffer; returns the loss."""
        loss = C.c_float()
        _check(self._lib.pgf_mlp_backward(self.m, _ptr(X), _ptr(T), X.shape[0], C.byref(loss)), "pgf_mlp_backward")
        return loss.value

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(self._lib.pgf_mlp_adam_step(self.m, lr, beta1, beta2, eps), "pgf_mlp_adam_step")
