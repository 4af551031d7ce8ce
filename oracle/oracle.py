"""ctypes loader of the CPU oracle (oracle/_build/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.  The
product (pgfem-3d_amd/) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class OrcMaterial(C.Structure):
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


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        _lib = C.CDLL(LIB_PATH)
        for fn in ("orc_generate_local_dof_ids", "orc_psparse_apai", "orc_basis", "orc_he_update",
                   "orc_j2_update", "orc_assemble", "orc_nint"):
            getattr(_lib, fn).restype = C.c_int
        _lib.orc_j2_update.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_j2d_update.restype = C.c_int
        _lib.orc_j2d_update.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                        C.c_double] + [C.c_void_p] * 7
        _lib.orc_weibull.restype = None
        _lib.orc_weibull.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
        _lib.orc_j2d_Y0.restype = C.c_double
        _lib.orc_j2d_Y0.argtypes = [C.c_void_p, C.c_void_p]
        _lib.orc_assemble.argtypes = [C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]
        _lib.orc_psparse_apai.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64,
                                          C.c_void_p]
        _lib.orc_generate_local_dof_ids.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64,
                                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_basis.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_he_update.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def material(m) -> OrcMaterial:
    """Copy a pgf_material-like ctypes struct (same field order) into an OrcMaterial."""
    o = OrcMaterial()
    for name, _ in OrcMaterial._fields_:
        setattr(o, name, getattr(m, name))
    return o


def generate_dof_ids(X, elem_ptr, elem_conn, supports, coel_ptr=None, coel_conn=None):
    ids = np.ascontiguousarray(supports, dtype=np.int64).copy()
    nd = C.c_int64()
    nco = 0 if coel_ptr is None else coel_ptr.shape[0] - 1
    rc = lib().orc_generate_local_dof_ids(X.shape[0], elem_ptr.shape[0] - 1, _p(elem_ptr), _p(elem_conn), nco,
                                          _p(coel_ptr), _p(coel_conn), _p(ids), C.byref(nd))
    if rc:
        raise RuntimeError("orc_generate_local_dof_ids failed")
    return ids, nd.value


def psparse_apai(X, elem_ptr, elem_conn, ids, ndof, coel_ptr=None, coel_conn=None):
    L = lib()
    nco = 0 if coel_ptr is None else coel_ptr.shape[0] - 1
    Ap = np.zeros(ndof + 1, dtype=np.int32)
    nnz = C.c_int64()
    L.orc_psparse_apai(X.shape[0], elem_ptr.shape[0] - 1, _p(elem_ptr), _p(elem_conn), nco, _p(coel_ptr),
                       _p(coel_conn), _p(ids), ndof, _p(Ap), None, 0, C.byref(nnz))
    Ai = np.zeros(nnz.value, dtype=np.int32)
    rc = L.orc_psparse_apai(X.shape[0], elem_ptr.shape[0] - 1, _p(elem_ptr), _p(elem_conn), nco, _p(coel_ptr),
                            _p(coel_conn), _p(ids), ndof, _p(Ap), _p(Ai), nnz.value, C.byref(nnz))
    if rc:
        raise RuntimeError("orc_psparse_apai failed")
    return Ap, Ai


def basis(nne, X, ip):
    N = np.zeros(nne)
    dNdX = np.zeros((nne, 3))
    w = C.c_double()
    rc = lib().orc_basis(nne, _p(np.ascontiguousarray(X, dtype=np.float64)), ip, _p(N), _p(dNdX), C.byref(w))
    return N, dNdX, w.value, rc


def nint(nne):
    return lib().orc_nint(nne)


def he_update(mat, F, want_L=True):
    S = np.zeros(9)
    L = np.zeros(81) if want_L else None
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    om = material(mat)
    rc = lib().orc_he_update(C.byref(om), _p(F), _p(S), _p(L))
    return S.reshape(3, 3), (L.reshape(3, 3, 3, 3) if want_L else None), rc


def j2_update(mat, F, Cpinv_n, ep_n, want_L=True):
    S = np.zeros(9)
    L = np.zeros(81) if want_L else None
    c6 = np.zeros(6)
    ep1, gam = C.c_double(), C.c_double()
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    cn = np.ascontiguousarray(Cpinv_n, dtype=np.float64)
    om = material(mat)
    rc = lib().orc_j2_update(C.byref(om), _p(F), _p(cn), float(ep_n), _p(c6), C.byref(ep1), C.byref(gam), _p(S),
                             _p(L))
    return dict(S=S.reshape(3, 3), L=(L.reshape(3, 3, 3, 3) if want_L else None), Cpinv=c6, ep=ep1.value,
                gam=gam.value, rc=rc)


def j2d_update(mat, F, Cpinv_n, ep_n, w_n, X_n, dt, want_L=True):
    """J2 with continuum damage at one point: dict of S, L, S0, Cpinv, ep, gam, w, X, H,
    damaged, rc."""
    S = np.zeros(9)
    S0 = np.zeros(9)
    L = np.zeros(81) if want_L else None
    c6 = np.zeros(6)
    dam = np.zeros(4)
    ep1, gam = C.c_double(), C.c_double()
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    cn = np.ascontiguousarray(Cpinv_n, dtype=np.float64)
    om = material(mat)
    rc = lib().orc_j2d_update(C.byref(om), _p(F), _p(cn), float(ep_n), float(w_n), float(X_n), float(dt), _p(c6),
                              C.byref(ep1), C.byref(gam), _p(dam), _p(S), _p(L), _p(S0))
    return dict(S=S.reshape(3, 3), L=(L.reshape(3, 3, 3, 3) if want_L else None), S0=S0.reshape(3, 3), Cpinv=c6,
                ep=ep1.value, gam=gam.value, w=dam[0], X=dam[1], H=dam[2], damaged=int(dam[3]), rc=rc)


def weibull(mat, Y):
    """(G(Y), H(Y)) of the J2 damage's Weibull law."""
    g, h = C.c_double(), C.c_double()
    lib().orc_weibull(C.byref(material(mat)), float(Y), C.byref(g), C.byref(h))
    return g.value, h.value


def j2d_Y0(mat, F):
    return lib().orc_j2d_Y0(C.byref(material(mat)), _p(np.ascontiguousarray(F, dtype=np.float64).reshape(9)))


class OracleState:
    """SoA state arrays [comp][n_qp] in mesh QP order, as the oracle's assembly expects."""

    def __init__(self, n_qp, has_j2):
        self.n_qp = n_qp
        eye = np.array([1.0, 0, 0, 0, 1, 0, 0, 0, 1])
        self.F_np1 = np.ascontiguousarray(np.repeat(eye[:, None], n_qp, axis=1))
        c6 = np.array([1.0, 1, 1, 0, 0, 0])
        self.P_n = np.ascontiguousarray(np.repeat(c6[:, None], n_qp, axis=1))
        self.P_np1 = self.P_n.copy()
        # scalars: ep, gam, then the J2 damage variables w, X, H, damaged
        self.S_n = np.zeros((6, n_qp))
        self.S_np1 = np.zeros((6, n_qp))

    def commit(self):
        self.P_n[:] = self.P_np1
        self.S_n[:] = self.S_np1


def assemble(mesh, mats, ids, ndof, Ap, Ai, state: OracleState, r, d_r=None, defl=None, defl_d=None, dt=1.0,
             residual=True, integrate=True, tangent=True):
    """fd_residuals_MP + stiffmat_fd_MP restated on the CPU. Returns (f, vals, err)."""
    L = lib()
    om = (OrcMaterial * len(mats))(*[material(m) for m in mats])
    f = np.zeros(ndof)
    vals = np.zeros(Ap[-1])
    em = mesh.elem_mat if mesh.elem_mat is not None else None
    err = L.orc_assemble(mesh.n_nodes, _p(mesh.X), mesh.n_elem, _p(mesh.elem_ptr), _p(mesh.elem_conn), _p(em), om,
                         _p(ids), ndof, _p(Ap), _p(Ai), _p(r), _p(d_r), _p(defl), _p(defl_d), dt, int(residual),
                         int(integrate), int(tangent), _p(state.F_np1), _p(state.P_n), _p(state.P_np1),
                         _p(state.S_n), _p(state.S_np1), _p(f), _p(vals))
    return f, vals, err


# ----------------------------------------------------------------------------- crystal plasticity
def _cp_lib():
    L = lib()
    if not getattr(L, "_cp_ready", False):
        L.orc_cp_slip_systems.argtypes = [C.c_void_p]
        L.orc_cp_slip_systems.restype = None
        L.orc_cp_grain.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_cp_grain.restype = None
        L.orc_cp_update.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                    C.c_double, C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_cp_update.restype = C.c_int
        L.orc_assemble_cp.argtypes = [C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_double, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 14
        L.orc_assemble_cp.restype = C.c_int
        L._cp_ready = True
    return L


def cp_slip_systems():
    P0 = np.zeros((12, 3, 3))
    _cp_lib().orc_cp_slip_systems(_p(P0))
    return P0


def cp_grain(mat, R):
    P = np.zeros((12, 3, 3))
    Lg = np.zeros((3, 3, 3, 3))
    om = material(mat)
    _cp_lib().orc_cp_grain(C.byref(om), _p(np.ascontiguousarray(R, dtype=np.float64)), _p(P), _p(Lg))
    return P, Lg


def cp_update(mat, P, Lg, F, pFn, g_n, lam_n, dt, x0=None, solve=True, tangent=True, Fn=None):
    om = material(mat)
    pF1 = np.zeros(9)
    S = np.zeros(9)
    tau = np.zeros(12)
    gdot = np.zeros(12)
    dMdF = np.zeros(81) if tangent else None
    g1, l1, its = C.c_double(), C.c_double(), C.c_int()
    x0a = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
    rc = _cp_lib().orc_cp_update(C.byref(om), _p(np.ascontiguousarray(P)), _p(np.ascontiguousarray(Lg)),
                                 _p(np.ascontiguousarray(F, dtype=np.float64).reshape(9)),
                                 _p(None if Fn is None else np.ascontiguousarray(Fn, dtype=np.float64).reshape(9)),
                                 _p(np.ascontiguousarray(pFn, dtype=np.float64).reshape(9)), g_n, lam_n, dt, _p(x0a),
                                 int(solve), _p(pF1), C.byref(g1), C.byref(l1), _p(S), _p(tau), _p(gdot), _p(dMdF),
                                 C.byref(its))
    return dict(rc=rc, pF=pF1.reshape(3, 3), g=g1.value, lam=l1.value, S=S.reshape(3, 3), tau=tau, gdot=gdot,
                dMdF=None if dMdF is None else dMdF.reshape(9, 9), iters=its.value)


class CPState:
    """Crystal-plasticity state [comp][n_qp] in mesh QP order (oracle layout)."""

    def __init__(self, n_qp, g0):
        eye = np.array([1.0, 0, 0, 0, 1, 0, 0, 0, 1])
        self.n_qp = n_qp
        self.F_np1 = np.ascontiguousarray(np.repeat(eye[:, None], n_qp, axis=1))
        self.F_n = self.F_np1.copy()
        self.pF_n = self.F_np1.copy()
        self.pF_np1 = self.F_np1.copy()
        self.g_n = np.full(n_qp, float(g0))
        self.g_np1 = self.g_n.copy()
        self.lam_n = np.zeros(n_qp)
        self.lam_np1 = np.zeros(n_qp)
        self.tau = np.zeros((12, n_qp))
        self.gdot = np.zeros((12, n_qp))
        self.gdot_n = np.zeros((12, n_qp))
        self.iters = np.zeros(n_qp, dtype=np.int32)

    def commit(self):
        self.F_n[:] = self.F_np1
        self.pF_n[:] = self.pF_np1
        self.g_n[:] = self.g_np1
        self.lam_n[:] = self.lam_np1
        self.gdot_n[:] = self.gdot


def assemble_cp(mesh, mats, elem_grain, grain_P, grain_L, ids, ndof, Ap, Ai, st: CPState, r, d_r=None, defl=None,
                defl_d=None, dt=1.0, residual=True, integrate=True, tangent=True, warm=None):
    L = _cp_lib()
    om = (OrcMaterial * len(mats))(*[material(m) for m in mats])
    f = np.zeros(ndof)
    vals = np.zeros(Ap[-1])
    em = mesh.elem_mat if mesh.elem_mat is not None else None
    eg = np.ascontiguousarray(elem_grain, dtype=np.int32)
    gP = np.ascontiguousarray(grain_P, dtype=np.float64)
    gL = np.ascontiguousarray(grain_L, dtype=np.float64)
    err = L.orc_assemble_cp(mesh.n_nodes, _p(mesh.X), mesh.n_elem, _p(mesh.elem_ptr), _p(mesh.elem_conn), _p(em), om,
                            _p(eg), _p(gP), _p(gL), _p(ids), ndof, _p(Ap), _p(Ai), _p(r), _p(d_r), _p(defl),
                            _p(defl_d), dt, int(residual), int(integrate), int(tangent), _p(st.F_np1), _p(st.pF_n),
                            _p(st.pF_np1), _p(st.g_n), _p(st.g_np1), _p(st.lam_n), _p(st.lam_np1), _p(st.tau),
                            _p(st.gdot), _p(st.iters), _p(st.F_n), _p(warm), _p(f), _p(vals))
    return f, vals, err


# ----------------------------------------------------------------------------- cohesive
class OrcCohesive(C.Structure):
    _fields_ = [("sig_c", C.c_double), ("chi_c", C.c_double), ("beta", C.c_double), ("alpha", C.c_double)]


def _coh_lib():
    L = lib()
    if not getattr(L, "_coh_ready", False):
        L.orc_coh_traction.argtypes = [C.c_void_p] * 5
        L.orc_coh_traction.restype = C.c_int
        L.orc_coh_tangents.argtypes = [C.c_void_p] * 6
        L.orc_coh_tangents.restype = C.c_int
        L.orc_coh_element.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 6
        L.orc_coh_element.restype = C.c_int
        L.orc_assemble_coh.argtypes = ([C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_int64] + [C.c_void_p] * 10)
        L.orc_assemble_coh.restype = C.c_int
        L._coh_ready = True
    return L


def _coh(props):
    return OrcCohesive(props.sig_c, props.chi_c, props.beta, props.alpha)


def coh_traction(props, jump, n, vars_n=(0.0, 0.0)):
    t = np.zeros(3)
    v = np.ascontiguousarray(vars_n, dtype=np.float64)
    p = _coh(props)
    rc = _coh_lib().orc_coh_traction(C.byref(p), _p(np.ascontiguousarray(jump, dtype=np.float64)),
                                     _p(np.ascontiguousarray(n, dtype=np.float64)), _p(v), _p(t))
    return t, rc


def coh_tangents(props, jump, n, vars_n=(0.0, 0.0)):
    mt = np.zeros(9)
    gt = np.zeros(9)
    v = np.ascontiguousarray(vars_n, dtype=np.float64)
    p = _coh(props)
    _coh_lib().orc_coh_tangents(C.byref(p), _p(np.ascontiguousarray(jump, dtype=np.float64)),
                                _p(np.ascontiguousarray(n, dtype=np.float64)), _p(v), _p(mt), _p(gt))
    return mt.reshape(3, 3), gt.reshape(3, 3)


def coh_element(props, x0, u, vars_n=None):
    """One interface element with nn = len(x0)/2 nodes per side:
    (fe[6nn], K[6nn,6nn], vars_np1[nn,2], n_failed)."""
    x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(-1, 3)
    nn = x0.shape[0] // 2
    nd = 6 * nn
    fe = np.zeros(nd)
    K = np.zeros(nd * nd)
    v1 = np.zeros(2 * nn)
    vn = np.zeros(2 * nn) if vars_n is None else np.ascontiguousarray(vars_n, dtype=np.float64).reshape(2 * nn)
    p = _coh(props)
    err = _coh_lib().orc_coh_element(C.byref(p), nn, _p(x0), _p(np.ascontiguousarray(u, dtype=np.float64).reshape(nd)),
                                     _p(vn), _p(fe), _p(K), _p(v1))
    return fe, K.reshape(nd, nd), v1.reshape(nn, 2), err


class CohState:
    """Interface state [2][n_cqp] (X_eff_max, t_max) at n and n+1, cohesive-QP order (nn QPs
    per element, element by element)."""

    def __init__(self, n_cqp):
        self.vars_n = np.zeros((2, n_cqp))
        self.vars_np1 = np.zeros((2, n_cqp))

    def commit(self):
        self.vars_n[:] = self.vars_np1


def assemble_coh(mesh, props_list, coel_mat, ids, ndof, Ap, Ai, st: CohState, r, f, vals, d_r=None, defl=None,
                 defl_d=None):
    """Add the interface elements' residual into f and tangent into vals (in place)."""
    L = _coh_lib()
    pr = (OrcCohesive * len(props_list))(*[_coh(p) for p in props_list])
    cm = None if coel_mat is None else np.ascontiguousarray(coel_mat, dtype=np.int32)
    err = L.orc_assemble_coh(_p(mesh.X), mesh.n_coel, _p(mesh.coel_ptr), _p(mesh.coel_conn), _p(cm), pr, _p(ids),
                             ndof, _p(Ap), _p(Ai), _p(r), _p(d_r), _p(defl), _p(defl_d), _p(st.vars_n),
                             _p(st.vars_np1), _p(f), _p(vals))
    return err


def line_search_gamma(gam, phi, prev=None):
    """Next step length of the global Newton line search, LINE_S3_MP (src/LINE.cc:58-250).

    phi is ||R(u + gam du)||^2 / ||R(u)||^2, the ratio LS2/LS1 of the reference (LS scaled by
    LS1, so phi(0) = 1 and the slope of the Newton direction is -2).  The first cut minimises
    the quadratic through phi(0), phi'(0), phi(gam = 1) (LINE.cc:101-102); later cuts minimise
    the cubic through the last two trials (LINE.cc:104-116), capped at half the current step;
    a proposal below a tenth of the current step is replaced by a tenth (LINE.cc:120-124).
    prev = (gam, phi) of the trial before."""
    slope = -2.0
    if prev is None:
        lam = -slope / (2.0 * (phi - 1.0 - slope))
    else:
        g2, f2 = prev
        r1 = phi - 1.0 - gam * slope
        r2 = f2 - 1.0 - g2 * slope
        a = (r1 / (gam * gam) - r2 / (g2 * g2)) / (gam - g2)
        b = (-g2 * r1 / (gam * gam) + gam * r2 / (g2 * g2)) / (gam - g2)
        if a == 0.0:
            lam = -slope / (2.0 * b)
        else:
            disc = b * b - 3.0 * a * slope
            if disc < 0.0:
                lam = 0.5 * gam
            elif b <= 0.0:
                lam = (-b + np.sqrt(disc)) / (3.0 * a)
            else:
                lam = -slope / (b + np.sqrt(disc))
        lam = min(lam, 0.5 * gam)
    return lam if lam >= 0.1 * gam else 0.1 * gam


# ----------------------------------------------------------------------------- three-field (CM3F)
def _tf_lib():
    L = lib()
    if not getattr(L, "_tf_ready", False):
        vp, i64 = C.c_void_p, C.c_int64
        L.orc_tf_element.argtypes = [vp, C.c_int, vp, vp, C.c_double, C.c_double] + [vp] * 8
        L.orc_assemble_tf.argtypes = [i64, vp, vp, vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, C.c_int, vp, i64,
                                      vp, vp, vp, vp, vp, vp]
        L.orc_tf_update.argtypes = [i64] + [vp] * 15 + [i64, vp, vp]
        L.orc_tf_initial.argtypes = [i64] + [vp] * 8
        for fn in ("orc_tf_element", "orc_assemble_tf", "orc_tf_update", "orc_tf_initial"):
            getattr(L, fn).restype = C.c_int
        L._tf_ready = True
    return L


def tf_element(mat, nne, x, u, theta, P, Cpn=None, epn=None):
    """One CM3F element (material form): its blocks and the condensed K and f (J2: the state at
    n per QP, Cpn [nint, 6] and epn [nint]; None = the virgin state)."""
    nd = 3 * nne
    out = dict(Kuu=np.zeros((nd, nd)), Ru=np.zeros(nd), Kup=np.zeros(nd), sc=np.zeros(4), Kc=np.zeros((nd, nd)),
               fc=np.zeros(nd))
    om = material(mat)
    f64 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    rc = _tf_lib().orc_tf_element(C.byref(om), nne, _p(np.ascontiguousarray(x, dtype=np.float64)),
                                  _p(np.ascontiguousarray(u, dtype=np.float64)), float(theta), float(P),
                                  *[_p(out[k]) for k in ("Kuu", "Ru", "Kup", "sc", "Kc", "fc")], _p(f64(Cpn)),
                                  _p(f64(epn)))
    out["Rp"], out["Rt"], out["Ktt"], out["Ktp"] = out.pop("sc")
    out["rc"] = rc
    return out


def _mesh_args(mesh, mats):
    om = (OrcMaterial * len(mats))(*[material(m) for m in mats])
    em = mesh.elem_mat if mesh.elem_mat is not None else None
    return om, em


def assemble_tf(mesh, mats, ids, ndof, Ap, Ai, theta, P, r, d_r=None, defl=None, defl_d=None, integrate=True,
                F_np1=None, tangent=True, state: "OracleState" = None):
    """Condensed CM3F residual and tangent over the mesh (J2 materials take their state from
    state, and with integrate its n+1 arrays receive the new state). Returns (f, vals, err)."""
    om, em = _mesh_args(mesh, mats)
    f = np.zeros(ndof)
    vals = np.zeros(Ap[-1]) if tangent else None
    nq = F_np1.shape[1] if F_np1 is not None else (state.n_qp if state is not None else 0)
    st = state
    err = _tf_lib().orc_assemble_tf(mesh.n_elem, _p(mesh.X), _p(mesh.elem_ptr), _p(mesh.elem_conn), _p(em), om,
                                    _p(ids), ndof, _p(Ap), _p(Ai), _p(r), _p(d_r), _p(defl), _p(defl_d),
                                    _p(np.ascontiguousarray(theta, dtype=np.float64)),
                                    _p(np.ascontiguousarray(P, dtype=np.float64)), int(integrate), _p(F_np1), nq,
                                    _p(f), _p(vals), _p(None if st is None else st.P_n),
                                    _p(None if st is None else st.P_np1), _p(None if st is None else st.S_n),
                                    _p(None if st is None else st.S_np1))
    return f, vals, err


def tf_update(mesh, mats, ids, theta, P, du, r, d_r=None, defl=None, defl_d=None, state: "OracleState" = None):
    """Increments (ddV, ddP) of the elements' volume and pressure for the displacement increment du."""
    om, em = _mesh_args(mesh, mats)
    ddV, ddP = np.zeros(mesh.n_elem), np.zeros(mesh.n_elem)
    st = state
    err = _tf_lib().orc_tf_update(mesh.n_elem, _p(mesh.X), _p(mesh.elem_ptr), _p(mesh.elem_conn), _p(em), om,
                                  _p(ids), _p(r), _p(d_r), _p(defl), _p(defl_d),
                                  _p(np.ascontiguousarray(theta, dtype=np.float64)),
                                  _p(np.ascontiguousarray(P, dtype=np.float64)),
                                  _p(np.ascontiguousarray(du, dtype=np.float64)), _p(ddV), _p(ddP),
                                  0 if st is None else st.n_qp, _p(None if st is None else st.P_n),
                                  _p(None if st is None else st.S_n))
    return ddV, ddP, err


def tf_initial(mesh, mats, u_nodal):
    """Per element volume-averaged J and kappa U'(J) at nodal displacements [n_nodes, 3]."""
    om, em = _mesh_args(mesh, mats)
    V, Pr = np.zeros(mesh.n_elem), np.zeros(mesh.n_elem)
    err = _tf_lib().orc_tf_initial(mesh.n_elem, _p(mesh.X), _p(mesh.elem_ptr), _p(mesh.elem_conn), _p(em), om,
                                   _p(np.ascontiguousarray(u_nodal, dtype=np.float64)), _p(V), _p(Pr))
    return V, Pr, err


# ----------------------------------------------------------------------------- CM3F, crystal plasticity
def _cptf_lib():
    L = lib()
    if not getattr(L, "_cptf_ready", False):
        vp, i64, d = C.c_void_p, C.c_int64, C.c_double
        L.orc_cp_tf_element.argtypes = [vp, vp, vp, C.c_int, vp, vp, d, d, d] + [vp] * 18
        L.orc_assemble_cp_tf.argtypes = ([vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp, d, vp,
                                          vp, C.c_int, i64] + [vp] * 13)
        L.orc_cp_tf_update.argtypes = [vp, i64] + [vp] * 12 + [d, vp, vp, i64] + [vp] * 10
        for fn in ("orc_cp_tf_element", "orc_assemble_cp_tf", "orc_cp_tf_update"):
            getattr(L, fn).restype = C.c_int
        L._cptf_ready = True
    return L


def cp_tf_element(mat, gP, gL, nne, x, u, theta, P, dt, pFn, g_n, lam_n, Fn=None):
    """One CM3F crystal-plasticity element from the state at n ([nint, 3, 3] pF(n), F(n);
    [nint] g(n), lambda(n)), integrated at (u, theta, P): its n+1 state, blocks and condensed
    K and f."""
    nd = 3 * nne
    nint = len(g_n)
    eye = np.repeat(np.eye(3)[None], nint, axis=0)
    st = dict(Fb=np.zeros((nint, 9)), pF1=np.zeros((nint, 9)), g1=np.zeros(nint), lam1=np.zeros(nint),
              tau=np.zeros((nint, 12)), gdot=np.zeros((nint, 12)))
    out = dict(Kuu=np.zeros((nd, nd)), Ru=np.zeros(nd), Kup=np.zeros(nd), Kut=np.zeros(nd), Ktu=np.zeros(nd),
               sc=np.zeros(4), Kc=np.zeros((nd, nd)), fc=np.zeros(nd))
    om = material(mat)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    rc = _cptf_lib().orc_cp_tf_element(
        C.byref(om), _p(f64(gP)), _p(f64(gL)), nne, _p(f64(x)), _p(f64(u)), float(theta), float(P), float(dt),
        _p(f64(pFn).reshape(-1)), _p(f64(eye if Fn is None else Fn).reshape(-1)), _p(f64(g_n)), _p(f64(lam_n)),
        *[_p(st[k]) for k in ("Fb", "pF1", "g1", "lam1", "tau", "gdot")],
        *[_p(out[k]) for k in ("Kuu", "Ru", "Kup", "Kut", "Ktu", "sc", "Kc", "fc")])
    out["Rp"], out["Rt"], out["Ktt"], out["Ktp"] = out.pop("sc")
    out.update(st)
    out["rc"] = rc
    return out


def assemble_cp_tf(mesh, mats, elem_grain, grain_P, grain_L, ids, ndof, Ap, Ai, st: "CPState", theta, P, r,
                   d_r=None, defl=None, defl_d=None, dt=1.0, integrate=True, tangent=True):
    """Condensed CM3F residual and tangent of a crystal-plasticity mesh (state in st, [comp][n_qp]
    mesh QP order; with integrate the return map runs at (theta/J)^(1/3) F and st.F_np1 holds
    that F). Returns (f, vals, err)."""
    om, em = _mesh_args(mesh, mats)
    f = np.zeros(ndof)
    vals = np.zeros(Ap[-1]) if tangent else None
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    err = _cptf_lib().orc_assemble_cp_tf(
        _p(mesh.X), mesh.n_elem, _p(mesh.elem_ptr), _p(mesh.elem_conn), _p(em), om,
        _p(np.ascontiguousarray(elem_grain, dtype=np.int32)), _p(f64(grain_P)), _p(f64(grain_L)), _p(ids), ndof,
        _p(Ap), _p(Ai), _p(r), _p(d_r), _p(defl), _p(defl_d), float(dt), _p(f64(theta)), _p(f64(P)), int(integrate),
        st.n_qp, _p(st.F_np1), _p(st.F_n), _p(st.pF_n), _p(st.pF_np1), _p(st.g_n), _p(st.g_np1), _p(st.lam_n),
        _p(st.lam_np1), _p(st.tau), _p(st.gdot), _p(st.iters), _p(f), _p(vals))
    return f, vals, err


def cp_tf_update(mesh, mats, elem_grain, grain_P, grain_L, ids, st: "CPState", theta, P, du, r, d_r=None,
                 defl=None, defl_d=None, dt=1.0):
    """Increments (ddV, ddP) of the CM3F crystal-plasticity elements for the displacement
    increment du, at the stored state."""
    om, em = _mesh_args(mesh, mats)
    ddV, ddP = np.zeros(mesh.n_elem), np.zeros(mesh.n_elem)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    err = _cptf_lib().orc_cp_tf_update(
        _p(mesh.X), mesh.n_elem, _p(mesh.elem_ptr), _p(mesh.elem_conn), _p(em), om,
        _p(np.ascontiguousarray(elem_grain, dtype=np.int32)), _p(f64(grain_P)), _p(f64(grain_L)), _p(ids), _p(r),
        _p(d_r), _p(defl), _p(defl_d), float(dt), _p(f64(theta)), _p(f64(P)), st.n_qp, _p(st.F_np1), _p(st.pF_n),
        _p(st.pF_np1), _p(st.g_n), _p(st.g_np1), _p(st.lam_n), _p(st.lam_np1), _p(f64(du)), _p(ddV), _p(ddP))
    return ddV, ddP, err


# ----------------------------------------------------------------------------- output variables
OUT_ORDER = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]  # XX YY ZZ YZ XZ XY


def output_qp(mat, F, plastic=None, w=0.0, L_grain=None):
    """PK2 S and Cauchy sigma (3x3) of one QP from its stored state (orc_output_qp)."""
    L = lib()
    if not getattr(L, "_out_ready", False):
        L.orc_output_qp.restype = C.c_int
        L.orc_output_qp.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                    C.c_void_p]
        L._out_ready = True
    S = np.zeros(9)
    sig = np.zeros(9)
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    pl = None if plastic is None else np.ascontiguousarray(plastic, dtype=np.float64)
    Lg = None if L_grain is None else np.ascontiguousarray(L_grain, dtype=np.float64)
    rc = L.orc_output_qp(C.byref(material(mat)), _p(Lg), _p(F), _p(pl), float(w), _p(S), _p(sig))
    if rc:
        raise RuntimeError("orc_output_qp failed")
    return S.reshape(3, 3), sig.reshape(3, 3)


def output_variables(mesh, mats, F, plastic=None, w=None, elem_grain=None, grain_L=None):
    """constitutive_model_update_output_variables over a mesh from state arrays [comp][n_qp]
    (mesh QP order): per QP S and E [n_qp, 6], per element the detJ w averages of the Cauchy
    stress and of the Almansi strain with doubled shears [n_elem, 6], order XX YY ZZ YZ XZ XY."""
    nqp = F.shape[1]
    S_out = np.zeros((nqp, 6))
    E_out = np.zeros((nqp, 6))
    sig_el = np.zeros((mesh.n_elem, 6))
    eps_el = np.zeros((mesh.n_elem, 6))
    q = 0
    for e in range(mesh.n_elem):
        conn = mesh.elem_conn[mesh.elem_ptr[e]:mesh.elem_ptr[e + 1]]
        x = mesh.X[conn]
        mat = mats[mesh.elem_mat[e] if mesh.elem_mat is not None else 0]
        vol = 0.0
        for ip in range(nint(len(conn))):
            _, _, wq, _ = basis(len(conn), x, ip)
            Fq = F[:, q].reshape(3, 3)
            pl = None if plastic is None else plastic[:, q]
            Lg = None if grain_L is None else grain_L[elem_grain[e]]
            S, sig = output_qp(mat, Fq, pl, 0.0 if w is None else w[q], Lg)
            E = 0.5 * (Fq.T @ Fq - np.eye(3))
            ea = 0.5 * (np.eye(3) - np.linalg.inv(Fq @ Fq.T))
            for c, (i, j) in enumerate(OUT_ORDER):
                S_out[q, c] = S[i, j]
                E_out[q, c] = E[i, j]
                sig_el[e, c] += wq * sig[i, j]
                eps_el[e, c] += wq * (1.0 if c < 3 else 2.0) * ea[i, j]
            vol += wq
            q += 1
        sig_el[e] /= vol
        eps_el[e] /= vol
    return S_out, E_out, sig_el, eps_el
