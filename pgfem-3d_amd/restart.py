"""Checkpoint and resume of the device state in the reference's restart files.

save() writes, for every partition of a merged deck, the restart file the reference's rank
would write at that step (write_restart_mechanical -> write_restart_constitutive_model,
src/restart.cc:187-245: nodal u(n-1) and u(n), then per element its QPs' model records) and
the step's time-step info (write_time_step_info, :140-165).  load() reads them back into a
handle: the records set the state at n and n-1 and, as the models' read_restart do
(CP_PARAM::read_restart, src/plasticity_model.cc:913-942; HE_PARAM::read_restart,
src/plasticity_model_none.cc:314-327), the state at n+1 equal to n.

With the handle in the three-field (CM3F) mode, each file also holds the elements' volume and
pressure at n and n-1 (src/restart.cc:241-254, read back at :324-340 with n+1 = n).

Records exist for crystal plasticity (F, pF at n and n-1; tau and gamma_dot at n; g and lambda
at n and n-1) and for the hyperelastic model (F at n and n-1).  The reference's J2 record
(CM_J2P_PARAM::write_restart, src/cm_j2_plasticity.cc:784-808) holds the un-vendored GCM
integrator's own tensors, which this build's J2 state (Cp^-1, equivalent plastic strain) does
not map onto, so J2 meshes are refused.
"""
from __future__ import annotations

import numpy as np

import pgfem3d_amd as P
import ref_io as R

_CP_FIELDS = {"F_n": (P.STATE_F_N, None), "F_nm1": (P.STATE_F_NM1, None),
              "pF_n": (P.STATE_PLASTIC_N, None), "pF_nm1": (P.STATE_PLASTIC_NM1, None),
              "tau_n": (P.STATE_TAU_N, None), "gdot_n": (P.STATE_GDOT_N, None),
              "g_n": (P.STATE_SCALAR_N, 0), "lam_n": (P.STATE_SCALAR_N, 1),
              "g_nm1": (P.STATE_SCALAR_NM1, 0), "lam_nm1": (P.STATE_SCALAR_NM1, 1)}


def _kind(h):
    models = {m.model for m in h.materials}
    if P.MODEL_J2_PLASTICITY in models:
        raise NotImplementedError("J2 restart records hold the GCM integrator's tensors; not mapped")
    return "cp" if P.MODEL_CRYSTAL_PLASTICITY in models else "he"


def _qp_ptr(elem_ptr):
    nqp = np.array([{4: 1, 10: 4, 8: 8}[int(k)] for k in np.diff(elem_ptr)], dtype=np.int64)
    return np.concatenate([[0], np.cumsum(nqp)]), nqp


def records(h, elem_ptr):
    """Per mesh element, the restart records of its QPs from the handle's state at n."""
    kind = _kind(h)
    qp, _ = _qp_ptr(elem_ptr)
    if kind == "he":
        Fn, Fm = h.get_state(P.STATE_F_N), h.get_state(P.STATE_F_NM1)
        return [[R.HERecord(Fn[:, q].copy(), Fm[:, q].copy()) for q in range(qp[e], qp[e + 1])]
                for e in range(len(qp) - 1)]
    cache = {}
    for f, _ in _CP_FIELDS.values():
        if f not in cache:
            cache[f] = h.get_state(f)

    def val(name, q):
        f, c = _CP_FIELDS[name]
        a = cache[f]
        return float(a[c, q]) if c is not None else a[:, q].copy()

    return [[R.CPRecord(val("F_n", q), val("F_nm1", q), val("pF_n", q), val("pF_nm1", q), val("tau_n", q),
                        val("gdot_n", q), val("g_n", q), val("g_nm1", q), val("lam_n", q), val("lam_nm1", q))
             for q in range(qp[e], qp[e + 1])] for e in range(len(qp) - 1)]


def load_records(h, elem_ptr, elems):
    """Set the handle's state from per-element records: n and n-1 from the records, n+1 = n."""
    kind = _kind(h)
    qp, nqp = _qp_ptr(elem_ptr)
    if [len(r) for r in elems] != list(nqp):
        raise ValueError("records do not match the mesh's QPs")
    flat = [r for recs in elems for r in recs]
    if kind == "he":
        Fn = np.stack([r.F_n for r in flat], axis=1)
        Fm = np.stack([r.F_nm1 for r in flat], axis=1)
        for f, v in ((P.STATE_F_NP1, Fn), (P.STATE_F_N, Fn), (P.STATE_F_NM1, Fm)):
            h.set_state(f, v)
        return
    st = {n: np.stack([getattr(r, n) for r in flat], axis=1) if np.ndim(getattr(flat[0], n)) else
          np.array([getattr(r, n) for r in flat]) for n in _CP_FIELDS}
    sc_n = np.stack([st["g_n"], st["lam_n"]])
    sc_m = np.stack([st["g_nm1"], st["lam_nm1"]])
    for f, v in ((P.STATE_F_NP1, st["F_n"]), (P.STATE_F_N, st["F_n"]), (P.STATE_F_NM1, st["F_nm1"]),
                 (P.STATE_PLASTIC_NP1, st["pF_n"]), (P.STATE_PLASTIC_N, st["pF_n"]),
                 (P.STATE_PLASTIC_NM1, st["pF_nm1"]), (P.STATE_TAU, st["tau_n"]), (P.STATE_TAU_N, st["tau_n"]),
                 (P.STATE_GDOT, st["gdot_n"]), (P.STATE_GDOT_N, st["gdot_n"]), (P.STATE_SCALAR_NP1, sc_n),
                 (P.STATE_SCALAR_N, sc_n), (P.STATE_SCALAR_NM1, sc_m)):
        h.set_state(f, v)


def save(opath, ofname, step, deck, h, u_nm1, u_n, t3, norm, physics="Mechanical"):
    """Write the step's restart files of every partition of deck (u_nm1, u_n: nodal [n_nodes, 3]
    on the merged mesh) and its time-step info (t3 = t(n-1), t(n), t(n+1); norm = NORM)."""
    elems = records(h, deck.elem_ptr)
    tf = [h.tf_get(f) for f in (P.TF_V_N, P.TF_V_NM1, P.TF_P_N, P.TF_P_NM1)] if h.three_field else None
    for p, part in enumerate(deck.parts):
        nodes, els = deck.node_of[p], deck.elem_of[p]
        R.write_restart(R.restart_file(opath, ofname, p, step, physics), u_nm1[nodes], u_n[nodes],
                        [elems[e] for e in els], None if tf is None else [a[els] for a in tf])
    R.write_time_step_info(R.time_step_info_file(opath, step), t3, [norm], [t3[1]])


def load(opath, ofname, step, deck, h, physics="Mechanical"):
    """Read a step's restart files into the handle; returns (u_nm1, u_n, t3, NORM), the nodal
    fields on the merged mesh."""
    kind = _kind(h)
    n_nodes = deck.X.shape[0]
    u_nm1 = np.zeros((n_nodes, 3))
    u_n = np.zeros((n_nodes, 3))
    ne = len(deck.elem_ptr) - 1
    elems = [None] * ne
    tf = np.zeros((4, ne))
    _, nqp = _qp_ptr(deck.elem_ptr)
    for p, part in enumerate(deck.parts):
        els = deck.elem_of[p]
        out = R.read_restart(R.restart_file(opath, ofname, p, step, physics), part.nn, 3,
                             [kind] * part.ne, nqp[els], three_field=h.three_field)
        a, b, recs = out[:3]
        u_nm1[deck.node_of[p]] = a
        u_n[deck.node_of[p]] = b
        for k, e in enumerate(els):
            elems[e] = recs[k]
        if h.three_field:
            for i in range(4):
                tf[i, els] = out[3][i]
    load_records(h, deck.elem_ptr, elems)
    if h.three_field:  # n+1 = n, increments zero
        zero = np.zeros(ne)
        for f, v in ((P.TF_V_NP1, tf[0]), (P.TF_V_N, tf[0]), (P.TF_V_NM1, tf[1]), (P.TF_P_NP1, tf[2]),
                     (P.TF_P_N, tf[2]), (P.TF_P_NM1, tf[3]), (P.TF_DV, zero), (P.TF_DDV, zero), (P.TF_DP, zero),
                     (P.TF_DDP, zero)):
            h.tf_set(f, v)
    t3, norms, _ = R.read_time_step_info(R.time_step_info_file(opath, step))
    return u_nm1, u_n, t3, norms[0]
