"""Readers and writers of the reference's input and restart files for the element path.

Formats (one rank per partition file, as PGFem3D reads them):
  <base><p>.in       mesh, supports, prescribed values, elements, materials
                     (read_input_file, src/read_input_file.cc:280-410: read_nodes,
                     src/node.cc:78-140; read_Dirichlet_BCs / _values, src/in.cc:29-168;
                     build_elem / read_elem, src/element.cc:34-330; read_material,
                     src/in.cc:195-227; read_matgeom, :229-245; nodal and element loads,
                     :247-300)
  <base>.in.st       Newton settings, times, output steps, load steps and increments
                     (read_solver_file, src/read_input_file.cc:556-700; read_time_steps,
                     :517-538; solver_file_read_load, src/solver_file.cc:32-49)
  model_params.in    constitutive-model parameters; crystal plasticity as CP_PARAM::read_param
                     (src/plasticity_model.cc:945-1060)
  CO/co_<p>.in       crystal orientations per element and QP (plasticity_model_read_orientations,
                     src/plasticity_model.cc:1262-1330)
  <base>_<p>_<step>.res  restart: nodal u(n-1), u(n), then per element its QP count and each
                     QP's model record (write_restart_constitutive_model,
                     src/restart.cc:187-245; CP_PARAM::write_restart,
                     src/plasticity_model.cc:883-912; HE_PARAM::write_restart,
                     src/plasticity_model_none.cc:300-312), kept under
                     <opath>/restart/<physics>/STEP_<step>/ (:199-208)
  time_step_info_<step>.res  t(n-1) t(n) t(n+1), then NORM and t(n) per physics
                     (write_time_step_info / read_time_step_info, src/restart.cc:82-165)

Partitions are merged into one mesh: a node with a global number (Gnn >= 0) is the same node
in every partition that lists it; a node with Gnn = -1 belongs to its partition alone.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np


class _Tokens:
    """Whitespace tokens of a file, '#' comments removed, with typed pulls."""

    def __init__(self, path, keep_braces=True):
        with open(path) as fh:
            text = fh.read()
        toks = []
        for line in text.splitlines():
            line = line.split("#", 1)[0]
            if not keep_braces:
                line = line.replace("{", " ").replace("}", " ")
            toks.extend(line.split())
        self.t, self.i, self.path = toks, 0, path

    def more(self):
        return self.i < len(self.t)

    def word(self):
        if self.i >= len(self.t):
            raise ValueError(f"{self.path}: unexpected end of file")
        w = self.t[self.i]
        self.i += 1
        return w

    def int(self):
        return int(self.word())

    def float(self):
        return float(self.word())

    def ints(self, n):
        return [self.int() for _ in range(n)]

    def floats(self, n):
        return [self.float() for _ in range(n)]


NODES_OF_TYPE = {4: 4, 10: 10, 8: 8}


@dataclass
class Partition:
    """One rank's input file."""
    nn: int
    ne: int
    ndofn: int
    gnn: np.ndarray          # [nn] global number of a shared node, -1 for a partition node
    dom: np.ndarray          # [nn] owning partition
    X: np.ndarray            # [nn, 3]
    support: np.ndarray      # [nn, ndofn] 0 free, 1 fixed, -k prescribed value k
    defl: np.ndarray         # prescribed values (step-0 increment)
    toe: np.ndarray          # [ne] nodes per element
    conn: list               # [ne] arrays of local node ids
    elem_mat: np.ndarray     # [ne] material index
    materials: list          # dicts: Ex Ey Ez Gyz Gxz Gxy nyz nxz nxy ax ay az sig devPotFlag volPotFlag
    lin_maxit: int = 0
    lin_err: float = 0.0
    lim_zero: float = 0.0


def read_in(path) -> Partition:
    """One <base><p>.in file (non-legacy format)."""
    t = _Tokens(path)
    nn, ndofn, ne = t.ints(3)
    lin_maxit = t.int()
    lin_err, lim_zero = t.float(), t.float()
    nmat, n_conc, n_orient = t.ints(3)
    toe = np.array(t.ints(ne), dtype=np.int64)
    gnn = np.full(nn, -1, np.int64)
    dom = np.zeros(nn, np.int64)
    X = np.zeros((nn, 3))
    for _ in range(nn):
        g, d, i = t.ints(3)
        gnn[i], dom[i] = g, d
        X[i] = t.floats(3)
        t.ints(3)  # model_type, model_id, pr
    support = np.zeros((nn, ndofn), np.int64)
    for _ in range(t.int()):
        n = t.int()
        support[n] = t.ints(ndofn)
    defl = np.array(t.floats(t.int()))
    conn = [None] * ne
    for _ in range(ne):
        i = t.int()
        nne = NODES_OF_TYPE[int(toe[i])]
        conn[i] = np.array(t.ints(nne), np.int64)
        t.ints(3)  # face type, face id, pr
        t.ints(2 * (6 if nne == 8 else 4))  # boundary ids and types of the faces
    elem_mat = np.zeros(ne, np.int64)
    for _ in range(ne):
        i = t.int()
        mat = t.ints(2)
        t.ints(2)  # homogenisation ids
        elem_mat[i] = mat[0]
    keys = ["Ex", "Ey", "Ez", "Gyz", "Gxz", "Gxy", "nyz", "nxz", "nxy", "ax", "ay", "az", "sig"]
    materials = []
    for _ in range(nmat):
        m = dict(zip(keys, t.floats(13)))
        m["devPotFlag"], m["volPotFlag"] = t.ints(2)
        materials.append(m)
    t.floats(n_conc)
    t.floats(9 * n_orient)
    t.int()
    t.floats(2)  # SH a1 a2
    for _ in range(t.int()):  # nodal loads
        t.int()
        t.floats(ndofn)
    for _ in range(t.int()):  # surface loads: element, face, load
        t.ints(2)
        t.floats(ndofn)
    t.int()  # volume loads (count)
    return Partition(nn, ne, ndofn, gnn, dom, X, support, defl, toe, conn, elem_mat, materials, lin_maxit, lin_err,
                     lim_zero)


@dataclass
class Deck:
    """Partitions merged into one mesh."""
    parts: list
    X: np.ndarray              # [n_nodes, 3]
    elem_ptr: np.ndarray       # [n_elem+1]
    elem_conn: np.ndarray
    elem_mat: np.ndarray       # [n_elem]
    support: np.ndarray        # [n_nodes, 3]
    defl: np.ndarray           # prescribed values
    node_of: list = field(default_factory=list)   # per partition: local node -> merged node
    elem_of: list = field(default_factory=list)   # per partition: local element -> merged element


def merge(parts) -> Deck:
    key_to_node, X, support = {}, [], []
    node_of, elem_of = [], []
    defl = parts[0].defl
    for p, part in enumerate(parts):
        if part.defl.shape != defl.shape or not np.array_equal(part.defl, defl):
            raise ValueError(f"partition {p}: prescribed values differ from partition 0")
        m = np.empty(part.nn, np.int64)
        for i in range(part.nn):
            key = ("g", int(part.gnn[i])) if part.gnn[i] >= 0 else ("p", p, i)
            n = key_to_node.get(key)
            if n is None:
                n = len(X)
                key_to_node[key] = n
                X.append(part.X[i])
                support.append(part.support[i].copy())
            else:
                if not np.allclose(X[n], part.X[i], rtol=0, atol=1e-12):
                    raise ValueError(f"partition {p}: shared node {part.gnn[i]} has other coordinates")
                s = support[n]
                if np.any(part.support[i] != 0) and np.any(s != 0) and np.any(part.support[i] != s):
                    raise ValueError(f"partition {p}: shared node {part.gnn[i]} has other supports")
                if np.any(s == 0) and np.any(part.support[i] != 0):
                    support[n] = np.where(s != 0, s, part.support[i])
            m[i] = n
        node_of.append(m)
    ptr, conn, mat = [0], [], []
    for p, part in enumerate(parts):
        e_of = np.empty(part.ne, np.int64)
        for e in range(part.ne):
            e_of[e] = len(ptr) - 1
            conn.extend(node_of[p][part.conn[e]])
            ptr.append(len(conn))
            mat.append(part.elem_mat[e])
        elem_of.append(e_of)
    return Deck(parts, np.array(X), np.array(ptr, np.int64), np.array(conn, np.int64), np.array(mat, np.int32),
                np.array(support, np.int64), defl, node_of, elem_of)


def read_partitioned(directory, base, nparts) -> Deck:
    """<directory>/<base><p>.in for p = 0..nparts-1, merged."""
    return merge([read_in(os.path.join(directory, f"{base}{p}.in")) for p in range(nparts)])


@dataclass
class SolverFile:
    nor_min: float
    iter_max: int
    npres: int
    FNR: int
    times: np.ndarray          # [nt+1]
    print_steps: list
    load_steps: list           # steps whose increments the file holds
    increments: dict           # step -> prescribed-value increments

    def increment(self, step, step0):
        """Prescribed-value increment of a step; step 0's is the mesh file's values."""
        if step == 0:
            return np.asarray(step0, float)
        return self.increments.get(step, np.zeros(len(step0)))


def read_solver_file(path, n_prescribed) -> SolverFile:
    t = _Tokens(path)
    nor_min = t.float()
    iter_max, npres, FNR = t.ints(3)
    if FNR in (2, 3):
        t.floats(2)  # arc length
    if FNR == 4:
        raise NotImplementedError("multiphysics staggering (FNR 4)")
    nt = t.int()
    times = np.array(t.floats(nt + 1))
    print_steps = t.ints(t.int())
    load_steps = t.ints(t.int())
    increments = {s: np.array(t.floats(n_prescribed)) for s in load_steps if s > 0}
    return SolverFile(nor_min, iter_max, npres, FNR, times, print_steps, load_steps, increments)


def read_model_params(path) -> list:
    """Constitutive-model blocks: [{'mat': id, 'type': model type, ...}]; crystal plasticity
    (type 1) keys follow CP_PARAM::read_param."""
    t = _Tokens(path, keep_braces=False)
    out = []
    for _ in range(t.int()):
        mat, typ = t.ints(2)
        d = {"mat": mat, "type": typ}
        if typ == 1:
            d.update(zip(["gamma_dot_0", "m", "G0", "g0", "gs_0", "gamma_dot_s", "w"], t.floats(7)))
            d["unit_cell"] = t.int()
            d["ort_option"] = t.int()
            if d["ort_option"] == 0:
                d["ort_seed"] = t.int()
            elif d["ort_option"] == 2:
                d["ort_file"] = t.word()
            elif d["ort_option"] == 3:
                d["ort_angles"] = t.floats(3)
            if t.more() and t.int() == 1:
                (d["max_itr_stag"], d["max_itr_hardening"], d["max_itr_M"], d["max_subdivision"]) = t.ints(4)
                d["tol_hardening"], d["tol_M"], d["computer_zero"] = t.floats(3)
        else:
            raise NotImplementedError(f"model type {typ}")
        out.append(d)
    return out


def read_orientations(path):
    """(Euler angle type, {(element, ip): (phi, theta, psi)}) of one CO/co_<p>.in file."""
    angle_type, ang = 0, {}
    with open(path) as fh:
        for line in fh:
            if line.startswith("#"):
                if line[1:].startswith(" EulerAngleType:"):
                    angle_type = int(line.split(":", 1)[1].split()[0])
                continue
            w = line.split()
            if len(w) >= 5:
                ang[(int(w[0]), int(w[1]))] = tuple(float(v) for v in w[2:5])
    return angle_type, ang


@dataclass
class CPRecord:
    """One QP's crystal-plasticity restart record."""
    F_n: np.ndarray
    F_nm1: np.ndarray
    pF_n: np.ndarray
    pF_nm1: np.ndarray
    tau_n: np.ndarray
    gdot_n: np.ndarray
    g_n: float
    g_nm1: float
    lam_n: float
    lam_nm1: float


@dataclass
class HERecord:
    """One QP's hyperelastic restart record."""
    F_n: np.ndarray
    F_nm1: np.ndarray


def restart_file(opath, ofname, rank, step, physics="Mechanical"):
    """Where the reference keeps a rank's restart file of a step."""
    return os.path.join(opath, "restart", physics, f"STEP_{step:06d}", f"{ofname}_{rank}_{step}.res")


def time_step_info_file(opath, step):
    return os.path.join(opath, "restart", f"time_step_info_{step:06d}.res")


def read_restart(path, nn, ndofn, kinds, nqp=None, three_field=False):
    """(u_nm1 [nn, ndofn], u_n [nn, ndofn], [[record per QP] per element]); kinds[e] is "cp"
    or "he", the model of element e; nqp[e], when given, its QP count, which the file must
    hold (read_restart_constitutive_model checks it, src/restart.cc:300-309).  three_field
    (a CM3F run): the per-element volume and pressure records follow (src/restart.cc:324-340,
    one of each per element), returned fourth as (V_n, V_nm1, P_n, P_nm1) arrays [ne]."""
    t = _Tokens(path)
    u_nm1 = np.zeros((nn, ndofn))
    u_n = np.zeros((nn, ndofn))
    for a in range(nn):
        for b in range(ndofn):
            u_nm1[a, b], u_n[a, b] = t.float(), t.float()
    elems = []
    for e, kind in enumerate(kinds):
        n_ip = t.int()
        if nqp is not None and n_ip != nqp[e]:
            raise ValueError(f"{path}: element {e} has {n_ip} QPs in the file, {nqp[e]} in the mesh")
        recs = []
        for _ in range(n_ip):
            if kind == "cp":
                F_n, F_nm1, pF_n, pF_nm1 = (np.array(t.floats(9)) for _ in range(4))
                tau, gd = np.array(t.floats(12)), np.array(t.floats(12))
                recs.append(CPRecord(F_n, F_nm1, pF_n, pF_nm1, tau, gd, *t.floats(4)))
            elif kind == "he":
                recs.append(HERecord(np.array(t.floats(9)), np.array(t.floats(9))))
            else:
                raise ValueError(f"no restart record of model kind {kind!r}")
        elems.append(recs)
    if not three_field:
        return u_nm1, u_n, elems
    tf = np.zeros((4, len(kinds)))
    for e in range(len(kinds)):
        tf[0, e], tf[1, e] = t.float(), t.float()
        tf[2, e], tf[3, e] = t.float(), t.float()
    return u_nm1, u_n, elems, (tf[0], tf[1], tf[2], tf[3])


def read_restart_cp(path, nn, ndofn, ne):
    """read_restart with crystal plasticity on every element."""
    return read_restart(path, nn, ndofn, ["cp"] * ne)


def _row(v, trail):
    return " ".join(f"{x:.17e}" for x in v) + (" \n" if trail else "\n")


def write_restart(path, u_nm1, u_n, elems, tf=None):
    """The same layout, %.17e as the reference writes it (tensors one row each with no
    trailing blank, slip values with one): a file read by read_restart is written back
    byte for byte.  tf = (V_n, V_nm1, P_n, P_nm1) per element appends the CM3F records
    (write_restart_constitutive_model, src/restart.cc:241-254)."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as fh:
        for a in range(u_n.shape[0]):
            for b in range(u_n.shape[1]):
                fh.write(f"{u_nm1[a, b]:.17e} {u_n[a, b]:.17e} ")
            fh.write("\n")
        for recs in elems:
            fh.write(f"{len(recs)}\n")
            for r in recs:
                if isinstance(r, CPRecord):
                    for v in (r.F_n, r.F_nm1, r.pF_n, r.pF_nm1):
                        fh.write(_row(v, False))
                    fh.write(_row(r.tau_n, True))
                    fh.write(_row(r.gdot_n, True))
                    fh.write(f"{r.g_n:.17e} {r.g_nm1:.17e} {r.lam_n:.17e} {r.lam_nm1:.17e}\n")
                else:
                    fh.write(_row(r.F_n, False))
                    fh.write(_row(r.F_nm1, False))
        if tf is not None:
            for e in range(len(elems)):
                fh.write(f"{tf[0][e]:.17e} {tf[1][e]:.17e} \n{tf[2][e]:.17e} {tf[3][e]:.17e} \n")


write_restart_cp = write_restart


def write_time_step_info(path, t3, norms, tns):
    """t3 = (t(n-1), t(n), t(n+1)); norms and tns per physics."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as fh:
        fh.write(f"{t3[0]:.17e} {t3[1]:.17e} {t3[2]:.17e} ")
        for nv, tv in zip(norms, tns):
            fh.write(f"\n{nv:.17e} {tv:.17e}")
        fh.write("\n")


def read_time_step_info(path, nphysics=1):
    """(t3, norms, tns); a file without t(n) per physics (an older layout) gives tns None."""
    t = _Tokens(path)
    t3 = tuple(t.floats(3))
    norms, tns = [], []
    for _ in range(nphysics):
        norms.append(t.float())
        tns.append(t.float() if t.more() else None)
    return t3, norms, tns


def euler_rotation(phi, theta, psi):
    """R = Az(psi) Ay(theta) Ax(phi) with right-handed rotations about the axes (the product
    named at src/plasticity_model.cc:1252-1255; the matrices themselves are built in a library
    absent from the reference tree, so their sign convention is this one, stated)."""
    c, s = np.cos, np.sin
    Ax = np.array([[1, 0, 0], [0, c(phi), -s(phi)], [0, s(phi), c(phi)]])
    Ay = np.array([[c(theta), 0, s(theta)], [0, 1, 0], [-s(theta), 0, c(theta)]])
    Az = np.array([[c(psi), -s(psi), 0], [s(psi), c(psi), 0], [0, 0, 1]])
    return Az @ Ay @ Ax
