"""Benchmark of the finite-strain element hot path on one MI355X.

Headline workload (value): J2 finite-strain plasticity with linear isotropic hardening on a
unit cube of 48^3 cells, each split into 6 quadratic 10-node tets (663,552 elements,
2,654,208 QPs, 4-point quadrature), E=70 GPa, nu=0.33, sigma_y=250 MPa, H=1 GPa, FP64.
Bottom face fixed, top face stretched; the displacement is a seeded 1.5% uniaxial stretch
(plastic at every QP).

Second workload (the "cp" object): FCC crystal plasticity on 55^3 cells x 6 linear tets
(998,250 elements and QPs), 64 Voronoi grains with uniformly random orientations, power-law
flow (gamma_dot0=1e-3/s, m=0.05), saturation hardening (g0=16, gs=148, h0=180 MPa), cubic
elasticity (168.4/121.4/75.4 GPa), strain rate 1e-3/s, dt = 1 s, plain local-Newton start.
The load history is driven to 0.5% (five committed steps), then the sixth step is timed.
The learned warm start ("cp.surrogate"): the MLP (34 -> 4 x 128 GELU -> 12, BF16 on the
matrix cores) is trained by the library's own kernels on the QP states logged over the five
history steps (seeded 90/10 split), attached, and the sixth step is timed again; the
converged state is compared with the plain start's.

One step = one global Newton iteration's element work: deformation gradient, constitutive
update (return map with its local Newton) and state update at every QP, element residual
and tangent, and their assembly into the residual vector and the CSR values (Psparse_ApAi
pattern), all resident in HBM.  The linear solve is not part of the step.
value = QPs / second of wall time per step.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("pgfem-3d_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

METRIC = "QP constitutive updates + tangent assembly /sec (FP64); surrogate vs plain speedup"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
CP_STEP = 1e-3  # strain per 1 s step at 1e-3/s


def algorithmic_bytes(n_qp, nnz, ndof, n_nodes, n_elem, nne):
    """Unique HBM traffic of one J2 step (bytes): state read (Cp^-1_n 6 + ep_n 1 doubles) and
    written (F 9 + Cp^-1 6 + ep, gam 2 doubles), each CSR value and residual entry written
    once, coordinates and displacements (6 doubles per node) and connectivity (4 B per
    element node) read once."""
    state = n_qp * 8 * (7 + 17)
    return state + 8 * nnz + 8 * ndof + 8 * 6 * n_nodes + 4 * nne * n_elem


def flops_per_qp(nne):
    """FP64 operations per QP of the fused J2 kernel, counted from the source (FMA = 2):
    phase 1 ~ (9*nne*2 kinematics + 9*nne h_a + ~700 constitutive/tangent), phase 2 per
    element row: 18 MACs (the three strain entries of the row against c) + 9 (tau h_a) + 3
    (residual) + 9 (geometric term folded into the row block) + 9*nne (the block times h_w)."""
    ndofe = 3 * nne
    phase1 = 2 * (9 * nne * 2 + 9 * nne) + 700
    phase2 = ndofe * 2 * (18 + 9 + 3 + 9 + 9 * nne)
    return phase1 + phase2


def cp_flops_per_qp(mean_iters, nne=4):
    """FP64 operations per QP of the crystal-plasticity pass, counted from the kernels (FMA = 2,
    divide, pow and cbrt as one each; the flow rule's |tau/g|^19 is eight multiplies of binary
    powering, counted as one, which the ~1e4 totals absorb):
      material-point evaluation (cp_point) ~1.4e3: eF, C_e, S = L:E, Mandel stress, 12 tau_a,
        12 power-law rates, R_M;
      Jacobian ~6.2e3: 12 x (d tau_a/dM 260 + rank-1 update of the 9x9 block 162 + borders
        36), lambda terms 400, Schur complement 162;
      9x9 LU 500 and two solves 324;
      => Newton iteration ~8.4e3 (one evaluation: a line-search trial's evaluation is the next
      iteration's point);
      start ~3.5e3 (up to 10 tau-ratio evaluations of ~350);
      tangent pass ~1.2e4 (evaluation, Jacobian, LU, nine bordered solves, d tau/dF);
      element rows: nine row-basis vectors per QP x ~650 (B, Y = L:B, Z, z = Z:D, W, Fr W),
        then 3*nne rows x 2 (50 + 9 nne) (S g_a, residual, the three-basis combination, the
        geometric term and the row of K)."""
    return 8.4e3 * mean_iters + 3.5e3 + 1.2e4 + 9 * 650 + 3 * nne * 2 * (50 + 9 * nne)


def measured_traffic(key):
    """HBM bytes per pass of the dominant kernel from the committed PMC summary (FETCH_SIZE and
    WRITE_SIZE, calibrated against a copy of known size), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as fh:
            return json.load(fh)[key]["elem_kernel_bytes_per_pass"]
    except (OSError, KeyError, ValueError):
        return None


def measured_flops(key):
    """FP64 FLOPs per pass from the committed counter summary (64 x (ADD + MUL + TRANS +
    2 FMA) F64 wave instructions per dispatch, summed over the pass's kernels), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "fp64_counters.json")) as fh:
            return json.load(fh)[key]["flops_per_pass"]
    except (OSError, KeyError, ValueError):
        return None


def _timed_passes(h, f, vals, rd, dd, steps, warmup, dt=1.0):
    import torch

    for _ in range(warmup):
        err = h.residual_tangent(f, vals, rd, defl=dd, dt=dt)
        assert err == 0, err
    torch.cuda.synchronize()
    pass_ms = []
    t1 = time.perf_counter()
    for _ in range(steps):
        err = h.residual_tangent(f, vals, rd, defl=dd, dt=dt)
        pass_ms.append(h.last_pass_ms()[0])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    assert err == 0, err
    return wall / steps, float(np.mean(pass_ms)), h.last_pass_ms()[1]


def host_info():
    """nproc (usable cores), the CPU model and the OpenMP thread count the all-core legs use."""
    import os

    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", nproc) or nproc)
    return {"nproc": nproc, "cpu_model": model, "omp_threads": min(threads, nproc)}


def _oracle_threads(n):
    import ctypes as C

    import oracle as O

    L = O.lib()
    L.orc_set_threads.restype = C.c_int
    L.orc_set_threads.argtypes = [C.c_int]
    return L.orc_set_threads(int(n))


def fp64_counter_fields(flops, kern_ms):
    """The counter-derived FP64 rate of a pass (executed FP64 instructions, profiles/
    fp64_counters.md) at the live pass time, beside the source-count rate."""
    if flops is None:
        return {"counter_flops_per_pass": None}
    t = flops / (kern_ms * 1e-3) / 1e12
    return {"counter_flops_per_pass": flops, "counter_tflops": t, "counter_frac": t / FP64_PEAK_TFLOPS,
            "counter_source": "profiles/fp64_counters.md"}


def fp64_entry(tflops, flops, kern_ms):
    return {"achieved_tflops": tflops, "peak_tflops": FP64_PEAK_TFLOPS, "frac": tflops / FP64_PEAK_TFLOPS,
            "basis": "source count (flops_per_qp)", **fp64_counter_fields(flops, kern_ms)}


def cpu_baseline(n_cells):
    """Oracle (CPU restatement of the reference path) on a bounded sample: the same element
    type and material on an n_cells^3 mesh; one fused residual + tangent pass, timed on one
    host thread and on all of them (OpenMP over element colours, same results bit for bit).
    value is the all-thread rate."""
    import oracle as O
    import pgfem3d_amd as P
    from problems import BoxProblem

    info = host_info()
    prob = BoxProblem(n_cells, P.PGF_TET10, P.j2_linear(70e3, 0.33, 250.0, 1000.0))
    r, defl = prob.displacement(0.015, 2e-4, seed=11)
    times = {}
    for th in (1, info["omp_threads"]):
        used = _oracle_threads(th)
        st = O.OracleState(prob.nqp, True)
        t0 = time.perf_counter()
        O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r, defl=defl)
        times[used] = time.perf_counter() - t0
    nt = max(times)
    return {"value": prob.nqp / times[nt], "unit": "QP updates/s", "cores": nt, "kind": "port",
            "single_thread": prob.nqp / times[1], **info,
            "sample": f"oracle orc_assemble, one fused residual+tangent pass, J2 P2 {n_cells}^3 cells "
                      f"({prob.mesh.n_elem} elements, {prob.nqp} QPs): {times[1]:.1f} s on 1 host thread, "
                      f"{times[nt]:.2f} s on {nt} (OpenMP over element colours)"}


def cpu_baseline_cp(n_cells, n_grains=64):
    """Oracle crystal plasticity on an n_cells^3 P1 sample with the same grains and load
    history as the GPU run: five committed steps untimed, the sixth timed on one host thread
    and on all of them (value: all threads)."""
    import oracle as O
    import pgfem3d_amd as P
    from problems import CPBoxProblem

    info = host_info()
    mat = P.fcc_crystal()
    prob = CPBoxProblem(n_cells, P.PGF_TET4, mat, n_grains=n_grains, seed=2)
    st = O.CPState(prob.nqp, mat.g0)
    _oracle_threads(info["omp_threads"])
    for s in range(1, 6):
        r, defl = prob.displacement(CP_STEP * s, 0.0, seed=s)
        _, _, err = prob.assemble(st, r, defl, tangent=False)
        assert err == 0
        st.commit()
    r, defl = prob.displacement(CP_STEP * 6, 0.0, seed=6)
    times = {}
    for th in (1, info["omp_threads"]):
        used = _oracle_threads(th)
        t0 = time.perf_counter()
        _, _, err = prob.assemble(st, r, defl)
        times[used] = time.perf_counter() - t0
        assert err == 0
    nt = max(times)
    return {"value": prob.nqp / times[nt], "unit": "QP updates/s", "cores": nt, "kind": "port",
            "single_thread": prob.nqp / times[1], "mean_local_iters": float(np.mean(st.iters)), **info,
            "sample": f"oracle orc_assemble_cp, one fused residual+tangent pass at 0.6% strain, CP P1 "
                      f"{n_cells}^3 cells ({prob.nqp} QPs, {n_grains} grains): {times[1]:.1f} s on 1 host "
                      f"thread, {times[nt]:.2f} s on {nt} (OpenMP over element colours)"}


def run_j2(args, dev):
    import torch

    import pgfem3d_amd as P
    from problems import BoxProblem

    t0 = time.perf_counter()
    prob = BoxProblem(args.n, P.PGF_TET10, P.j2_linear(70e3, 0.33, 250.0, 1000.0))
    t_setup = time.perf_counter() - t0
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    t_create = time.perf_counter() - t0 - t_setup
    print(f"[bench] J2: {prob.mesh.n_elem} P2 tets, {prob.mesh.n_nodes} nodes, ndof {prob.ndof}, nnz {h.nnz}, "
          f"QPs {h.n_qp}, colours {h.n_colors}; host set-up {t_setup:.1f}s, create {t_create:.1f}s",
          file=sys.stderr, flush=True)
    r, defl = prob.displacement(0.015, 2e-4, seed=11)
    rd = torch.from_numpy(r).to(dev)
    dd = torch.from_numpy(defl).to(dev)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=dev)
    vals = torch.empty(h.nnz, dtype=torch.float64, device=dev)
    h.set_profiling(True)
    modes = {"gather": P.ASSEMBLY_GATHER, "coloured": P.ASSEMBLY_COLOURED}
    h.set_assembly(modes[args.assembly])
    tuning = None
    if args.assembly == "gather" and not args.no_autotune:
        t0 = time.perf_counter()
        h.autotune(rd, defl=dd)  # the library's own tuner picks the gather tile
        tuning = h.tuning()
        print(f"[bench] J2: autotuned gather tile {tuning['gather_waves']} nodes/workgroup x "
              f"{tuning['gather_batch']} blocks in flight ({time.perf_counter() - t0:.1f}s)", file=sys.stderr,
              flush=True)
    s_step, kern_ms, n_launch = _timed_passes(h, f, vals, rd, dd, args.steps, args.warmup)
    nbytes = algorithmic_bytes(h.n_qp, h.nnz, prob.ndof, prob.mesh.n_nodes, prob.mesh.n_elem, 10)
    res = dict(n_qp=h.n_qp, nnz=h.nnz, ndof=prob.ndof, n_elem=prob.mesh.n_elem, s_step=s_step, kern_ms=kern_ms,
               n_launch=n_launch, achieved=nbytes / (kern_ms * 1e-3) / 1e9,
               tflops=flops_per_qp(10) * h.n_qp / (kern_ms * 1e-3) / 1e12, tuning=tuning)
    print(f"[bench] J2 ({args.assembly} assembly): {1e3 * s_step:.3f} ms/step wall, element pass {kern_ms:.3f} ms "
          f"over {n_launch} launches, {res['achieved']:.0f} GB/s algorithmic, {res['tflops']:.2f} TFLOP/s",
          file=sys.stderr, flush=True)
    other = "coloured" if args.assembly == "gather" else "gather"
    h.set_assembly(modes[other])
    s2, k2, n2 = _timed_passes(h, f, vals, rd, dd, max(3, args.steps // 4), 1)
    res["other"] = {"assembly": other, "ms_per_step": 1e3 * s2, "kernel_ms_per_pass": k2, "launches_per_pass": n2}
    h.commit_step()  # state roll-over: copies of known size (calibration of HBM counters)
    print(f"[bench] J2 ({other} assembly): {1e3 * s2:.3f} ms/step wall, element pass {k2:.3f} ms over {n2} "
          "launches", file=sys.stderr, flush=True)
    h.close()
    return res


def run_tf(args, dev):
    """Three-field (CM3F) passes of a neo-Hookean mesh of the J2 workload's size (48^3 P2 tets,
    E=70 GPa, nu=0.33), theta and P at their initial conditions for the displacement; the same
    mesh's displacement-formulation pass is timed beside it.  Algorithmic bytes per pass: F(n+1)
    written (9 doubles per QP), per element theta, P and their increments read and the update
    data written (4 + 34 doubles), the CSR values, residual, nodes and connectivity."""
    import torch

    import pgfem3d_amd as P
    from problems import BoxProblem

    prob = BoxProblem(args.n, P.PGF_TET10, P.neo_hookean(70e3, 0.33))
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    r, defl = prob.displacement(0.015, 2e-4, seed=11)
    rd = torch.from_numpy(r).to(dev)
    dd = torch.from_numpy(defl).to(dev)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=dev)
    vals = torch.empty(h.nnz, dtype=torch.float64, device=dev)
    h.set_profiling(True)
    steps = max(3, args.steps // 2)
    s_1f, k_1f, _ = _timed_passes(h, f, vals, rd, dd, steps, 1)
    h.set_three_field(True)
    h.tf_initial(rd, dd)
    s_tf, k_tf, n_l = _timed_passes(h, f, vals, rd, dd, steps, 1)
    ne, nq = prob.mesh.n_elem, h.n_qp
    nbytes = 8 * (9 * nq + 38 * ne + h.nnz + prob.ndof + 6 * prob.mesh.n_nodes) + 4 * 10 * ne
    res = dict(n_qp=nq, n_elem=ne, s_step=s_tf, kern_ms=k_tf, n_launch=n_l, kern_ms_1f=k_1f, s_step_1f=s_1f,
               achieved=nbytes / (k_tf * 1e-3) / 1e9)
    print(f"[bench] CM3F neo-Hookean {ne} P2 tets: {1e3 * s_tf:.3f} ms/step wall, pass {k_tf:.3f} ms over {n_l} "
          f"launches ({res['achieved']:.0f} GB/s algorithmic); displacement formulation {k_1f:.3f} ms",
          file=sys.stderr, flush=True)
    h.close()
    return res


def run_cp_tf(args, dev):
    """CM3F crystal plasticity: the CP workload's mesh (55^3 P1 tets, 64 grains) with the
    hyperelastic elastic law, five load steps of 1e-3 then the sixth timed, in the three-field
    mode (theta = 1, P = 0 at the start of each step) and in the displacement formulation on a
    handle of its own with the same history."""
    import torch

    import pgfem3d_amd as P
    from problems import CPBoxProblem, fcc_crystal_he

    mat = fcc_crystal_he()
    prob = CPBoxProblem(args.cp_n, P.PGF_TET4, mat, n_grains=64, seed=2)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=dev)
    out = {}
    for tf in (False, True):
        h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
        h.set_orientations(prob.R, prob.grain)
        vals = torch.empty(h.nnz, dtype=torch.float64, device=dev)
        if tf:
            h.set_three_field(True)
            h.tf_initial(torch.zeros(prob.ndof, dtype=torch.float64, device=dev),
                         torch.zeros(1, dtype=torch.float64, device=dev))
        for s in range(1, 6):
            r, defl = prob.displacement(CP_STEP * s, 0.0, seed=s)
            err = h.residual(f, torch.from_numpy(r).to(dev), defl=torch.from_numpy(defl).to(dev), dt=1.0)
            assert err == 0, (tf, s, err)
            h.commit_step()
        r, defl = prob.displacement(CP_STEP * 6, 0.0, seed=6)
        h.set_profiling(True)
        s_step, kern_ms, n_launch = _timed_passes(h, f, vals, torch.from_numpy(r).to(dev),
                                                  torch.from_numpy(defl).to(dev), args.cp_steps, 1)
        out["tf" if tf else "1f"] = (s_step, kern_ms, n_launch, float(h.get_state(P.STATE_ITERS)[0].mean()))
        h.close()
    s_tf, k_tf, n_l, its = out["tf"]
    res = dict(n_qp=prob.nqp, s_step=s_tf, kern_ms=k_tf, n_launch=n_l, mean_iters=its, kern_ms_1f=out["1f"][1],
               s_step_1f=out["1f"][0])
    print(f"[bench] CM3F CP (hyperelastic law) {prob.mesh.n_elem} P1 tets: {1e3 * s_tf:.3f} ms/step wall, pass "
          f"{k_tf:.3f} ms over {n_l} launches, local iterations {its:.2f}; displacement formulation "
          f"{out['1f'][1]:.3f} ms", file=sys.stderr, flush=True)
    return res


def run_cp(args, dev):
    import torch

    import pgfem3d_amd as P
    from problems import CPBoxProblem

    mat = P.fcc_crystal()
    t0 = time.perf_counter()
    prob = CPBoxProblem(args.cp_n, P.PGF_TET4, mat, n_grains=64, seed=2)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.R, prob.grain)
    print(f"[bench] CP: {prob.mesh.n_elem} P1 tets, ndof {prob.ndof}, nnz {h.nnz}, QPs {h.n_qp}, colours "
          f"{h.n_colors}; set-up {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=dev)
    vals = torch.empty(h.nnz, dtype=torch.float64, device=dev)
    log = None
    if not args.no_surrogate:
        # training data: the plain configuration-3 run logged over its load history
        # (args.sur_steps steps of 1e-3, 20 by default: 20M QP states), on a handle of its own
        import surrogate as S

        n_log = args.sur_steps
        log = (torch.empty((n_log * h.n_qp, S.N_FEAT), dtype=torch.float32, device=dev),
               torch.empty((n_log * h.n_qp, S.N_OUT), dtype=torch.float32, device=dev))
        hl = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
        hl.set_orientations(prob.R, prob.grain)
        row = 0
        t0 = time.perf_counter()
        for s in range(1, n_log + 1):
            r, defl = prob.displacement(CP_STEP * s, 0.0, seed=s)
            err = hl.residual(f, torch.from_numpy(r).to(dev), defl=torch.from_numpy(defl).to(dev), dt=1.0)
            assert err == 0, (s, err)
            row = S.log_step(hl, log[0], log[1], row)
            hl.commit_step()
        hl.close()
        print(f"[bench] CP: logged {row} QP states over {n_log} plain steps in {time.perf_counter() - t0:.1f}s",
              file=sys.stderr, flush=True)
    n_hist = 5
    for s in range(1, n_hist + 1):  # load history to 0.5%
        r, defl = prob.displacement(CP_STEP * s, 0.0, seed=s)
        err = h.residual(f, torch.from_numpy(r).to(dev), defl=torch.from_numpy(defl).to(dev), dt=1.0)
        assert err == 0, (s, err)
        h.commit_step()
    r, defl = prob.displacement(CP_STEP * 6, 0.0, seed=6)
    rd = torch.from_numpy(r).to(dev)
    dd = torch.from_numpy(defl).to(dev)
    h.set_profiling(True)
    tuning = None
    if not args.no_autotune:  # points per wavefront of the work queue and the gather tile
        h.autotune(rd, defl=dd)
        tuning = h.tuning()
        print(f"[bench] CP: autotuned {tuning['cp_chunk']} QPs per wave range, gather tile "
              f"{tuning['gather_waves']} x {tuning['gather_batch']}", file=sys.stderr, flush=True)
    s_step, kern_ms, n_launch = _timed_passes(h, f, vals, rd, dd, args.cp_steps, 1)
    its = h.get_state(P.STATE_ITERS)[0]
    mean_it = float(its.mean())
    tflops = cp_flops_per_qp(mean_it) * h.n_qp / (kern_ms * 1e-3) / 1e12
    res = dict(n_qp=h.n_qp, s_step=s_step, kern_ms=kern_ms, n_launch=n_launch, mean_iters=mean_it,
               max_iters=int(its.max()), tflops=tflops, surrogate=None, tuning=tuning)
    print(f"[bench] CP: {1e3 * s_step:.3f} ms/step wall, kernels {kern_ms:.3f} ms/pass over {n_launch} launches, "
          f"local Newton iterations mean {mean_it:.2f} max {int(its.max())}, {tflops:.2f} TFLOP/s",
          file=sys.stderr, flush=True)
    if log is not None:
        # learned warm start: trained on the logged plain run by the library's kernels (seeded
        # 90/10 split), then the same step timed with the network in the loop
        tau_plain = h.get_state(P.STATE_TAU)
        pF_plain = h.get_state(P.STATE_PLASTIC_NP1)
        n_samples = int(log[0].shape[0])
        t0 = time.perf_counter()
        mlp, nrm, hist = S.train(log[0], log[1], epochs=args.sur_epochs, batch=8192, lr=2e-3, seed=0)
        t_train = time.perf_counter() - t0
        del log
        S.attach(h, mlp, nrm)
        s_sur, kern_sur, _ = _timed_passes(h, f, vals, rd, dd, args.cp_steps, 1)
        its_s = h.get_state(P.STATE_ITERS)[0]
        n_learned, n_fallback = h.cp_start_stats()
        tau_s = h.get_state(P.STATE_TAU)
        dtau = float(np.abs(tau_s - tau_plain).max() / np.abs(tau_plain).max())
        dpf = float(np.abs(h.get_state(P.STATE_PLASTIC_NP1) - pF_plain).max())
        h.set_surrogate(None)
        mlp.close()
        # the bound any start can reach: the converged point itself as the start (M = pF(n)
        # pF(n+1)^-1, lambda, g in device QP order)
        order = h.qp_order()
        pFn = h.get_state(P.STATE_PLASTIC_N).T.reshape(-1, 3, 3)
        pF1 = pF_plain.T.reshape(-1, 3, 3)
        sc = h.get_state(P.STATE_SCALAR_NP1)
        M = pFn @ np.linalg.inv(pF1)
        warm_mesh = np.concatenate([M.reshape(-1, 9).T, sc[1][None], sc[0][None]], axis=0)
        warm = torch.from_numpy(np.ascontiguousarray(warm_mesh[:, order])).to(dev)
        h.set_warm_start(warm)
        s_ub, kern_ub, _ = _timed_passes(h, f, vals, rd, dd, args.cp_steps, 1)
        its_ub = h.get_state(P.STATE_ITERS)[0]
        h.set_warm_start(None)
        res["surrogate"] = dict(
            value=h.n_qp / s_sur, ms_per_step=1e3 * s_sur, kernel_ms_per_pass=kern_sur, speedup=s_step / s_sur,
            mean_local_iters=float(its_s.mean()), iterations_saved=mean_it - float(its_s.mean()),
            qps_from_learned_start=n_learned, qps_fallen_back=n_fallback,
            max_rel_tau_diff_vs_plain=dtau, max_abs_pF_diff_vs_plain=dpf, train_s=t_train,
            train_samples=n_samples, logged_steps=args.sur_steps, split="seeded 90/10", epochs=args.sur_epochs,
            val_loss=[float(x[2]) for x in hist],
            upper_bound=dict(start="the converged point", ms_per_step=1e3 * s_ub, kernel_ms_per_pass=kern_ub,
                             mean_local_iters=float(its_ub.mean())),
            upper_bound_speedup=s_step / s_ub)
        print(f"[bench] CP learned start: {1e3 * s_sur:.3f} ms/step, speedup {s_step / s_sur:.3f}x, iterations "
              f"mean {its_s.mean():.2f} (plain {mean_it:.2f}), {n_learned} QPs from the learned start, "
              f"{n_fallback} fell back, max rel tau diff {dtau:.1e}, trained on {n_samples} states in "
              f"{t_train:.1f}s, val loss {hist[0][2]:.3e}->{hist[-1][2]:.3e}; converged-point start "
              f"{1e3 * s_ub:.3f} ms/step ({its_ub.mean():.2f} iterations, bound {s_step / s_ub:.3f}x)",
              file=sys.stderr, flush=True)
    h.close()
    return res


def run_mixed(args, dev):
    """Mixed batch: 41^3 x 6 linear tets (413,526), 32^3 x 6 quadratic tets (196,608) and
    46^3 hexes (97,336, 8 QPs) in 64 crystal-plasticity grains, interface elements (exponential
    law, sig_max = 500 MPa, delta_c = 1 um) on the grain boundaries of the linear-tet and hex
    boxes; one pass = every element type and the interfaces.  Timed eager and as one captured
    hipGraph per pass, after autotuning the local-Newton work queue."""
    import torch

    import pgfem3d_amd as P
    from problems import MixedProblem

    t0 = time.perf_counter()
    coh = P.Cohesive(sig_c=500.0, chi_c=1e-3, beta=0.5, alpha=1.0)
    prob = MixedProblem(args.mix[0], args.mix[1], args.mix[2], P.fcc_crystal(), coh, n_grains=64, seed=2)
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    h.set_orientations(prob.R, prob.grain)
    h.set_cohesive_props(prob.coh)
    t_set = time.perf_counter() - t0
    counts = np.unique(np.diff(prob.mesh.elem_ptr), return_counts=True)
    print(f"[bench] mixed: elements {dict(zip(counts[0].tolist(), counts[1].tolist()))}, interfaces "
          f"{prob.mesh.n_coel}, QPs {h.n_qp} + {h.n_cqp} interface, ndof {prob.ndof}, nnz {h.nnz}; set-up "
          f"{t_set:.1f}s", file=sys.stderr, flush=True)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=dev)
    vals = torch.empty(h.nnz, dtype=torch.float64, device=dev)
    for s in range(1, 4):
        r, defl = prob.displacement(CP_STEP * s, 0.0, seed=s)
        assert h.residual(f, torch.from_numpy(r).to(dev), defl=torch.from_numpy(defl).to(dev)) == 0
        h.commit_step()
    r, defl = prob.displacement(CP_STEP * 4, 0.0, seed=4)
    rd, dd = torch.from_numpy(r).to(dev), torch.from_numpy(defl).to(dev)
    chunk = h.autotune(rd, defl=dd)
    h.set_profiling(True)
    s_eager, k_eager, nl = _timed_passes(h, f, vals, rd, dd, args.cp_steps, 1)
    h.set_graph(True)
    s_graph, k_graph, _ = _timed_passes(h, f, vals, rd, dd, args.cp_steps, 1)
    h.set_graph(False)
    its = h.get_state(P.STATE_ITERS)[0]
    nq = h.n_qp + h.n_cqp
    res = dict(n_qp=h.n_qp, n_cqp=h.n_cqp, n_coel=int(prob.mesh.n_coel), elements=counts[1].tolist(),
               launches=nl, chunk=chunk, eager=dict(value=nq / s_eager, ms_per_step=1e3 * s_eager, kernel_ms=k_eager),
               graph=dict(value=nq / s_graph, ms_per_step=1e3 * s_graph, kernel_ms=k_graph),
               mean_local_iters=float(its.mean()), max_local_iters=int(its.max()))
    print(f"[bench] mixed: eager {1e3 * s_eager:.3f} ms/step, graph {1e3 * s_graph:.3f} ms/step over {nl} launches, "
          f"autotuned chunk {chunk}, {nq / s_graph:.3e} QP/s", file=sys.stderr, flush=True)
    h.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=48, help="cells per cube edge (J2 workload)")
    ap.add_argument("--cp-n", type=int, default=55, help="cells per cube edge (crystal-plasticity workload)")
    ap.add_argument("--cp-steps", type=int, default=10)
    ap.add_argument("--cpu-n", type=int, default=18, help="cells per edge of the J2 CPU baseline sample")
    ap.add_argument("--cpu-cp-n", type=int, default=34, help="cells per edge of the CP CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cp", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-tf", action="store_true", help="skip the three-field (CM3F) passes")
    ap.add_argument("--mix", type=int, nargs=3, default=[41, 32, 46],
                    help="cells per edge of the linear-tet, quadratic-tet and hex boxes of the mixed batch")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the J2 and CP passes at the default launch configuration")
    ap.add_argument("--sur-epochs", type=int, default=3)
    ap.add_argument("--sur-steps", type=int, default=20,
                    help="plain load steps logged as training data for the learned start (20: 20M QP states)")
    ap.add_argument("--assembly", choices=["gather", "coloured"], default="gather",
                    help="assembly mode of the timed passes (the other mode is timed for comparison)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (multi-GPU sharding is a "
              "design note only)", file=sys.stderr)
        sys.exit(2)

    import torch

    dev = torch.device("cuda:0")
    j2 = run_j2(args, dev)
    cp = None if args.no_cp else run_cp(args, dev)
    mixed = None if args.no_mixed else run_mixed(args, dev)
    tf = None if args.no_tf else run_tf(args, dev)
    cptf = None if (args.no_tf or args.no_cp) else run_cp_tf(args, dev)

    cpu = cpu_cp = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_n)
        print(f"[bench] cpu baseline {cpu['value']:.0f} QP/s ({cpu['sample']})", file=sys.stderr, flush=True)
        if cp is not None:
            cpu_cp = cpu_baseline_cp(args.cpu_cp_n)
            print(f"[bench] cpu baseline CP {cpu_cp['value']:.0f} QP/s ({cpu_cp['sample']})", file=sys.stderr,
                  flush=True)

    out = {
        "metric": METRIC,
        "value": j2["n_qp"] / j2["s_step"],
        "unit": "QP updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * j2["s_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered box mesh and displacement)",
        "config": {
            "workload": f"J2 plasticity (linear isotropic hardening), {args.n}^3 cells x 6 quadratic tets "
                        f"({j2['n_elem']} elements, {j2['n_qp']} QPs, {j2['ndof']} free DOFs, nnz {j2['nnz']}); "
                        "one fused residual + tangent + CSR assembly pass per step",
            "element": "tet10, 4 QPs",
            "model": "J2 E=70GPa nu=0.33 sigma_y=250MPa H=1GPa",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": j2["achieved"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": j2["achieved"] / HBM_PEAK_GBS,
            "traffic": measured_traffic(f"j2_tet10_{args.n}cube"),
        },
        "fp64": fp64_entry(j2["tflops"], measured_flops(f"j2_tet10_{args.n}cube"), j2["kern_ms"]),
        "tuning": j2["tuning"],
        "kernel_ms_per_pass": j2["kern_ms"],
        "launches_per_pass": j2["n_launch"],
        "assembly": args.assembly,
        "other_assembly": j2["other"],
        "cpu_baseline": cpu,
    }
    if cp is not None:
        out["cp"] = {
            "workload": f"FCC crystal plasticity, {args.cp_n}^3 cells x 6 linear tets ({cp['n_qp']} QPs), 64 "
                        "Voronoi grains, step 6 of 1e-3/s at dt = 1 s, plain local-Newton start",
            "value": cp["n_qp"] / cp["s_step"],
            "unit": "QP updates/s",
            "ms_per_step": 1e3 * cp["s_step"],
            "tuning": cp["tuning"],
            "kernel_ms_per_pass": cp["kern_ms"],
            "mean_local_iters": cp["mean_iters"],
            "max_local_iters": cp["max_iters"],
            "roofline": {"bound": "fp64", "achieved": cp["tflops"], "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": cp["tflops"] / FP64_PEAK_TFLOPS,
                         **fp64_counter_fields(measured_flops(f"cp_tet4_{args.cp_n}cube"), cp["kern_ms"])},
            "surrogate_speedup": None if cp["surrogate"] is None else cp["surrogate"]["speedup"],
            "surrogate": cp["surrogate"],
            "cpu_baseline": cpu_cp,
        }
    if mixed is not None:
        out["mixed"] = {
            "workload": f"{args.mix[0]}^3 x 6 tet4 + {args.mix[1]}^3 x 6 tet10 + {args.mix[2]}^3 hex8 in 64 CP grains, "
                        f"{mixed['n_coel']} interface elements on tet4/hex8 grain boundaries; step 4 of 1e-3/s",
            "value": mixed["graph"]["value"], "unit": "QP updates/s (volume + interface QPs)",
            "eager": mixed["eager"], "graph": mixed["graph"], "launches_per_pass": mixed["launches"],
            "autotuned_chunk": mixed["chunk"], "mean_local_iters": mixed["mean_local_iters"],
        }
    if tf is not None:
        out["three_field"] = {
            "workload": f"CM3F (one volume and one pressure per element, condensed) neo-Hookean E=70GPa nu=0.33, "
                        f"{args.n}^3 cells x 6 quadratic tets ({tf['n_elem']} elements, {tf['n_qp']} QPs); one fused "
                        "residual + tangent + CSR assembly pass per step",
            "value": tf["n_qp"] / tf["s_step"], "unit": "QP updates/s", "ms_per_step": 1e3 * tf["s_step"],
            "kernel_ms_per_pass": tf["kern_ms"], "launches_per_pass": tf["n_launch"],
            "displacement_formulation_kernel_ms_per_pass": tf["kern_ms_1f"],
            "roofline": {"bound": "hbm", "achieved": tf["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": tf["achieved"] / HBM_PEAK_GBS},
        }
    if cptf is not None:
        flops = measured_flops(f"cp_tf_tet4_{args.cp_n}cube")
        ach = None if flops is None else flops / (cptf["kern_ms"] * 1e-3) / 1e12
        out["three_field_cp"] = {
            "workload": f"CM3F FCC crystal plasticity with the hyperelastic elastic law, {args.cp_n}^3 cells x 6 "
                        f"linear tets ({cptf['n_qp']} QPs), 64 Voronoi grains, step 6 of 1e-3/s at dt = 1 s "
                        "(theta = 1, P = 0 at the step's start); one fused residual + tangent + CSR assembly pass",
            "value": cptf["n_qp"] / cptf["s_step"], "unit": "QP updates/s", "ms_per_step": 1e3 * cptf["s_step"],
            "kernel_ms_per_pass": cptf["kern_ms"], "launches_per_pass": cptf["n_launch"],
            "mean_local_iters": cptf["mean_iters"],
            "displacement_formulation_kernel_ms_per_pass": cptf["kern_ms_1f"],
            "roofline": {"bound": "fp64", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": None if ach is None else ach / FP64_PEAK_TFLOPS,
                         "basis": "executed FP64 instructions (profiles/fp64_counters.md) over the live pass time"},
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
