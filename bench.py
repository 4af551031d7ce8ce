"""Benchmark of the finite-strain element hot path on one MI355X.

Workload (default): J2 finite-strain plasticity with linear isotropic hardening on a unit cube
of 48^3 cells, each split into 6 quadratic 10-node tets (663,552 elements, 2,654,208 QPs,
4-point quadrature), E=70 GPa, nu=0.33, sigma_y=250 MPa, H=1 GPa, FP64.  Bottom face fixed,
top face stretched; the displacement is a seeded 1.5% uniaxial stretch (plastic at every QP).

One step = one global Newton iteration's element work: deformation gradient, J2 return map
and state update at every QP, element residual and tangent, and their assembly into the
residual vector and the CSR values (Psparse_ApAi pattern), all resident in HBM.  The linear
solve is not part of the step.  value = QPs / second of wall time per step.
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


def algorithmic_bytes(n_qp, nnz, ndof, n_nodes, n_elem, nne):
    """Unique HBM traffic of one step (bytes): J2 state read (Cp^-1_n 6 + ep_n 1 doubles) and
    written (F 9 + Cp^-1 6 + ep, gam 2 doubles), each CSR value and residual entry written
    once, coordinates and displacements (6 doubles per node) and connectivity (4 B per
    element node) read once."""
    state = n_qp * 8 * (7 + 17)
    return state + 8 * nnz + 8 * ndof + 8 * 6 * n_nodes + 4 * nne * n_elem


def flops_per_qp(nne):
    """FP64 operations per QP of the fused kernel, counted from the source (FMA = 2):
    phase 1 ~ (9*nne*2 kinematics + 9*nne h_a + ~700 constitutive/tangent), phase 2 per
    element row: 36 MACs (b^T c) + 9 (tau h_a) + 12*nne (R h_w and geometric term)."""
    ndofe = 3 * nne
    phase1 = 2 * (9 * nne * 2 + 9 * nne) + 700
    phase2 = ndofe * 2 * (36 + 9 + 3 + 12 * nne)
    return phase1 + phase2


def cpu_baseline(n_cells, timeout_s=30.0):
    """Oracle (CPU restatement of the reference path, one thread) on a bounded sample: the
    same element type and material on an n_cells^3 mesh; one fused residual + tangent pass."""
    import oracle as O
    import pgfem3d_amd as P
    from problems import BoxProblem

    prob = BoxProblem(n_cells, P.PGF_TET10, P.j2_linear(70e3, 0.33, 250.0, 1000.0))
    r, defl = prob.displacement(0.015, 2e-4, seed=11)
    st = O.OracleState(prob.nqp, True)
    t0 = time.perf_counter()
    O.assemble(prob.mesh, prob.mats, prob.ids, prob.ndof, prob.Ap, prob.Ai, st, r, defl=defl)
    dt = time.perf_counter() - t0
    return {"value": prob.nqp / dt, "unit": "QP updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle orc_assemble, one fused residual+tangent pass, J2 P2 {n_cells}^3 cells "
                      f"({prob.mesh.n_elem} elements, {prob.nqp} QPs), {dt:.1f} s on 1 host thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=48, help="cells per cube edge")
    ap.add_argument("--cpu-n", type=int, default=18, help="cells per edge of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (multi-GPU sharding is a "
              "design note only)", file=sys.stderr)
        sys.exit(2)

    import torch

    import pgfem3d_amd as P
    from problems import BoxProblem

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    mat = P.j2_linear(70e3, 0.33, 250.0, 1000.0)
    prob = BoxProblem(args.n, P.PGF_TET10, mat)
    t_setup = time.perf_counter() - t0
    h = P.Handle(prob.mesh, prob.ids, prob.ndof, prob.Ap, prob.Ai, prob.mats)
    t_create = time.perf_counter() - t0 - t_setup
    print(f"[bench] mesh {prob.mesh.n_elem} P2 tets, {prob.mesh.n_nodes} nodes, ndof {prob.ndof}, "
          f"nnz {h.nnz}, QPs {h.n_qp}, colours {h.n_colors}; host set-up {t_setup:.1f}s, create {t_create:.1f}s",
          file=sys.stderr, flush=True)
    r, defl = prob.displacement(0.015, 2e-4, seed=11)
    rd = torch.from_numpy(r).to(dev)
    dd = torch.from_numpy(defl).to(dev)
    f = torch.empty(prob.ndof, dtype=torch.float64, device=dev)
    vals = torch.empty(h.nnz, dtype=torch.float64, device=dev)
    h.set_profiling(True)
    for _ in range(args.warmup):
        err = h.residual_tangent(f, vals, rd, defl=dd)
        assert err == 0, err
    torch.cuda.synchronize()
    pass_ms = []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        err = h.residual_tangent(f, vals, rd, defl=dd)
        pass_ms.append(h.last_pass_ms()[0])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    n_launch = h.last_pass_ms()[1]
    assert err == 0
    ms_step = 1e3 * wall / args.steps
    value = h.n_qp / (wall / args.steps)
    kern_ms = float(np.mean(pass_ms))  # GPU time of the element launches of one pass (HIP events)
    nbytes = algorithmic_bytes(h.n_qp, h.nnz, prob.ndof, prob.mesh.n_nodes, prob.mesh.n_elem, 10)
    achieved = nbytes / (kern_ms * 1e-3) / 1e9
    tflops = flops_per_qp(10) * h.n_qp / (kern_ms * 1e-3) / 1e12
    print(f"[bench] {ms_step:.3f} ms/step wall, element kernels {kern_ms:.3f} ms/pass over {n_launch} launches, "
          f"{achieved:.0f} GB/s algorithmic, {tflops:.2f} TFLOP/s", file=sys.stderr, flush=True)
    h.close()
    del vals

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_n)
        print(f"[bench] cpu baseline {cpu['value']:.0f} QP/s ({cpu['sample']})", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "QP updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered box mesh and displacement)",
        "config": {
            "workload": f"J2 plasticity (linear isotropic hardening), {args.n}^3 cells x 6 quadratic tets "
                        f"({prob.mesh.n_elem} elements, {h.n_qp} QPs, {prob.ndof} free DOFs, nnz {h.nnz}); "
                        "one fused residual + tangent + CSR assembly pass per step",
            "element": "tet10, 4 QPs",
            "model": "J2 E=70GPa nu=0.33 sigma_y=250MPa H=1GPa",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
        "fp64": {"achieved_tflops": tflops, "peak_tflops": FP64_PEAK_TFLOPS, "frac": tflops / FP64_PEAK_TFLOPS},
        "kernel_ms_per_pass": kern_ms,
        "launches_per_pass": n_launch,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
