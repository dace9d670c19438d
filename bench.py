#!/usr/bin/env python3
"""Benchmark: Tet10 assemble + Jacobi-PCG on one MI355X.

Headline workload: steady -div(k grad T) = Q on a 1000 x 1000 x 600 km box,
6-tet Kuhn Tet10 mesh at nx=ny=64, nz=32 (786,432 elements, 1,081,665 DOFs),
k=3, Q=1e-6, T=1280 at z=0 and T=0 at z=600.  Synthetic input (generated
mesh), resident in HBM before timing.

One step = numeric assembly (element kernels and the cached system plan's
fused gather: K_bc values in CSR and SELL layouts, Dirichlet lift, load,
Jacobi diagonal) + Jacobi-PCG to rel-res 1e-8 by the reference's criterion
||r|| < 1e-8 ||b||.  Timed for the plain FP64 path and the mixed path (FP64
flexible CG around an FP32 Jacobi-PCG recursion).  The symbolic phase (CSR
pattern, contributor lists, system plan, SELL layout) runs once per mesh and
DOF map in the warm-up and is reported separately.

Also reported (same JSON line): the elastic column at nx=ny=48, nz=24
(1,383,123 DOFs), and the adjoint inversion gradients (thermal log k, elastic
log E) with finite-difference checks.

Prints ONE JSON line.  value = mixed-precision assemble+PCG seconds per step
of the headline workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: (nx, ny, nz)
    "thermal_64x64x32": (64, 64, 32),
    "thermal_32x32x16": (32, 32, 16),
    "thermal_6x6x6": (6, 6, 6),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sync():
    import torch
    torch.cuda.synchronize()


def host_cpu():
    """The host the CPU baseline ran on: CPU model, logical CPUs the OS
    reports, and those this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline(nx, ny, nz, threads):
    """Oracle (C restatement of the reference algorithm) on the same workload:
    parallel element loop, serial triplet merge / CSR / Dirichlet / CG, as in
    the reference, all timed: assembly, and Jacobi-CG to rel-res 1e-8.  A few
    BiCGSTAB passes and GMRES steps give their per-pass costs."""
    from oracle import oracle as O
    xyz, conn = O.generate_cube(nx, ny, nz, 1000.0, 1000.0, 600.0)
    mask, vals = O.thermal_dirichlet(xyz, 600.0)
    t0 = time.perf_counter()
    K = O.assemble_thermal(xyz, conn, 3.0, nthreads=threads)
    f = O.assemble_thermal_load(xyz, conn, 1e-6)
    Kbc, fbc = O.apply_dirichlet(K, f, mask, vals)
    t_asm = time.perf_counter() - t0
    del K
    t0 = time.perf_counter()
    _, st = O.cg_solve(Kbc, fbc, max_iter=5000, tol=1e-8)
    t_cg = time.perf_counter() - t0
    per_iter = t_cg / max(1, st.iterations + 1)
    # a parallel CG beside the reference's serial one (timing only)
    _, sp = O.cg_solve_omp(Kbc, fbc, max_iter=5000, tol=1e-8, nthreads=threads)
    t_cg_par = sp.solve_time
    t0 = time.perf_counter()
    _, sb = O.bicgstab_solve(Kbc, fbc, max_iter=5, tol=1e-8)
    bicg_per_iter = (time.perf_counter() - t0) / max(1, sb.iterations)
    t0 = time.perf_counter()
    _, sg = O.gmres_solve(Kbc, fbc, max_iter=10, restart=30, tol=1e-8)
    gm_per_step = (time.perf_counter() - t0) / max(1, sg.iterations)
    return (t_asm + t_cg, t_asm, t_cg, st, per_iter, bicg_per_iter, gm_per_step,
            (t_cg_par, sp.iterations))


def make_cg(G, prec, tol, inner_tol=None, abs_tol=None, max_iter=200000):
    cg = (G.ConjugateGradient().with_max_iterations(max_iter).with_tolerance(tol)
          .with_precision(prec))
    if abs_tol is not None:
        cg.with_abs_tolerance(abs_tol)
    if prec == "mixed" and inner_tol is not None:
        cg.with_inner_tolerance(inner_tol)
    return cg


def timed_steps(assemble, cg, steps, warmup):
    """warm-up, then `steps` timed assemble+solve steps; returns wall/step,
    mean assembly and solve seconds and the last (K, f, x, stats)."""
    for _ in range(warmup):
        K, f = assemble()
        cg.solve(K, f)
    sync()
    asm_t, sol_t, last = [], [], None
    t0 = time.perf_counter()
    for _ in range(steps):
        ta = time.perf_counter()
        K, f = assemble()
        sync()
        tb = time.perf_counter()
        x, st = cg.solve(K, f)
        sync()
        asm_t.append(tb - ta)
        sol_t.append(time.perf_counter() - tb)
        last = (K, f, x, st)
    wall = (time.perf_counter() - t0) / steps
    return wall, float(np.mean(asm_t)), float(np.mean(sol_t)), last


def run_headline(G, args, nx, ny, nz):
    import torch
    from geo_simulator_amd import workloads as W
    log(f"[bench] thermal {nx}x{ny}x{nz}: generating mesh")
    mesh, dof, exact = W.thermal_box(G, nx, ny, nz)
    n = mesh.num_nodes()
    ctx = G.default_context()
    sync()
    t0 = time.perf_counter()
    mesh.handle(ctx)
    dof.handle(ctx)
    sync()
    t_upload = time.perf_counter() - t0
    assemble = lambda: G.Assembler.assemble_thermal_system(mesh, dof, 3.0, 1e-6)
    t0 = time.perf_counter()
    K, f = assemble()
    sync()
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    K, f = assemble()
    sync()
    t_second = time.perf_counter() - t0
    t_symbolic = max(0.0, t_first - t_second)
    nnz = K.nnz()
    log(f"[bench] n={n} nnz={nnz} symbolic~{t_symbolic:.3f}s numeric asm {t_second*1e3:.1f} ms")

    results = {}
    for prec in ("fp64", "mixed"):
        cg = make_cg(G, prec, args.tol, args.inner_tol)
        wall, ta, ts, (K, f, x, st) = timed_steps(assemble, cg, args.steps, args.warmup)
        err = float(np.linalg.norm(x.cpu().numpy() - exact) / np.linalg.norm(exact))
        results[prec] = dict(step_s=wall, asm_s=ta, solve_s=ts, iterations=st.iterations,
                             outer=st.outer_iterations, spmv=st.spmv_count,
                             rel_res=st.relative_residual, converged=st.converged,
                             rel_l2_vs_closed_form=err)
        log(f"[bench] {prec}: {wall*1e3:.2f} ms/step  asm {ta*1e3:.2f} ms  "
            f"solve {ts*1e3:.2f} ms  it={st.iterations} outer={st.outer_iterations} "
            f"relres={st.relative_residual:.2e} relL2={err:.2e}")

    # FP64-equivalent accuracy: both paths to a tight residual
    tight = {}
    if not args.no_tight:
        for prec in ("fp64", "mixed"):
            cg = make_cg(G, prec, 1e-12, args.inner_tol)
            wall, ta, ts, (K, f, x, st) = timed_steps(assemble, cg, 1, 1)
            err = float(np.linalg.norm(x.cpu().numpy() - exact) / np.linalg.norm(exact))
            tight[prec] = dict(step_s=wall, iterations=st.iterations,
                               outer=st.outer_iterations, rel_res=st.relative_residual,
                               rel_l2_vs_closed_form=err)
            log(f"[bench] tight {prec}: {wall*1e3:.2f} ms it={st.iterations} relL2={err:.2e}")
        tight["speedup_mixed_vs_fp64"] = tight["fp64"]["step_s"] / tight["mixed"]["step_s"]

    # equal accuracy instead of equal residual: each path at the loosest
    # relative tolerance of a fixed ladder whose solution is within 1e-10
    # (rel L2) of the closed form, which the quadratic elements reproduce at
    # the nodes (so it is the converged FP64 solution to its own rounding)
    matched = {}
    if not args.no_tight:
        ladder = (1e-10, 3e-11, 1e-11, 5e-12, 3e-12, 2e-12, 1.5e-12, 1e-12, 8e-13, 6e-13,
                  5e-13, 3e-13, 2e-13, 1e-13)
        for prec in ("fp64", "mixed"):
            pick = None
            for tol in ladder:
                x, st = make_cg(G, prec, tol, args.inner_tol).solve(K, f)
                err = float(np.linalg.norm(x.cpu().numpy() - exact) / np.linalg.norm(exact))
                if err <= 1e-10:
                    pick = tol
                    break
            if pick is None:
                matched[prec] = dict(reached=False)
                continue
            cg = make_cg(G, prec, pick, args.inner_tol)
            wall, ta, ts, (_, _, x, st) = timed_steps(assemble, cg, 3, 1)
            err = float(np.linalg.norm(x.cpu().numpy() - exact) / np.linalg.norm(exact))
            matched[prec] = dict(reached=True, rel_tol=pick, step_s=wall, steps=3,
                                 iterations=st.iterations,
                                 outer=st.outer_iterations, rel_l2_vs_closed_form=err)
            log(f"[bench] rel L2 <= 1e-10 {prec}: tol {pick:g} {wall*1e3:.2f} ms "
                f"it={st.iterations} relL2={err:.2e}")
        if matched.get("fp64", {}).get("reached") and matched.get("mixed", {}).get("reached"):
            matched["speedup_mixed_vs_fp64"] = matched["fp64"]["step_s"] / matched["mixed"]["step_s"]

    # dominant kernels: the CG SpMV in each precision, timed with HIP events on
    # the library's own stream (the stream the kernels run on).  Algorithmic
    # bytes, CSR-equivalent (int32 column indices and row pointers): FP64
    # q = A p reads the operator and p once and writes q; the FP32
    # recursion's kernel (w = A r on the scaled operator, 4-byte fixed-point
    # values, then p = r + beta p,
    # s = w + beta s in place) reads the operator, r once, p and s, and
    # writes p and s.
    bytes64 = nnz * 12 + n * 20
    bytes32 = nnz * 8 + n * 24
    ms64 = K.time_spmv("fp64", args.spmv_reps)
    ms32 = K.time_spmv("mixed", args.spmv_reps)
    gbs64 = bytes64 / (ms64 * 1e-3) / 1e9
    gbs32 = bytes32 / (ms32 * 1e-3) / 1e9
    log(f"[bench] SpMV fp64 {ms64*1e3:.1f} us {gbs64:.0f} GB/s; fp32 {ms32*1e3:.1f} us {gbs32:.0f} GB/s")
    # the box's achievable HBM rates on buffers 4x the Infinity Cache
    from geo_simulator_amd.linalg import time_stream
    gib = 2 ** 30
    ms_r = time_stream(gib, "read", 10)
    ms_c = time_stream(gib, "copy", 10)
    ceiling = {"buffer_bytes": gib,
               "read_GBs": gib / (ms_r * 1e-3) / 1e9,
               "copy_GBs": 2 * gib / (ms_c * 1e-3) / 1e9,
               "kernels": "k_stream_sum (16-byte loads, FP64 sum), k_stream_copy"}
    log(f"[bench] HBM ceiling (1 GiB): read {ceiling['read_GBs']:.0f} GB/s, "
        f"copy {ceiling['copy_GBs']:.0f} GB/s")
    if args.copy_probe_mib > 0:
        # counter calibration: a copy of known size (k_stream_copy) in the
        # profiled command
        time_stream(args.copy_probe_mib * 2 ** 20, "copy", 3)
    traffic32 = traffic64 = traffic_tree = None
    pmc_path = os.path.join(ROOT, "profiles", f"spmv_{args.config}_pmc.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as fh:
            pmc = json.load(fh)
        traffic32 = pmc.get("fp32", {}).get("hbm_bytes")
        traffic64 = pmc.get("fp64", {}).get("hbm_bytes")
        traffic_tree = pmc.get("tree")
    spmv = {"fp64_us": ms64 * 1e3, "fp64_GBs": gbs64, "fp64_bytes": bytes64,
            "fp64_frac": gbs64 / HBM_PEAK_GBS, "fp64_traffic": traffic64,
            "fp32_us": ms32 * 1e3, "fp32_GBs": gbs32, "fp32_bytes": bytes32,
            "fp32_frac": gbs32 / HBM_PEAK_GBS, "fp32_traffic": traffic32,
            "traffic_measured_on": traffic_tree, "hbm_ceiling": ceiling}
    if not args.no_bicgstab:
        # BiCGSTAB (FP64, Jacobi) on the same system: the general-matrix solver
        # on the same SELL operator, two SpMVs per pass
        bg = (G.BiCGSTAB().with_tolerance(args.tol).with_max_iterations(20000))
        bg.solve(K, f)
        sync()
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            x, st = bg.solve(K, f)
            sync()
            ts.append(time.perf_counter() - t0)
        err = float(np.linalg.norm(x.cpu().numpy() - exact) / np.linalg.norm(exact))
        ms = float(np.median(ts)) * 1e3
        results["bicgstab_fp64"] = dict(
            solve_s=ms * 1e-3, iterations=st.iterations, spmv=st.spmv_count,
            rel_res=st.relative_residual, converged=st.converged,
            rel_l2_vs_closed_form=err, ms_per_pass=ms / max(1, st.iterations),
            note="solve only, on the assembled system of the fp64 step")
        log(f"[bench] BiCGSTAB fp64: solve {ms:.2f} ms it={st.iterations} "
            f"({ms / max(1, st.iterations) * 1e3:.1f} us/pass) relL2={err:.2e}")
        # GMRES(30), Jacobi, on the same system
        gm = G.GMRES().with_restart(30).with_tolerance(args.tol).with_max_iterations(20000)
        gm.solve(K, f)
        sync()
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            x, st = gm.solve(K, f)
            sync()
            ts.append(time.perf_counter() - t0)
        err = float(np.linalg.norm(x.cpu().numpy() - exact) / np.linalg.norm(exact))
        ms = float(np.median(ts)) * 1e3
        results["gmres30_fp64"] = dict(
            solve_s=ms * 1e-3, iterations=st.iterations, spmv=st.spmv_count,
            rel_res=st.relative_residual, converged=st.converged,
            rel_l2_vs_closed_form=err, ms_per_step=ms / max(1, st.iterations),
            note="solve only, restart 30, on the assembled system of the fp64 step")
        log(f"[bench] GMRES(30) fp64: solve {ms:.2f} ms it={st.iterations} "
            f"({ms / max(1, st.iterations) * 1e3:.1f} us/step) relL2={err:.2e}")
    if not args.no_amg:
        # AMG (the reference's Ruge-Stueben-style preconditioner, default
        # parameters 10 levels, coarse size 500, strength 0.25): set-up on the
        # host, V-cycle on the device, CG through solve_with_operator
        t0 = time.perf_counter()
        amg = G.AMGPreconditioner(K, 10, 500, 0.25)
        setup_s = time.perf_counter() - t0
        amg.apply(f)
        sync()
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            amg.apply(f)
            sync()
            ts.append(time.perf_counter() - t0)
        vc_ms = float(np.median(ts)) * 1e3
        cga = G.ConjugateGradient().with_tolerance(args.tol).with_max_iterations(20000)
        ts = []
        for _ in range(2):
            t0 = time.perf_counter()
            x, st = cga.solve_with_operator(K, f, amg)
            sync()
            ts.append(time.perf_counter() - t0)
        err = float(np.linalg.norm(x.cpu().numpy() - exact) / np.linalg.norm(exact))
        ms = float(np.median(ts)) * 1e3
        results["cg_amg_fp64"] = dict(
            setup_s=setup_s, vcycle_ms=vc_ms, solve_s=ms * 1e-3, iterations=st.iterations,
            rel_res=st.relative_residual, converged=st.converged, rel_l2_vs_closed_form=err,
            levels=[lv["rows"] for lv in amg.levels],
            note="AMGPreconditioner(K, 10, 500, 0.25) built on the host (caller numbering), "
                 "V-cycle on the device; CG via solve_with_operator (host-driven loop, one "
                 "operator and one preconditioner callback per iteration)")
        log(f"[bench] CG+AMG fp64: setup {setup_s:.2f} s, V-cycle {vc_ms:.2f} ms, solve "
            f"{ms:.2f} ms it={st.iterations} relL2={err:.2e} levels "
            f"{[lv['rows'] for lv in amg.levels]}")
        del amg
    # numeric assembly against its algorithmic bytes: coordinates n*24,
    # connectivity E*40, the K_bc values nnz_bc*8 and the right-hand side n*8
    ne = mesh.num_elements()
    asm_bytes = n * 24 + ne * 40 + nnz * 8 + n * 8
    asm_s = results["fp64"]["asm_s"]
    assembly = {"bytes": asm_bytes, "s": asm_s, "GBs": asm_bytes / asm_s / 1e9,
                "frac": asm_bytes / asm_s / 1e9 / HBM_PEAK_GBS,
                "note": "wall clock of the plan-driven numeric assembly (element kernel, "
                        "system gather, diagonal maximum, right-hand side and Jacobi "
                        "diagonal) per step, host synchronisation included"}
    info = dict(n=n, nnz=nnz, elements=ne, symbolic_s=t_symbolic,
                upload_s=t_upload, assembly=assembly)
    return results, tight, matched, spmv, info


def run_elastic(G, args):
    from geo_simulator_amd import workloads as W
    nx, ny, nz = 48, 48, 24
    mesh, dof, uz = W.elastic_box(G, nx, ny, nz)
    mat = G.IsotropicElasticity(1e11, 0.25)
    assemble = lambda: G.Assembler.assemble_elasticity_system(mesh, dof, mat, 3300.0,
                                                              [0.0, 0.0, -9.81])
    t0 = time.perf_counter()
    K, f = assemble()
    sync()
    t_first = time.perf_counter() - t0
    out = dict(workload=f"elastic Tet10 gravity column {nx}x{ny}x{nz} (metres, "
                        f"{mesh.num_elements()} elements, {3 * mesh.num_nodes()} DOFs, "
                        f"nnz_bc {K.nnz()})", first_assembly_s=t_first)
    for prec in ("fp64", "mixed"):
        for tol in (1e-8, 1e-12):
            cg = make_cg(G, prec, tol, args.inner_tol)
            wall, ta, ts, (K, f, u, st) = timed_steps(assemble, cg, 2 if tol == 1e-8 else 1, 1)
            uh = u.cpu().numpy()
            err = float(np.linalg.norm(uh[2::3] - uz) / np.linalg.norm(uz))
            key = f"{prec}" if tol == 1e-8 else f"tight_{prec}"
            out[key] = dict(step_s=wall, asm_s=ta, solve_s=ts, iterations=st.iterations,
                            outer=st.outer_iterations, rel_res=st.relative_residual,
                            converged=st.converged, rel_l2_uz_vs_closed_form=err,
                            max_abs_ux_m=float(np.abs(uh[0::3]).max()))
            log(f"[bench] elastic {key}: {wall*1e3:.1f} ms/step asm {ta*1e3:.1f} ms "
                f"it={st.iterations} outer={st.outer_iterations} relL2={err:.2e}")
    out["speedup_mixed_vs_fp64"] = out["fp64"]["step_s"] / out["mixed"]["step_s"]
    nn3, ne_, nnz_ = 3 * mesh.num_nodes(), mesh.num_elements(), K.nnz()
    asm_bytes = mesh.num_nodes() * 24 + ne_ * 40 + nnz_ * 8 + nn3 * 8
    asm_s = out["fp64"]["asm_s"]
    out["assembly"] = {"bytes": asm_bytes, "s": asm_s, "GBs": asm_bytes / asm_s / 1e9,
                       "frac": asm_bytes / asm_s / 1e9 / HBM_PEAK_GBS}
    out["tight_speedup_mixed_vs_fp64"] = (out["tight_fp64"]["step_s"] /
                                          out["tight_mixed"]["step_s"])
    nnz, n = K.nnz(), 3 * mesh.num_nodes()
    b64, b32 = nnz * 12 + n * 20, nnz * 8 + n * 12
    ms64, ms32 = K.time_spmv("fp64", 20), K.time_spmv("mixed", 20)
    out["spmv"] = {"fp64_us": ms64 * 1e3, "fp64_GBs": b64 / ms64 / 1e6,
                   "fp32_us": ms32 * 1e3, "fp32_GBs": b32 / ms32 / 1e6}
    log(f"[bench] elastic SpMV fp64 {ms64*1e3:.1f} us fp32 {ms32*1e3:.1f} us")
    # element kernel alone: K_e (900 doubles), F_e (30) and the connectivity
    # per element, each node's coordinates once
    ne, nn = mesh.num_elements(), mesh.num_nodes()
    kv = {"note": "B^T D B + gravity load of every element; valu = the assembly kernel "
                  "(FP64, one lane per 3x3 node block, bit-identical to the reference "
                  "order), mfma = v_mfma_f64_16x16x4f64, *_f32 = FP32 B^T D B (VALU, "
                  "v_mfma_f32_16x16x4f32) writing FP32 K_e; valu_rows = one lane per K_e "
                  "row (the earlier FP64 layout); rel_diff vs the FP64 valu K_e"}
    for v in ("valu", "mfma", "valu_f32", "mfma_f32", "valu_rows"):
        sz = 4 if v.endswith("_f32") else 8
        eb = ne * (900 * sz + 30 * sz + 40) + nn * 24
        ms, d = G.time_elastic_elements(mesh, mat, None, v, reps=20)
        kv[v] = {"ms": ms, "algorithmic_bytes": eb, "GBs": eb / ms / 1e6,
                 "frac": eb / ms / 1e6 / HBM_PEAK_GBS, "rel_diff_vs_valu": d}
        log(f"[bench] elastic element kernel {v}: {ms:.3f} ms ({eb / ms / 1e6:.0f} GB/s) "
            f"max rel diff {d:.1e}")
    out["element_kernel"] = kv
    return out


def run_adjoint(G, args):
    """Inversion gradients: thermal d L / d log k_e and elastic d L / d log E_e
    from the adjoint (forward solve, reverse solve with the same SPD operator,
    element-backward kernel), mixed and FP64 solves, checked by central FD."""
    import torch
    from geo_simulator_amd import workloads as W
    from geo_simulator_amd.adjoint import (ElasticProblem, ThermalProblem, elastic_solve,
                                           thermal_solve)
    from geo_simulator_amd.gradcheck import corrected_solver, fd_central, fd_direction
    dev = "cuda"
    out = {}
    cases = []
    if not args.no_adjoint_thermal:
        cases.append("thermal")
    if not args.no_adjoint_elastic:
        cases.append("elastic")
    for case in cases:
        if case == "thermal":
            mesh, dof, theta, theta_true, obs = W.thermal_adjoint_case(G)
            mk = lambda solver: ThermalProblem(mesh, dof, 1e-6, solver=solver)
            apply = thermal_solve
            desc = (f"thermal 64x64x32, log k_e from a Gaussian field (corr 100 km, k in "
                    f"[1,5]), T observed at {len(obs)} free nodes")
        else:
            mesh, dof, theta, theta_true, obs = W.elastic_adjoint_case(G)
            mk = lambda solver: ElasticProblem(mesh, dof, 0.25, 3300.0, [0.0, 0.0, -9.81],
                                               solver=solver)
            apply = elastic_solve
            desc = (f"elastic 48x48x24, log E_e log-normal (median 1e11 Pa, sigma 0.5), "
                    f"u_z observed at {len(obs)} top-face nodes")
        obs_t = torch.tensor(obs, device=dev)
        fp64 = make_cg(G, "fp64", 1e-12, abs_tol=0.0)
        mixed = make_cg(G, "mixed", 1e-12, args.inner_tol, abs_tol=0.0)
        with torch.no_grad():
            y_obs = apply(torch.tensor(theta_true, device=dev), mk(fp64))[obs_t].clone()
        res = {"workload": desc, "loss": "0.5 ||y_obs - y||^2", "solve_rel_res": 1e-12}
        grads = {}
        for name, solver in (("mixed", mixed), ("fp64", fp64)):
            prob = mk(solver)
            for _ in range(2):  # the second run is timed (the first builds the plan)
                th = torch.tensor(theta, device=dev, requires_grad=True)
                sync()
                t0 = time.perf_counter()
                y = apply(th, prob)
                L = 0.5 * ((y[obs_t] - y_obs) ** 2).sum()
                L.backward()
                sync()
                dt = time.perf_counter() - t0
            grads[name] = th.grad.cpu().numpy()
            res[name] = dict(gradient_s=dt, loss=L.item(),
                             forward_iterations=prob.forward_stats.iterations,
                             adjoint_iterations=prob.backward_stats.iterations,
                             forward_converged=prob.forward_stats.converged,
                             adjoint_converged=prob.backward_stats.converged)
            log(f"[bench] adjoint {case} {name}: gradient {dt*1e3:.1f} ms "
                f"(fwd it {prob.forward_stats.iterations}, adj it "
                f"{prob.backward_stats.iterations})")
        g, g64 = grads["mixed"], grads["fp64"]
        res["speedup_mixed_vs_fp64"] = res["fp64"]["gradient_s"] / res["mixed"]["gradient_s"]
        res["mixed_vs_fp64_max_rel"] = float(np.abs(g - g64).max() / np.abs(g64).max())
        # central finite differences (FP64 solves, residual-corrected)
        if args.fd_elems > 0:
            import torch as _t
            fdcg = make_cg(G, "fp64", 1e-13, abs_tol=0.0, max_iter=20000)
            y0_full = apply(_t.tensor(theta, device=dev), mk(fdcg)).detach()
            if case == "thermal":
                asm = lambda th: G.Assembler.assemble_thermal_system(
                    mesh, dof, 1.0, 1e-6, k_per_elem=_t.exp(th).contiguous())
            else:
                mat1 = G.IsotropicElasticity(1.0, 0.25)
                asm = lambda th: G.Assembler.assemble_elasticity_system(
                    mesh, dof, mat1, 3300.0, [0.0, 0.0, -9.81],
                    E_per_elem=_t.exp(th).contiguous())
            th0 = torch.tensor(theta, device=dev)
            corr = corrected_solver(asm, th0, y0_full, fdcg)
            solve_obs = lambda th: (corr(th)[obs_t], y0_full[obs_t])
            rng = np.random.Generator(np.random.PCG64(77))
            elems = np.sort(rng.choice(len(g), min(args.fd_elems, len(g)), replace=False))
            t0 = time.perf_counter()
            with torch.no_grad():
                fd = fd_central(solve_obs, th0, elems, 1e-6, y_obs)
                v = rng.standard_normal(len(g))
                fdv = fd_direction(solve_obs, th0, torch.tensor(v, device=dev), 1e-6, y_obs)
            scale = np.abs(g).max()
            rel = np.abs(fd - g[elems]) / np.maximum(np.abs(fd), 1e-3 * scale)
            rel_pure = np.abs(fd - g[elems]) / np.abs(fd)
            res["fd"] = dict(h=1e-6, elements=len(elems), solve_rel_res=1e-13,
                             residual="f - K(theta) y0 = r0 - (K(theta) - K(theta0)) y0, the "
                                      "difference from an O(1) change of the same elements",
                             selection="32 elements drawn with PCG64 seed 77 from all",
                             error="|fd - g| / max(|fd|, 1e-3 max|g|)",
                             max_rel_err=float(rel.max()), median_rel_err=float(np.median(rel)),
                             max_pure_rel_err=float(rel_pure.max()),
                             median_pure_rel_err=float(np.median(rel_pure)),
                             min_abs_fd=float(np.abs(fd).min()), max_abs_g=float(scale),
                             direction_rel_err=float(abs(fdv - g @ v) / abs(fdv)),
                             seconds=time.perf_counter() - t0)
            log(f"[bench] adjoint {case} FD ({len(elems)} elems): max rel {rel.max():.2e} "
                f"median {np.median(rel):.2e}; random direction {res['fd']['direction_rel_err']:.2e}")
        if args.oracle_adjoint and case == "thermal":
            res["oracle_check"] = oracle_thermal_adjoint(theta, theta_true, obs, g, g64)
        out[case] = res
    return out


def oracle_thermal_adjoint(theta, theta_true, obs, g_mixed, g_fp64):
    """The whole thermal inversion gradient recomputed by the oracle (C
    restatement, serial CG as in the reference) on the host: observation
    solve, forward solve, adjoint solve, element backward."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    threads = min(os.cpu_count() or 1, 16)
    xyz, conn = O.generate_cube(64, 64, 32, 1000.0, 1000.0, 600.0)
    mask, vals = O.thermal_dirichlet(xyz, 600.0)
    f = O.assemble_thermal_load(xyz, conn, 1e-6)

    def solve(k, rhs=None):
        K = O.assemble_thermal(xyz, conn, 1.0, k_elem=k, nthreads=threads)
        Kb, fb = O.apply_dirichlet(K, f, mask, vals)
        x, st = O.cg_solve(Kb, fb if rhs is None else rhs, 100000, 1e-12, 0.0)
        return Kb, x, st

    k_true, k = np.exp(theta_true), np.exp(theta)
    _, t_true, _ = solve(k_true)
    Kb, T, st_f = solve(k)
    rhs = np.zeros(len(T))
    rhs[obs] = T[obs] - t_true[obs]
    rhs[mask == 1] = 0.0
    lam, st_a = O.cg_solve(Kb, rhs, 100000, 1e-12, 0.0)
    g_or = O.thermal_element_backward(xyz, conn, k, T, lam)
    scale = np.abs(g_or).max()
    out = dict(seconds=time.perf_counter() - t0, threads_assembly=threads,
               forward_iterations=st_f.iterations, adjoint_iterations=st_a.iterations,
               mixed_vs_oracle_max_rel=float(np.abs(g_mixed - g_or).max() / scale),
               fp64_vs_oracle_max_rel=float(np.abs(g_fp64 - g_or).max() / scale))
    log(f"[bench] oracle thermal adjoint ({out['seconds']:.0f} s): mixed vs oracle "
        f"{out['mixed_vs_oracle_max_rel']:.2e}, fp64 vs oracle {out['fp64_vs_oracle_max_rel']:.2e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="thermal_64x64x32", choices=sorted(CONFIGS))
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--inner-tol", type=float, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tight", action="store_true")
    ap.add_argument("--no-elastic", action="store_true")
    ap.add_argument("--no-bicgstab", action="store_true",
                    help="skip the BiCGSTAB and GMRES timings")
    ap.add_argument("--no-amg", action="store_true",
                    help="skip the AMG-preconditioned CG timing")
    ap.add_argument("--no-adjoint-thermal", action="store_true")
    ap.add_argument("--no-adjoint-elastic", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the elastic and adjoint sections")
    ap.add_argument("--fd-elems", type=int, default=32)
    ap.add_argument("--oracle-adjoint", action="store_true",
                    help="also recompute the thermal inversion gradient with the CPU "
                         "oracle (about a minute of host time)")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a plain device copy of this size (counter "
                         "calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.headline_only:
        args.no_elastic = args.no_adjoint_thermal = args.no_adjoint_elastic = True

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the assemble-and-solve path is built and "
              "measured on a single MI355X (multi-GPU sharding is a design note "
              "only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import geo_simulator_amd as G

    if not torch.cuda.is_available():
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    nx, ny, nz = CONFIGS[args.config]
    results, tight, matched, spmv, info = run_headline(G, args, nx, ny, nz)
    elastic = None if args.no_elastic else run_elastic(G, args)
    adjoint = None
    if not (args.no_adjoint_thermal and args.no_adjoint_elastic):
        adjoint = run_adjoint(G, args)

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        log(f"[bench] CPU baseline (oracle, {threads} threads for the element loop)")
        total, t_asm, t_cg, cst, per_iter, bicg_per_iter, gm_per_step, par = cpu_baseline(
            nx, ny, nz, threads)
        cpu = {"value": total, "unit": "s", "cores": threads, "kind": "port",
               "sample": (f"oracle C restatement on the same {nx}x{ny}x{nz} workload, "
                          f"one full step timed: assembly+load+Dirichlet {t_asm:.2f} s "
                          f"(element loop on {threads} threads, merge/CSR/BC serial as "
                          f"in the reference) and serial Jacobi-CG to rel-res 1e-8 "
                          f"{t_cg:.2f} s ({cst.iterations} iterations, "
                          f"{per_iter*1e3:.1f} ms/iter; GPU FP64: "
                          f"{results['fp64']['iterations']})"),
               "cg_iterations": cst.iterations, "assembly_s": t_asm, "cg_s": t_cg,
               "cg_parallel": {"s": par[0], "iterations": par[1], "threads": threads,
                               "note": "the same Jacobi-CG loop with OpenMP rows and "
                                       "reductions (timing column; not the reference's "
                                       "serial order)"},
               "bicgstab_s_per_pass": bicg_per_iter,
               "gmres30_s_per_step": gm_per_step,
               "host": host_cpu()}
        log(f"[bench] CPU baseline {total:.1f} s (asm {t_asm:.1f} s, {per_iter*1e3:.1f} ms/iter)")

    mixed, fp64 = results["mixed"], results["fp64"]
    out = {
        "metric": "assemble+PCG time to rel-res 1e-8 (s), SpMV GB/s; mixed vs FP64 on same GPU",
        "value": mixed["step_s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["step_s"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-outer",
        "data": "synthetic",
        "config": {"workload": f"thermal Tet10 assemble+Jacobi-PCG {nx}x{ny}x{nz} "
                               f"({info['elements']} elements, {info['n']} DOFs, "
                               f"nnz_bc {info['nnz']})",
                   "rel_res": args.tol,
                   "path": "mixed (FP64 flexible CG around an FP32 Jacobi-PCG recursion)"},
        "fp64": fp64,
        "mixed": mixed,
        "speedup_mixed_vs_fp64": fp64["step_s"] / mixed["step_s"],
        "tight_rel_res_1e-12": tight,
        "accuracy_matched_rel_l2_1e-10": matched,
        "krylov_fp64": {k: results[k] for k in ("bicgstab_fp64", "gmres30_fp64", "cg_amg_fp64")
                        if k in results},
        "symbolic_s": info["symbolic_s"],
        "upload_s": info["upload_s"],
        "assembly": info["assembly"],
        "spmv": spmv,
        "roofline": {"bound": "hbm", "achieved": spmv["fp32_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": spmv["fp32_GBs"] / HBM_PEAK_GBS,
                     "traffic": spmv["fp32_traffic"],
                     "traffic_measured_on": spmv["traffic_measured_on"],
                     "cache_resident": True,
                     "cache_note": "the 4-byte operator (194 MB) fits the 256 MB Infinity "
                                   "Cache, so this kernel runs above the box's HBM "
                                   "ceiling (spmv.hbm_ceiling); the FP64 SpMV, which "
                                   "streams from HBM, is given beside it",
                     "fp64_spmv": {"achieved": spmv["fp64_GBs"],
                                   "frac": spmv["fp64_frac"],
                                   "frac_of_read_ceiling": spmv["fp64_GBs"] /
                                   spmv["hbm_ceiling"]["read_GBs"],
                                   "traffic": spmv["fp64_traffic"],
                                   "kernel": "k_cg_q<double> (SELL-64 SpMV, FP64)"},
                     "kernel": "k_cg2_dir<float> (SELL-64 SpMV of the FP32 recursion on "
                               "the Jacobi-scaled operator, int32 fixed-point values, + "
                               "direction update p, s and p.s, r.p partials; dominant "
                               "kernel of the mixed path)",
                     "algorithmic_bytes": spmv["fp32_bytes"]},
        "elastic_48x48x24": elastic,
        "adjoint": adjoint,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
